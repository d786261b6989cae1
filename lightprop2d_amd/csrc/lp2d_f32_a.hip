// Kernel instantiations: float, n = 2^1 .. 2^8.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f32_a(Registry<float>& r) {
  register_n<float, 2>(r);
  register_n<float, 4>(r);
  register_n<float, 8>(r);
  register_n<float, 16>(r);
  register_n<float, 32>(r);
  register_n<float, 64>(r);
  register_n<float, 128>(r);
  register_n<float, 256>(r);
}

}  // namespace lp2d
