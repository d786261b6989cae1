// Kernel instantiations: float, n = 2^11 .. 2^13.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f32_c(Registry<float>& r) {
  register_n<float, 2048>(r);
  register_n<float, 4096>(r);
  register_n<float, 8192>(r);
}

}  // namespace lp2d
