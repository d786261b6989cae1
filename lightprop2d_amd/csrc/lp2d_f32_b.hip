// Kernel instantiations: float, n = 2^9 .. 2^10.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f32_b(Registry<float>& r) {
  register_n<float, 512>(r);
  register_n<float, 1024>(r);
}

}  // namespace lp2d
