// Kernel instantiations: double, n = 2^9 .. 2^10.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f64_b(Registry<double>& r) {
  register_n<double, 512>(r);
  register_n<double, 1024>(r);
}

}  // namespace lp2d
