// Kernel instantiations: double, n = 2^11 .. 2^13.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f64_c(Registry<double>& r) {
  register_n<double, 2048>(r);
  register_n<double, 4096>(r);
  register_n<double, 8192>(r);
}

}  // namespace lp2d
