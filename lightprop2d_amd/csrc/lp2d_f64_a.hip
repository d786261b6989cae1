// Kernel instantiations: double, n = 2^1 .. 2^8.
#include "lp2d_inst.hpp"

namespace lp2d {

void register_f64_a(Registry<double>& r) {
  register_n<double, 2>(r);
  register_n<double, 4>(r);
  register_n<double, 8>(r);
  register_n<double, 16>(r);
  register_n<double, 32>(r);
  register_n<double, 64>(r);
  register_n<double, 128>(r);
  register_n<double, 256>(r);
}

}  // namespace lp2d
