// Point-source and light-curve kernels for 3 lens(es).
#include "ps_impl.hpp"

namespace caus {
CAUS_DEFINE_PS(3)
}  // namespace caus
