// Point-source and light-curve kernels for 1 lens(es).
#include "ps_impl.hpp"

namespace caus {
CAUS_DEFINE_PS(1)
}  // namespace caus
