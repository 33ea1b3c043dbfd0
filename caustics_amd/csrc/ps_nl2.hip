// Point-source and light-curve kernels for 2 lens(es).
#include "ps_impl.hpp"

namespace caus {
CAUS_DEFINE_PS(2)
}  // namespace caus
