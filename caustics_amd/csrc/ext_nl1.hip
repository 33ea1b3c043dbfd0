// Extended-source contour kernels for 1 lens(es).
#include "extended_impl.hpp"

namespace caus {
CAUS_DEFINE_EXT(1)
}  // namespace caus
