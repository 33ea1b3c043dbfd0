// Extended-source contour kernels for 3 lens(es).
#include "extended_impl.hpp"

namespace caus {
CAUS_DEFINE_EXT(3)
}  // namespace caus
