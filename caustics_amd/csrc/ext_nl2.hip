// Extended-source contour kernels for 2 lens(es).
#include "extended_impl.hpp"

namespace caus {
CAUS_DEFINE_EXT(2)
}  // namespace caus
