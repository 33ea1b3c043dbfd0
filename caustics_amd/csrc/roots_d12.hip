// Root-solver instantiation for degree 12.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(12)
}  // namespace caus
