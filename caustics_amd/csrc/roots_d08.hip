// Root-solver instantiation for degree 8.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(8)
}  // namespace caus
