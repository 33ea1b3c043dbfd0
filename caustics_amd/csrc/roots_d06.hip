// Root-solver instantiation for degree 6.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(6)
}  // namespace caus
