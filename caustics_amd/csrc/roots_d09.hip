// Root-solver instantiation for degree 9.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(9)
}  // namespace caus
