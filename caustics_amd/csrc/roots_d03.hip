// Root-solver instantiation for degree 3.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(3)
}  // namespace caus
