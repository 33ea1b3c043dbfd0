// Root-solver instantiation for degree 1.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(1)
}  // namespace caus
