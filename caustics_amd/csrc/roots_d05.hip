// Root-solver instantiation for degree 5.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(5)
}  // namespace caus
