// Root-solver instantiation for degree 2.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(2)
}  // namespace caus
