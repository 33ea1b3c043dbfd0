// Root-solver instantiation for degree 10.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(10)
}  // namespace caus
