// Root-solver instantiation for degree 4.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(4)
}  // namespace caus
