// Root-solver instantiation for degree 7.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(7)
}  // namespace caus
