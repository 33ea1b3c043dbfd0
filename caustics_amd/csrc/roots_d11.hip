// Root-solver instantiation for degree 11.
#include "roots_impl.hpp"

namespace caus {
CAUS_DEFINE_ROOTS_DEGREE(11)
}  // namespace caus
