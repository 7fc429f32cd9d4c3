// Kernel instantiation for dimension 2, 3 Gauss points per face direction.
#include "launch_impl.hpp"

namespace kfvm {
template struct Launch<2, 3>;
}  // namespace kfvm
