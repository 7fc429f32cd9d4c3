// Kernel instantiation for dimension 3, 3 Gauss points per face direction.
#include "launch_impl.hpp"

namespace kfvm {
template struct Launch<3, 3>;
}  // namespace kfvm
