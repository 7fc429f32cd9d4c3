// Kernel instantiation for dimension 3, 2 Gauss points per face direction.
#include "launch_impl.hpp"

namespace kfvm {
template struct Launch<3, 2>;
}  // namespace kfvm
