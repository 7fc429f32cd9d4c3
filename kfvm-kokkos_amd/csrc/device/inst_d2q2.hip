// Kernel instantiation for dimension 2, 2 Gauss points per face direction.
#include "launch_impl.hpp"

namespace kfvm {
template struct Launch<2, 2>;
}  // namespace kfvm
