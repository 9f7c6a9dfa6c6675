// FP64 instantiations of the engine kernels (gfx950).
#include "sqe_kernels.hpp"

namespace sqe {
template struct Launch<double>;
}  // namespace sqe
