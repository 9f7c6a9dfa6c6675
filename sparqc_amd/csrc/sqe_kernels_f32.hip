// FP32 instantiations of the engine kernels (gfx950).
#include "sqe_kernels.hpp"

namespace sqe {
template struct Launch<float>;
}  // namespace sqe
