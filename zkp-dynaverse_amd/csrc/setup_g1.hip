// G1 instantiation of the fixed-base setup kernels (see setup_impl.hpp).
#include "setup_impl.hpp"

namespace dvgpu {
template void fixed_base_enqueue<FqOps>(SetupState&, const uint32_t*, size_t, uint8_t*, hipStream_t);
template void fixed_base_mont_enqueue<FqOps>(SetupState&, const uint32_t*, size_t, uint32_t*, hipStream_t);
}  // namespace dvgpu
