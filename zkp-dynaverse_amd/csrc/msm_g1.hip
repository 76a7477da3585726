// G1 instantiation of the Pippenger MSM (see msm_impl.hpp).
#include "msm_impl.hpp"

namespace dvgpu {

int msm_default_window(size_t n) { return msm_default_window_impl(n); }
MsmPlan msm_plan(size_t n, const MsmConfig& cfg) { return msm_plan_impl(n, cfg); }

template MsmPlan msm_enqueue<FqOps>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t, const MsmConfig&, uint32_t*);
template Affine<FqOps> msm_run<FqOps>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t, const MsmConfig&);
template Affine<FqOps> msm_combine_windows<FqOps>(const uint32_t*, const MsmPlan&);
template void points_to_mont<FqOps>(const uint8_t*, uint32_t*, size_t, hipStream_t);
template void affine_to_bytes<FqOps>(const Affine<FqOps>&, uint8_t*);

}  // namespace dvgpu
