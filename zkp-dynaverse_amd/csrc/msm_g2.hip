// G2 instantiation of the Pippenger MSM (see msm_impl.hpp).
#include "msm_impl.hpp"

namespace dvgpu {

template MsmPlan msm_enqueue<Fq2Ops>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t, const MsmConfig&, uint32_t*);
template Affine<Fq2Ops> msm_run<Fq2Ops>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t, const MsmConfig&);
template Affine<Fq2Ops> msm_combine_windows<Fq2Ops>(const uint32_t*, const MsmPlan&);
template void points_to_mont<Fq2Ops>(const uint8_t*, uint32_t*, size_t, hipStream_t);
template void affine_to_bytes<Fq2Ops>(const Affine<Fq2Ops>&, uint8_t*);

}  // namespace dvgpu
