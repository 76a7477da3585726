// G2 instantiation of the Pippenger MSM (see msm_impl.hpp).
#include "msm_impl.hpp"

namespace dvgpu {

#define DVGPU_MSM_INSTANTIATE(F)                                                                              \
    template void msm_batch_enqueue<F>(const uint32_t*, const uint32_t*, const uint32_t*, const MsmBatchPlan&, \
                                       MsmWorkspace&, hipStream_t, uint32_t*);                                \
    template void msm_batch_combine<F>(const uint32_t*, const MsmBatchPlan&, Affine<F>*);                     \
    template MsmPlan msm_enqueue<F>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t,    \
                                    const MsmConfig&, uint32_t*, size_t);                                     \
    template Affine<F> msm_run<F>(const uint32_t*, const uint32_t*, size_t, MsmWorkspace&, hipStream_t,      \
                                  const MsmConfig&);                                                          \
    template Affine<F> msm_combine_windows<F>(const uint32_t*, const MsmPlan&);                               \
    template void points_to_mont<F>(const uint8_t*, uint32_t*, size_t, hipStream_t);                          \
    template void affine_to_bytes<F>(const Affine<F>&, uint8_t*);                                            \
    template void msm_batch_finalize_enqueue<F>(const uint32_t*, const MsmBatchPlan&, const MsmWorkspace&,       \
                                                uint8_t*, uint32_t*, hipStream_t);

DVGPU_MSM_INSTANTIATE(Fq2Ops)

}  // namespace dvgpu
