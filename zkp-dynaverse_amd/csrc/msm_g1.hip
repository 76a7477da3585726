// G1 instantiation of the Pippenger MSM (see msm_impl.hpp).
#include "msm_impl.hpp"

namespace dvgpu {

int msm_default_window(size_t n) { return msm_default_window_impl(n); }
int msm_window_for(size_t n, const WindowTable* table) { return msm_window_for_impl(n, table); }
MsmPlan msm_plan(size_t n, const MsmConfig& cfg) { return msm_plan_impl(n, cfg); }
MsmBatchPlan msm_batch_plan(const std::vector<MsmItem>& items, const MsmConfig& cfg, const WindowTable* table) {
    return msm_batch_plan_impl(items, cfg, table);
}
void msm_batch_prepare(const MsmBatchPlan& plan, int xyzz_words, MsmWorkspace& ws, hipStream_t stream) {
    msm_batch_prepare_impl(plan, xyzz_words, ws, stream);
}

uint64_t msm_run_workspace_bytes(size_t n, const MsmConfig& cfg, bool g2) {
    if (!n) return 0;
    const MsmBatchPlan plan = msm_single_plan_impl(MsmItem{n, 0, 0, cfg.window_bits}, cfg, g2);
    const int xw = g2 ? PointLayout<Fq2Ops>::XYZZ_WORDS : PointLayout<FqOps>::XYZZ_WORDS;
    const MsmBufferBytes b = msm_buffer_bytes(plan, xw);
    uint64_t total = 0;
    for (size_t part : {b.counts, b.offsets, b.digits, b.ranks, b.sorted, b.block_sums, b.total, b.tasks, b.partials,
                        b.heavy, b.chunks, b.hist, b.descs, b.windows, b.order})
        if (part) total += DeviceBuffer::grown(part);
    return total;
}

size_t msm_workspace_borrow(const MsmBatchPlan& plan, int xyzz_words, uint8_t* base, size_t capacity,
                            MsmWorkspace& ws) {
    return msm_workspace_borrow_impl(plan, xyzz_words, base, capacity, ws);
}

void device_fill_words(uint32_t* p, size_t words, uint32_t value, hipStream_t stream) {
    if (!words) return;
    k_fill<<<unsigned(std::min<size_t>(grid_for(words, 256), 4096)), 256, 0, stream>>>(p, words, value);
    HIP_CHECK_LAUNCH();
}

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

DVGPU_MSM_INSTANTIATE(FqOps)

}  // namespace dvgpu
