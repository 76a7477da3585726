// Natural-order BN254 Fr NTT / INTT on gfx950.
#pragma once
#include <vector>

#include "common.hpp"
#include "field.hpp"

namespace dvgpu {

constexpr int NTT_MAX_LOG = 20;  // domain built from a root of order 2^20

// Twiddles omega^k (k < 2^19) for one direction, in Montgomery memory form,
// built from the 2^20-th root the caller supplies.
struct NttDomain {
    uint32_t root_words[8] = {0};  // canonical root this domain was built from
    bool ready = false;
    DeviceBuffer fwd, inv;         // 2^19 x 8 words each
    // Butterfly stages per LDS pass by log2 length (0 = the maximum, 11).
    int pass_stages[NTT_MAX_LOG + 1] = {0};
    void release() {
        fwd.release();
        inv.release();
        ready = false;
    }
};

// Validate that `root` (canonical LE words) has multiplicative order exactly
// 2^20 in Fr.  Returns false otherwise.
bool ntt_root_has_order_2_20(const uint32_t (&root)[8]);

// Build (or reuse) the twiddle tables for `root` on `stream`.
void ntt_domain_prepare(NttDomain& d, const uint32_t (&root)[8], hipStream_t stream);

struct NttWorkspace {
    DeviceBuffer a, b;
    void attach(Quota* q) { a.quota = b.quota = q; }
    void release() { a.release(); b.release(); }
};

// rows independent transforms of length n = 2^log_n, each row contiguous.
// in: canonical LE words (8 per element); out: canonical LE words.
// in and out may alias.  `scale_inverse` applies 1/n on inverse.
void ntt_rows(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
              bool inverse, NttWorkspace& ws, hipStream_t stream);

// Same, but input and output in Montgomery memory form (resident data).
void ntt_rows_mont(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
                   bool inverse, NttWorkspace& ws, hipStream_t stream);

// ---- fused batches of transforms of several lengths --------------------------
// Every entry (rows x 2^log_n, canonical in and out) runs in the same launch
// sequence: one launch per pass for all entries, then one bit-reversal
// launch.  Entries must not overlap each other (each may run in place).
constexpr int NTT_BATCH_MAX_PASSES = 6;

struct NttBatchEntry {
    const uint32_t* in;
    uint32_t* out;
    int log_n;
    size_t rows;
    bool inverse;
};

struct NttBatchPlan {
    uint32_t count = 0, total_elems = 0;
    int max_passes = 0;
    uint32_t pass_blocks[NTT_BATCH_MAX_PASSES] = {0};
    uint32_t pass_threads[NTT_BATCH_MAX_PASSES] = {0};
    DeviceBuffer descs, work;
    HostBuffer host_descs;
    // Recorded after the descriptor upload: the next prepare waits for it
    // before rewriting host_descs (instead of for the whole stream).
    hipEvent_t uploaded = nullptr;
    bool upload_pending = false;
    void attach(Quota* q) { descs.quota = work.quota = q; }
    void release() {
        if (uploaded && upload_pending) (void)hipEventSynchronize(uploaded);
        if (uploaded) (void)hipEventDestroy(uploaded);
        uploaded = nullptr;
        upload_pending = false;
        descs.release();
        work.release();
        host_descs.release();
    }
};

// Plan, size the workspace and upload the descriptors (waits for this
// plan's previous descriptor upload first).  Must precede ntt_batch_enqueue; capture-safe afterwards.
void ntt_batch_prepare(const NttDomain& d, const std::vector<NttBatchEntry>& entries, NttBatchPlan& plan,
                       hipStream_t stream);
void ntt_batch_enqueue(const NttDomain& d, const NttBatchPlan& plan, hipStream_t stream);

}  // namespace dvgpu
