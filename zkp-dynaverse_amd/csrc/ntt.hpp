// Natural-order BN254 Fr NTT / INTT on gfx950.
#pragma once
#include <vector>

#include "common.hpp"
#include "field.hpp"

namespace dvgpu {

constexpr int NTT_MAX_LOG = 20;        // domain built from a root of order 2^20
constexpr int NTT_SINGLE_MAX_LOG = 11; // lengths up to 2^11 run as one LDS pass
constexpr int NTT_STAGE_TABLE_LOG = 11;

// Twiddle tables of one direction, built from the 2^20-th root the caller
// supplies (its inverse for the inverse direction):
//  * pow: omega^k for k < 2^19, Montgomery memory form (8 words each); the
//    group transforms of kzg.hip read it, and the tables below derive from it;
//  * stage: the sub-transform twiddles omega_(2^(s+1))^j, j < 2^s, for local
//    stages s < 11 (shared by every single pass and every second pass), as
//    nine 29-bit limb planes per stage;
//  * pass1[L] (2^12 <= 2^L <= 2^20, built on first use): the first pass of a
//    two-pass transform of length 2^L, per column x2 < 2^a (a = L - L/2),
//    stage s < L/2 and j < 2^s, omega_(2^(a+s+1))^(j 2^a + x2), as limb planes
//    grouped by tiles of columns so that a tile reads them in runs.
struct NttTables {
    DeviceBuffer pow, stage;
    DeviceBuffer pass1[NTT_MAX_LOG + 1];
    void set_quota(Quota* q) {
        pow.quota = stage.quota = q;
        for (auto& b : pass1) b.quota = q;
    }
    void release() {
        pow.release();
        stage.release();
        for (auto& b : pass1) b.release();
    }
};

struct NttDomain {
    uint32_t root_words[8] = {0};  // canonical root this domain was built from
    bool ready = false;
    NttTables fwd, inv;
    // Butterfly stages per LDS round trip by log2 length (the tuning table):
    // 1 = radix-2 rounds, 2 or more = radix-4 rounds, 0 = by launch size.
    // The pass split itself is fixed: one pass up to 2^11, two passes of
    // L/2 and L - L/2 stages above.
    int pass_stages[NTT_MAX_LOG + 1] = {0};
    void set_quota(Quota* q) {
        fwd.set_quota(q);
        inv.set_quota(q);
    }
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

// Device bytes of a domain's tables, both directions: the power and stage
// tables ntt_domain_prepare builds, plus (max_log >= 12) the first-pass
// tables of every length 2^12..2^max_log, built on first use.
uint64_t ntt_domain_bytes(int max_log);

struct NttWorkspace {
    DeviceBuffer a, b;
    void attach(Quota* q) { a.quota = b.quota = q; }
    void release() { a.release(); b.release(); }
};

// 1 / 2^log_n as canonical Montgomery limbs (a table made once per process).
bn254::Fr ntt_inverse_scale(int log_n);

// rows independent transforms of length n = 2^log_n, each row contiguous.
// The transform is linear and its twiddles are Montgomery values, so words
// in give words out in the same form: canonical in, canonical out
// (ntt_rows); Montgomery memory form in and out (ntt_rows_mont).  Outputs are
// fully reduced.  in and out may alias.  The inverse applies 1/n.
void ntt_rows(NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
              bool inverse, NttWorkspace& ws, hipStream_t stream);
void ntt_rows_mont(NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
                   bool inverse, NttWorkspace& ws, hipStream_t stream);

// ---- fused batches of transforms of several lengths --------------------------
// Every entry (rows x 2^log_n, canonical in and out) runs in the same launch
// sequence: one launch per pass for all entries (two at most).  Entries must
// not overlap each other (each may run in place).
constexpr int NTT_BATCH_MAX_PASSES = 2;

struct NttBatchEntry {
    const uint32_t* in;
    uint32_t* out;
    int log_n;
    size_t rows;
    bool inverse;
};

struct NttBatchPlan {
    uint32_t count = 0;
    int max_passes = 0;
    uint32_t pass_blocks[NTT_BATCH_MAX_PASSES] = {0};
    uint32_t pass_threads[NTT_BATCH_MAX_PASSES] = {0};
    bool pass_radix4[NTT_BATCH_MAX_PASSES] = {false, false};
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

// Plan, size the workspace, build any missing twiddle table and upload the
// descriptors (waits for this plan's previous descriptor upload first).
// Must precede ntt_batch_enqueue; capture-safe afterwards.
void ntt_batch_prepare(NttDomain& d, const std::vector<NttBatchEntry>& entries, NttBatchPlan& plan,
                       hipStream_t stream);
void ntt_batch_enqueue(const NttDomain& d, const NttBatchPlan& plan, hipStream_t stream);

}  // namespace dvgpu
