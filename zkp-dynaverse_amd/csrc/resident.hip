// Resident operator engine (reference: native/src/icicle_resident.cu,
// semantics of crates/protocol/src/incremental_rebuild/operators.rs:387-590).
//
// Owner-tagged device buffers of three kinds:
//   0 Fr      -- Montgomery words (32 B), converted at WRITE / READ;
//   1 G1      -- XYZZ Montgomery words (128 B); boundary form affine 64 B;
//   2 bases   -- immutable affine Montgomery words (64 B), MSM layout.
// All-zero words are Fr 0 and the G1 point at infinity, so ALLOC zero-fills
// with one memset.  Transactions save each element on its first write only
// (host bitmap per buffer, device journal arena), ROLLBACK restores in
// reverse order, COMMIT clears only the touched marks.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

#include "g1mem.hpp"
#include "gfft.hpp"
#include "owner.hpp"
#include "setup_impl.hpp"

namespace dvgpu {

namespace {

constexpr uint64_t DOMAIN_BYTES = (uint64_t(DVGPU_MAX_NTT_SIZE) + 1 + 6 + 64) * 32;
constexpr uint64_t WORKSPACE_MARGIN = 64ull << 20;
std::atomic<uint32_t> next_tag{1};

__global__ void k_fr_in(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[8];
    for (int k = 0; k < 8; ++k) w[k] = in[i * 8 + k];
    stfr(out + i * 8, canon(to_mont<FrParams>(w)));
}

__global__ void k_fr_out(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[8];
    from_mont(ldfr(in + i * 8), w);
    for (int k = 0; k < 8; ++k) out[i * 8 + k] = w[k];
}

// canonical affine 64 B (zero = infinity) -> XYZZ words
__global__ void k_g1_in(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x[8], y[8], any = 0;
    for (int k = 0; k < 8; ++k) {
        x[k] = in[i * 16 + k];
        y[k] = in[i * 16 + 8 + k];
        any |= x[k] | y[k];
    }
    G1 p = xyzz_inf<FqOps>();
    if (any) p = {canon(to_mont<FqParams>(x)), canon(to_mont<FqParams>(y)), fp_one<FqParams>(), fp_one<FqParams>()};
    stg1(out + i * 32, p);
}

// XYZZ -> affine; CANON: canonical 64 B, else Montgomery words (MSM bases).
template <bool CANON>
__global__ void k_g1_out(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Affine<FqOps> a = xyzz_to_affine<FqOps>(ldg1(in + i * 32));
    uint32_t* o = out + i * 16;
    if (a.inf) {
        for (int k = 0; k < 16; ++k) o[k] = 0;
        return;
    }
    if (CANON) {
        uint32_t w[8];
        from_mont(a.x, w);
        for (int k = 0; k < 8; ++k) o[k] = w[k];
        from_mont(a.y, w);
        for (int k = 0; k < 8; ++k) o[8 + k] = w[k];
    } else {
        stfq(o, a.x);
        stfq(o + 8, a.y);
    }
}

// canonical affine 64 B -> Montgomery affine words (packed MSM bases)
__global__ void k_bases_in(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t x[8], y[8], any = 0;
    for (int k = 0; k < 8; ++k) {
        x[k] = in[i * 16 + k];
        y[k] = in[i * 16 + 8 + k];
        any |= x[k] | y[k];
    }
    if (!any) {
        for (int k = 0; k < 16; ++k) out[i * 16 + k] = 0;
        return;
    }
    stfq(out + i * 16, canon(to_mont<FqParams>(x)));
    stfq(out + i * 16 + 8, canon(to_mont<FqParams>(y)));
}

// packed bases -> canonical affine 64 B
__global__ void k_bases_out(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t any = 0;
    for (int k = 0; k < 16; ++k) any |= in[i * 16 + k];
    if (!any) {
        for (int k = 0; k < 16; ++k) out[i * 16 + k] = 0;
        return;
    }
    uint32_t w[8];
    from_mont(ldfq(in + i * 16), w);
    for (int k = 0; k < 8; ++k) out[i * 16 + k] = w[k];
    from_mont(ldfq(in + i * 16 + 8), w);
    for (int k = 0; k < 8; ++k) out[i * 16 + 8 + k] = w[k];
}

__global__ void k_reverse(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n, int words) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* s = in + (n - 1 - i) * words;
    uint32_t* d = out + i * words;
    for (int k = 0; k < words; ++k) d[k] = s[k];
}

// out[col*rows + row] = in[i] * root^(row*col)  (twiddle optional)
__global__ void k_transpose_fr(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t rows,
                               size_t cols, size_t start, size_t n, int twiddle, Fr root) {
    const size_t off = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (off >= n) return;
    const size_t i = start + off, row = i / cols, col = i % cols;
    Fr v = ldfr(in + i * 8);
    if (twiddle) v = mul(v, fr_pow64(root, uint64_t(row) * col));
    stfr(out + (col * rows + row) * 8, v);
}

__global__ void k_transpose_g1(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t rows,
                               size_t cols, size_t start, size_t n, int twiddle, Fr root) {
    const size_t off = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (off >= n) return;
    const size_t i = start + off, row = i / cols, col = i % cols;
    G1 v = ldg1(in + i * 32);
    if (twiddle) v = g1_scale(v, fr_pow64(root, uint64_t(row) * col));
    stg1(out + (col * rows + row) * 32, v);
}

__global__ void k_zip(const uint32_t* __restrict__ a, const uint32_t* __restrict__ b, uint32_t* __restrict__ out,
                      size_t n, uint32_t op) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const Fr x = ldfr(a + i * 8), y = ldfr(b + i * 8);
    const Fr r = op == 0 ? add(x, y) : (op == 1 ? sub(x, y) : mul(x, y));
    stfr(out + i * 8, canon(r));
}

__global__ void k_scale(const uint32_t* __restrict__ pts, const uint32_t* __restrict__ sc, uint32_t* __restrict__ out,
                        size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    stg1(out + i * 32, g1_scale(ldg1(pts + i * 32), ldfr(sc + i * 8)));
}

// out[row][k] = x[row][rev(k)]
__global__ void k_ecntt_bitrev(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, int L, size_t total) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const size_t n = size_t(1) << L, row = i >> L, k = i & (n - 1);
    const uint32_t src = L ? (__brev(uint32_t(k)) >> (32 - L)) : 0;
    stg1(out + i * 32, ldg1(in + (row * n + src) * 32));
}

// out[i] = [c] x[i] for one scalar c shared by every lane.
__global__ void __launch_bounds__(64) k_g1_mul_const(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                     size_t total, Fr c) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= total) return;
    uint32_t k[8];
    fr_bits(c, k);
    stg1(out + i * 32, g1_mul_scalar(ldg1(in + i * 32), k, true));
}

// Montgomery Fr -> canonical words (MSM scalars)
__global__ void k_fr_canon(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t n) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t w[8];
    from_mont(ldfr(in + i * 8), w);
    for (int k = 0; k < 8; ++k) out[i * 8 + k] = w[k];
}

uint64_t checked_product(uint64_t a, uint64_t b) {
    require(!a || b <= UINT64_MAX / a, "resident size overflow");
    return a * b;
}

bool overlaps(uint64_t a, uint64_t b, uint64_t n) { return n && (a < b ? b - a < n : a - b < n); }

void alias(uint64_t src, uint64_t so, uint64_t dst, uint64_t doff, uint64_t n, bool exact_ok) {
    require(src != dst || !overlaps(so, doff, n) || (exact_ok && so == doff), "overlapping resident ranges");
}

struct RBuffer {
    void* ptr = nullptr;
    uint64_t length = 0;
    uint32_t kind = 0;
    bool writable = false;
    std::vector<uint8_t> marks;
    size_t width() const { return kind == 0 ? 32 : (kind == 1 ? 128 : 64); }
    size_t boundary_width() const { return kind == 0 ? 32 : 64; }
    bool marked(size_t i) const { return (marks[i / 8] >> (i % 8)) & 1; }
    void mark(size_t i, bool on) {
        const uint8_t bit = uint8_t(1u << (i % 8));
        if (on) marks[i / 8] |= bit;
        else marks[i / 8] &= uint8_t(~bit);
    }
    uint8_t* at(uint64_t off) const { return static_cast<uint8_t*>(ptr) + off * width(); }
};

struct JournalRange {
    uint64_t buffer, offset, length, arena;
};

}  // namespace

struct ResidentState {
    Owner* owner = nullptr;
    uint32_t tag = 0;
    uint64_t limit = 0, journal_capacity = 0, journal_used = 0, max_length = 0;
    bool sealed = false, transaction = false;
    uint8_t* journal = nullptr;
    std::vector<RBuffer> buffers;
    std::vector<JournalRange> undo;
    DvgpuResidentStats stats{};
    DvgpuResidentDiagnostics diag{};
    hipEvent_t started = nullptr, finished = nullptr, prof[4] = {};
    DvgpuResidentProfile* profile = nullptr;
    bool op_active = false;
    std::chrono::steady_clock::time_point op_started{};
    NttDomain domain;
    NttWorkspace ntt;
    DeviceBuffer affine_ws, scalar_ws, g2_ws, host_stage_dev;
    HostBuffer pinned;

    hipStream_t stream() const { return owner->stream; }

    RBuffer& buffer(uint64_t id) {
        require((id >> 32) == tag, "resident handle belongs to another owner");
        const uint64_t slot = id & 0xffffffffu;
        require(slot && slot <= buffers.size(), "unknown resident handle");
        return buffers[slot - 1];
    }
    RBuffer& range(uint64_t id, uint64_t off, uint64_t n, int kind = -1, bool write = false) {
        RBuffer& b = buffer(id);
        require(kind < 0 || b.kind == unsigned(kind), "resident buffer kind mismatch");
        require(off <= b.length && n <= b.length - off, "resident range out of bounds");
        require(!write || !sealed || b.writable, "resident buffer is immutable after seal");
        return b;
    }

    void save(uint64_t id, uint64_t off, uint64_t n) {
        if (!transaction || !n) return;
        const auto entered = std::chrono::steady_clock::now();
        if (profile && !(profile->event_mask & 1)) {
            HIP_CHECK(hipEventRecord(prof[0], stream()));
            profile->event_mask |= 1;
        }
        RBuffer& b = buffer(id);
        const uint64_t end = off + n;
        uint64_t entries = 0, ranges = 0;
        bool prior = false;
        for (uint64_t i = off; i < end; ++i) {
            const bool fresh = !b.marked(i);
            if (fresh) {
                ++entries;
                if (!prior) ++ranges;
            }
            prior = fresh;
        }
        const uint64_t bytes = checked_product(entries, b.width());
        require(bytes <= journal_capacity - journal_used, "resident transaction journal exhausted");
        require(ranges <= undo.capacity() - undo.size(), "resident transaction range limit exceeded");
        uint64_t i = off;
        while (i < end) {
            if (b.marked(i)) {
                ++i;
                continue;
            }
            const uint64_t first = i;
            while (i < end && !b.marked(i)) ++i;
            const uint64_t count = i - first, size = count * b.width();
            HIP_CHECK(hipMemcpyAsync(journal + journal_used, b.at(first), size, hipMemcpyDeviceToDevice, stream()));
            undo.push_back({id, first, count, journal_used});
            for (uint64_t j = first; j < i; ++j) b.mark(j, true);
            journal_used += size;
            diag.journal_saved_bytes += size;
            diag.explicit_d2d_bytes += size;
            ++diag.journal_saved_ranges;
        }
        diag.journal_peak_bytes = std::max(diag.journal_peak_bytes, journal_used);
        if (profile) {
            HIP_CHECK(hipEventRecord(prof[1], stream()));
            profile->journal_host_ns += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                     std::chrono::steady_clock::now() - entered)
                                                     .count());
        }
    }

    // Journal space for first writes to two whole buffers, checked before
    // either is saved (icicle_resident.cu:142-160).
    void preflight_pair(uint64_t first, uint64_t second, uint64_t n) {
        if (!transaction) return;
        uint64_t bytes = 0, ranges = 0;
        for (const uint64_t id : {first, second}) {
            RBuffer& b = buffer(id);
            uint64_t entries = 0;
            bool prior = false;
            for (uint64_t i = 0; i < n; ++i) {
                const bool fresh = !b.marked(i);
                if (fresh) {
                    ++entries;
                    if (!prior) ++ranges;
                }
                prior = fresh;
            }
            bytes += checked_product(entries, b.width());
        }
        require(bytes <= journal_capacity - journal_used, "resident packet journal exhausted");
        require(ranges <= undo.capacity() - undo.size(), "resident packet range limit exceeded");
    }

    void op_begin() {
        if (!profile) return;
        op_started = std::chrono::steady_clock::now();
        op_active = true;
        HIP_CHECK(hipEventRecord(prof[2], stream()));
        profile->event_mask |= 2;
    }
    void op_end() {
        if (!profile || !op_active) return;
        HIP_CHECK(hipEventRecord(prof[3], stream()));
        profile->operation_host_ns += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                   std::chrono::steady_clock::now() - op_started)
                                                   .count());
        op_active = false;
    }
    void wait() {
        const auto entered = std::chrono::steady_clock::now();
        const hipError_t r = hipStreamSynchronize(stream());
        if (profile)
            profile->wait_host_ns += uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(
                                                  std::chrono::steady_clock::now() - entered)
                                                  .count());
        hip_check(r, "resident stream synchronize");
    }
    // Peak-memory admission before any allocation (icicle_resident.cu:219-229).
    void quota(uint64_t extra, uint64_t next_max) {
        const uint64_t workspace = WORKSPACE_MARGIN + checked_product(next_max, 1024);
        const uint64_t allocated = owner->quota.used;
        require(allocated <= limit && extra <= limit - allocated && workspace <= limit - allocated - extra,
                "resident peak memory quota exceeded before allocation");
        size_t free_bytes = 0, total_bytes = 0;
        HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
        require(extra <= free_bytes && workspace <= free_bytes - extra,
                "insufficient device memory for resident buffers and bounded workspace");
    }
    void complete() {
        op_end();
        HIP_CHECK_LAUNCH();
        HIP_CHECK(hipEventRecord(finished, stream()));
        wait();
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, started, finished));
        diag.device_operation_ns += uint64_t(double(ms) * 1e6);
        ++diag.device_operations;
        if (profile) {
            profile->event_total_ns = uint64_t(double(ms) * 1e6);
            profile->event_mask |= 4;
            if (profile->event_mask & 1) {
                HIP_CHECK(hipEventElapsedTime(&ms, prof[0], prof[1]));
                profile->event_journal_ns = uint64_t(double(ms) * 1e6);
            }
            if (profile->event_mask & 2) {
                HIP_CHECK(hipEventElapsedTime(&ms, prof[2], prof[3]));
                profile->event_operation_ns = uint64_t(double(ms) * 1e6);
            }
        }
    }
    void publish(bool seal) {
        if (seal) {
            sealed = true;
            return;
        }
        for (const auto& r : undo) {
            RBuffer& b = buffer(r.buffer);
            for (uint64_t i = r.offset; i < r.offset + r.length; ++i) b.mark(i, false);
        }
        undo.clear();
        journal_used = 0;
        transaction = false;
    }
    ~ResidentState() {
        if (owner && owner->stream) (void)hipStreamSynchronize(owner->stream);
        for (auto& b : buffers)
            if (b.ptr) {
                (void)hipFree(b.ptr);
                owner->quota.release(b.length * b.width());
            }
        if (journal) {
            (void)hipFree(journal);
            owner->quota.release(journal_capacity);
        }
        for (hipEvent_t e : {started, finished, prof[0], prof[1], prof[2], prof[3]})
            if (e) (void)hipEventDestroy(e);
    }

    uint64_t execute(const DvgpuResidentOp& op);
    void ecntt_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse);
    void fr_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse);
};

void resident_destroy(ResidentState* s) noexcept { delete s; }
DvgpuResidentStats resident_stats(const ResidentState* s) { return s ? s->stats : DvgpuResidentStats{}; }
DvgpuResidentDiagnostics resident_diagnostics(const ResidentState* s) { return s ? s->diag : DvgpuResidentDiagnostics{}; }

void ResidentState::fr_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse) {
    ntt_rows_mont(domain, in, out, L, rows, inverse, ntt, stream());
}

// Natural order in and out: the G1 transform of gfft_rows (wave-shared
// twiddles, GLV products) with one permutation pass, and for the inverse the
// 1/n scaling as one shared-scalar product per point.
void ResidentState::ecntt_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse) {
    const size_t n = size_t(1) << L, total = n * rows;
    uint32_t* work = ntt.b.as<uint32_t>(total * 32);
    if (!inverse) {
        HIP_CHECK(hipMemcpyAsync(work, in, total * 128, hipMemcpyDeviceToDevice, stream()));
        gfft_rows(domain, work, L, rows, false, stream());
        k_ecntt_bitrev<<<grid_for(total, 64), 64, 0, stream()>>>(work, out, L, total);
        HIP_CHECK_LAUNCH();
        return;
    }
    k_ecntt_bitrev<<<grid_for(total, 64), 64, 0, stream()>>>(in, work, L, total);
    HIP_CHECK_LAUNCH();
    gfft_rows(domain, work, L, rows, true, stream());
    uint32_t nw[8] = {0};
    nw[L / 32] = 1u << (L % 32);
    const Fr ninv = canon(inv(to_mont<FrParams>(nw)));
    k_g1_mul_const<<<grid_for(total, 64), 64, 0, stream()>>>(work, out, total, ninv);
    HIP_CHECK_LAUNCH();
}

static int log2_exact(uint64_t n) {
    int l = 0;
    while ((uint64_t(1) << l) < n) ++l;
    return l;
}

uint64_t ResidentState::execute(const DvgpuResidentOp& op) {
    ++stats.calls;
    const size_t n = op.length;
    if (op.opcode == DVGPU_RESIDENT_LENGTH) return buffer(op.a).length;
    if (op.opcode == DVGPU_RESIDENT_KIND) return buffer(op.a).kind;
    if (op.opcode == DVGPU_RESIDENT_BEGIN) {
        require(!transaction, "resident transaction already open");
        transaction = true;
        journal_used = 0;
        undo.clear();
        return 0;
    }
    HIP_CHECK(hipEventRecord(started, stream()));
    hipStream_t st = stream();
    if (op.opcode == DVGPU_RESIDENT_ALLOC || op.opcode == DVGPU_RESIDENT_ALLOC_BASES) {
        const bool packed = op.opcode == DVGPU_RESIDENT_ALLOC_BASES;
        require(!sealed && !transaction, "resident allocation after seal or transaction start");
        require(op.kind <= 1 && buffers.size() < 1024, "resident kind or buffer limit invalid");
        require(!packed || !n || op.input, "null affine basis input");
        RBuffer b;
        b.length = op.length;
        b.kind = packed ? 2u : op.kind;
        b.writable = !packed && op.flags != 0;
        const uint64_t bytes = checked_product(n, b.width());
        quota(bytes, std::max<uint64_t>(max_length, op.length));
        b.marks.assign(n / 8 + (n % 8 != 0), 0);
        if (bytes) {
            owner->quota.charge(bytes);
            const hipError_t e = hipMalloc(&b.ptr, bytes);
            if (e != hipSuccess) {
                owner->quota.release(bytes);
                throw DeviceError(std::string("hipMalloc: ") + hipGetErrorString(e));
            }
            ++diag.resident_allocations;
        }
        buffers.push_back(std::move(b));
        RBuffer& nb = buffers.back();
        stats.bytes_allocated += bytes;
        max_length = std::max<uint64_t>(max_length, op.length);
        if (packed && n) {
            uint32_t* stage = host_stage_dev.as<uint32_t>(n * 16);
            HIP_CHECK(hipMemcpyAsync(stage, op.input, n * 64, hipMemcpyHostToDevice, st));
            k_bases_in<<<grid_for(n, 128), 128, 0, st>>>(stage, static_cast<uint32_t*>(nb.ptr), n);
            stats.h2d_bytes += n * 64;
            diag.physical_h2d_bytes += n * 64;
        } else if (bytes) {
            HIP_CHECK(hipMemsetAsync(nb.ptr, 0, bytes, st));
        }
        complete();
        return (uint64_t(tag) << 32) | buffers.size();
    }
    if (op.opcode == DVGPU_RESIDENT_ROLLBACK) {
        require(transaction, "no resident transaction open");
        for (auto it = undo.rbegin(); it != undo.rend(); ++it) {
            RBuffer& b = buffer(it->buffer);
            HIP_CHECK(hipMemcpyAsync(b.at(it->offset), journal + it->arena, it->length * b.width(),
                                     hipMemcpyDeviceToDevice, st));
            diag.explicit_d2d_bytes += it->length * b.width();
        }
        complete();
        publish(false);
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_WRITE || op.opcode == DVGPU_RESIDENT_READ) {
        const bool write = op.opcode == DVGPU_RESIDENT_WRITE;
        RBuffer& b = range(op.a, op.a_offset, n, -1, write);
        require(b.kind == op.kind || (b.kind == 2 && op.kind == 1), "resident buffer kind mismatch");
        require(!write || b.kind != 2, "packed MSM bases are immutable");
        require(write || (op.flags <= DVGPU_READ_DEVICE_AFFINE && (!op.flags || op.kind == 1)),
                "invalid resident read flags");
        require(!n || (write ? op.input != nullptr : op.output != nullptr), "null resident host buffer");
        if (!n) {
            complete();
            return 0;
        }
        if (write) save(op.a, op.a_offset, n);
        const size_t bw = b.boundary_width();
        uint32_t* stage = host_stage_dev.as<uint32_t>(n * bw / 4);
        uint8_t* pin = pinned.as<uint8_t>(n * bw);
        op_begin();
        if (write) {
            HIP_CHECK(hipMemcpyAsync(stage, op.input, n * bw, hipMemcpyHostToDevice, st));
            if (b.kind == 0)
                k_fr_in<<<grid_for(n, 128), 128, 0, st>>>(stage, reinterpret_cast<uint32_t*>(b.at(op.a_offset)), n);
            else
                k_g1_in<<<grid_for(n, 128), 128, 0, st>>>(stage, reinterpret_cast<uint32_t*>(b.at(op.a_offset)), n);
            HIP_CHECK_LAUNCH();
        } else {
            const uint32_t* src = reinterpret_cast<const uint32_t*>(b.at(op.a_offset));
            if (b.kind == 0) k_fr_out<<<grid_for(n, 128), 128, 0, st>>>(src, stage, n);
            else if (b.kind == 1) k_g1_out<true><<<grid_for(n, 128), 128, 0, st>>>(src, stage, n);
            else k_bases_out<<<grid_for(n, 128), 128, 0, st>>>(src, stage, n);
            HIP_CHECK_LAUNCH();
            HIP_CHECK(hipMemcpyAsync(pin, stage, n * bw, hipMemcpyDeviceToHost, st));
        }
        op_end();
        complete();
        if (!write) std::memcpy(op.output, pin, n * bw);
        if (write) {
            stats.h2d_bytes += n * bw;
            diag.physical_h2d_bytes += n * bw;
        } else {
            stats.d2h_bytes += n * bw;
            diag.physical_d2h_bytes += n * bw;
        }
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_COPY || op.opcode == DVGPU_RESIDENT_REVERSE) {
        RBuffer& src = range(op.a, op.a_offset, n);
        require(src.kind <= 1, "packed MSM bases cannot be copied or reversed");
        RBuffer& dst = range(op.c, op.c_offset, n, int(src.kind), true);
        if (op.opcode == DVGPU_RESIDENT_REVERSE) require(src.kind == op.kind, "reverse kind mismatch");
        alias(op.a, op.a_offset, op.c, op.c_offset, n, op.opcode == DVGPU_RESIDENT_COPY || n <= 1);
        if (!n || (op.opcode == DVGPU_RESIDENT_COPY && op.a == op.c && op.a_offset == op.c_offset)) {
            complete();
            return 0;
        }
        save(op.c, op.c_offset, n);
        op_begin();
        if (op.opcode == DVGPU_RESIDENT_COPY) {
            HIP_CHECK(hipMemcpyAsync(dst.at(op.c_offset), src.at(op.a_offset), n * src.width(),
                                     hipMemcpyDeviceToDevice, st));
            diag.explicit_d2d_bytes += n * src.width();
        } else {
            k_reverse<<<grid_for(n, 128), 128, 0, st>>>(reinterpret_cast<const uint32_t*>(src.at(op.a_offset)),
                                                       reinterpret_cast<uint32_t*>(dst.at(op.c_offset)), n,
                                                       int(src.width() / 4));
        }
        complete();
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_TRANSFORM_PACKET) {
        require(op.kind <= 1 && op.flags <= 1 && n && !(n & (n - 1)) && n <= DVGPU_MAX_TRANSFORM_PACKET_SIZE &&
                    op.a != op.b && op.a != op.c && op.b != op.c,
                "invalid resident packet shape, kind or alias");
        RBuffer& src = range(op.a, 0, n, int(op.kind));
        RBuffer& scr = range(op.b, 0, n, int(op.kind), true);
        RBuffer& dst = range(op.c, 0, n, int(op.kind), true);
        // w_n (or its inverse) from the configured 2^20-th root.
        const int L = log2_exact(n);
        Fr root = to_mont<FrParams>(domain.root_words);
        for (int k = L; k < 20; ++k) root = sqr(root);
        if (op.flags) root = inv(root);
        root = canon(root);
        preflight_pair(op.b, op.c, n);
        save(op.b, 0, n);
        save(op.c, 0, n);
        const size_t a = size_t(1) << (L / 2), bdim = n / a;
        const bool inv_dir = op.flags != 0;
        op_begin();
        auto* in = reinterpret_cast<const uint32_t*>(src.ptr);
        auto* tmp = reinterpret_cast<uint32_t*>(scr.ptr);
        auto* out = reinterpret_cast<uint32_t*>(dst.ptr);
        const unsigned blocks = grid_for(n, 128);
        if (op.kind == 0) {
            k_transpose_fr<<<blocks, 128, 0, st>>>(in, out, a, bdim, 0, n, 0, root);
            fr_rows(out, tmp, log2_exact(a), bdim, inv_dir);
            k_transpose_fr<<<blocks, 128, 0, st>>>(tmp, out, bdim, a, 0, n, 1, root);
            fr_rows(out, tmp, log2_exact(bdim), a, inv_dir);
            k_transpose_fr<<<blocks, 128, 0, st>>>(tmp, out, a, bdim, 0, n, 0, root);
            stats.fr_transform_elements += 2 * n;
        } else {
            k_transpose_g1<<<blocks, 128, 0, st>>>(in, out, a, bdim, 0, n, 0, root);
            ecntt_rows(out, tmp, log2_exact(a), bdim, inv_dir);
            k_transpose_g1<<<blocks, 128, 0, st>>>(tmp, out, bdim, a, 0, n, 1, root);
            ecntt_rows(out, tmp, log2_exact(bdim), a, inv_dir);
            k_transpose_g1<<<blocks, 128, 0, st>>>(tmp, out, a, bdim, 0, n, 0, root);
            stats.g1_transform_elements += 2 * n;
        }
        HIP_CHECK_LAUNCH();
        diag.explicit_d2d_bytes += ((a == 1) + (bdim == 1)) * n * src.width();
        complete();
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_TRANSPOSE) {
        require(op.kind <= 1, "packed MSM bases cannot be transposed");
        const size_t rows = op.rows, cols = op.cols, start = op.a_offset;
        const uint64_t total = checked_product(rows, cols);
        require(rows && cols && total <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE && start <= total && n <= total - start &&
                    op.a != op.c,
                "invalid resident transpose shape or alias");
        RBuffer& src = range(op.a, 0, total, int(op.kind));
        RBuffer& dst = range(op.c, 0, total, int(op.kind), true);
        if (!n) {
            complete();
            return 0;
        }
        Fr root = fp_one<FrParams>();
        if (op.flags) {
            require(op.input != nullptr, "null transpose root");
            uint32_t w[8];
            std::memcpy(w, op.input, 32);
            root = canon(to_mont<FrParams>(w));
        }
        // Transposed targets are contiguous within each column: journal only those.
        const size_t end = start + n, first = start / cols, last = (end - 1) / cols;
        const size_t count_cols = std::min(cols, n);
        for (size_t step = 0; step < count_cols; ++step) {
            const size_t col = (start % cols + step) % cols;
            const size_t lo = first + (col < start % cols);
            const size_t hi = last - (col > (end - 1) % cols);
            if (lo <= hi && hi < rows) save(op.c, col * rows + lo, hi - lo + 1);
        }
        op_begin();
        if (op.kind == 0)
            k_transpose_fr<<<grid_for(n, 128), 128, 0, st>>>(static_cast<const uint32_t*>(src.ptr),
                                                            static_cast<uint32_t*>(dst.ptr), rows, cols, start, n,
                                                            op.flags ? 1 : 0, root);
        else
            k_transpose_g1<<<grid_for(n, 64), 64, 0, st>>>(static_cast<const uint32_t*>(src.ptr),
                                                          static_cast<uint32_t*>(dst.ptr), rows, cols, start, n,
                                                          op.flags ? 1 : 0, root);
        complete();
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_TRANSFORM) {
        require(op.kind <= 1, "packed MSM bases cannot be transformed");
        const size_t row_len = op.cols, count = op.rows;
        const uint64_t offset = checked_product(op.a_offset, row_len), total = checked_product(row_len, count);
        require(row_len && !(row_len & (row_len - 1)) && row_len <= DVGPU_MAX_NTT_SIZE && count <= INT_MAX &&
                    total <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE && op.a != op.c,
                "invalid resident transform shape or alias");
        RBuffer& src = range(op.a, offset, total, int(op.kind));
        RBuffer& dst = range(op.c, offset, total, int(op.kind), true);
        if (!total) {
            complete();
            return 0;
        }
        save(op.c, offset, total);
        op_begin();
        const int L = log2_exact(row_len);
        if (op.kind == 0) {
            fr_rows(reinterpret_cast<const uint32_t*>(src.at(offset)), reinterpret_cast<uint32_t*>(dst.at(offset)), L,
                    count, op.flags != 0);
            stats.fr_transform_elements += total;
        } else {
            ecntt_rows(reinterpret_cast<const uint32_t*>(src.at(offset)), reinterpret_cast<uint32_t*>(dst.at(offset)), L,
                       count, op.flags != 0);
            stats.g1_transform_elements += total;
        }
        complete();
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_ZIP || op.opcode == DVGPU_RESIDENT_SCALE) {
        const bool scale = op.opcode == DVGPU_RESIDENT_SCALE;
        RBuffer& left = range(op.a, op.a_offset, n, scale ? 1 : 0);
        RBuffer& right = range(op.b, op.b_offset, n, 0);
        RBuffer& dst = range(op.c, op.c_offset, n, scale ? 1 : 0, true);
        require(scale || op.flags <= 2, "invalid resident field binary operation");
        alias(op.a, op.a_offset, op.c, op.c_offset, n, true);
        alias(op.b, op.b_offset, op.c, op.c_offset, n, true);
        if (!n) {
            complete();
            return 0;
        }
        save(op.c, op.c_offset, n);
        op_begin();
        if (scale)
            k_scale<<<grid_for(n, 64), 64, 0, st>>>(reinterpret_cast<const uint32_t*>(left.at(op.a_offset)),
                                                   reinterpret_cast<const uint32_t*>(right.at(op.b_offset)),
                                                   reinterpret_cast<uint32_t*>(dst.at(op.c_offset)), n);
        else
            k_zip<<<grid_for(n, 128), 128, 0, st>>>(reinterpret_cast<const uint32_t*>(left.at(op.a_offset)),
                                                   reinterpret_cast<const uint32_t*>(right.at(op.b_offset)),
                                                   reinterpret_cast<uint32_t*>(dst.at(op.c_offset)), n, op.flags);
        complete();
        return 0;
    }
    if (op.opcode == DVGPU_RESIDENT_MSM_G1 || op.opcode == DVGPU_RESIDENT_MSM_G2) {
        RBuffer& sc = range(op.b, op.b_offset, n, 0);
        require(n <= DVGPU_MAX_MSM_SIZE && op.output, "invalid resident MSM length or output");
        uint32_t* canon_sc = scalar_ws.as<uint32_t>(std::max<size_t>(n, 1) * 8);
        if (op.opcode == DVGPU_RESIDENT_MSM_G1) {
            RBuffer& pts = range(op.a, op.a_offset, n);
            require(pts.kind == 1 || pts.kind == 2, "resident MSM point kind mismatch");
            uint8_t result[64] = {0};
            if (n) {
                op_begin();
                const uint32_t* bases = reinterpret_cast<const uint32_t*>(pts.at(op.a_offset));
                if (pts.kind == 1) {
                    uint32_t* aff = affine_ws.as<uint32_t>(n * 16);
                    k_g1_out<false><<<grid_for(n, 128), 128, 0, st>>>(bases, aff, n);
                    bases = aff;
                }
                k_fr_canon<<<grid_for(n, 128), 128, 0, st>>>(reinterpret_cast<const uint32_t*>(sc.at(op.b_offset)),
                                                            canon_sc, n);
                HIP_CHECK_LAUNCH();
                const Affine<FqOps> r = msm_run<FqOps>(bases, canon_sc, n, owner->msm, st, owner->msm_config);
                affine_to_bytes<FqOps>(r, result);
                op_end();
                diag.physical_d2h_bytes += 64;
            }
            complete();
            std::memcpy(op.output, result, 64);
            stats.msm_g1_points += n;
            stats.d2h_bytes += 64;
        } else {
            require(!n || op.input, "null resident G2 bases");
            uint8_t result[128] = {0};
            if (n) {
                op_begin();
                uint8_t* stage = g2_ws.as<uint8_t>(n * 128 * 2);
                HIP_CHECK(hipMemcpyAsync(stage, op.input, n * 128, hipMemcpyHostToDevice, st));
                uint32_t* mont = reinterpret_cast<uint32_t*>(stage + n * 128);
                points_to_mont<Fq2Ops>(stage, mont, n, st);
                k_fr_canon<<<grid_for(n, 128), 128, 0, st>>>(reinterpret_cast<const uint32_t*>(sc.at(op.b_offset)),
                                                            canon_sc, n);
                HIP_CHECK_LAUNCH();
                const Affine<Fq2Ops> r = msm_run<Fq2Ops>(mont, canon_sc, n, owner->msm, st, owner->msm_config);
                affine_to_bytes<Fq2Ops>(r, result);
                op_end();
                diag.physical_h2d_bytes += n * 128;
                diag.physical_d2h_bytes += 128;
            }
            complete();
            std::memcpy(op.output, result, 128);
            stats.msm_g2_points += n;
            stats.h2d_bytes += n * 128;
            stats.d2h_bytes += 128;
        }
        return 0;
    }
    throw InvalidArgument("unknown resident opcode");
}

}  // namespace dvgpu

using namespace dvgpu;

namespace {
Owner* as_owner(void* p) { return static_cast<Owner*>(p); }
}  // namespace

extern "C" {

// icicle_msm.cu:267-293
int dvgpu_resident_configure(void* backend, uint64_t limit, uint64_t journal, const uint8_t* root, char* error,
                             size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(root && !o->resident, "resident configuration already set or root absent");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "resident NTT root must have order 2^20");
        require(limit >= DOMAIN_BYTES + WORKSPACE_MARGIN && journal <= limit - DOMAIN_BYTES - WORKSPACE_MARGIN,
                "resident quota cannot contain the fixed domain, journal and workspace reserve");
        auto s = std::make_unique<ResidentState>();
        s->owner = o;
        s->tag = next_tag.fetch_add(1);
        require(s->tag != 0, "resident owner tag exhausted");
        s->limit = limit;
        s->journal_capacity = journal;
        s->buffers.reserve(1024);
        s->undo.reserve(std::max<uint64_t>(1, journal / 32));
        s->diag.journal_metadata_capacity_bytes = s->undo.capacity() * sizeof(JournalRange);
        for (hipEvent_t* e : {&s->started, &s->finished, &s->prof[0], &s->prof[1], &s->prof[2], &s->prof[3]})
            HIP_CHECK(hipEventCreate(e));
        // From here on every owner allocation counts against the limit.
        const uint64_t previous_limit = o->quota.limit;
        o->quota.limit = limit;
        try {
            s->domain.set_quota(&o->quota);
            s->ntt.attach(&o->quota);
            for (DeviceBuffer* b : {&s->affine_ws, &s->scalar_ws, &s->g2_ws, &s->host_stage_dev}) b->quota = &o->quota;
            s->quota(journal, 0);
            std::memcpy(s->domain.pass_stages, o->domain.pass_stages, sizeof(s->domain.pass_stages));
            ntt_domain_prepare(s->domain, rw, o->stream);
            s->stats.bytes_allocated = DOMAIN_BYTES;
            s->diag.resident_allocations = 2;
            if (journal) {
                o->quota.charge(journal);
                const hipError_t e = hipMalloc(&s->journal, journal);
                if (e != hipSuccess) {
                    o->quota.release(journal);
                    s->journal = nullptr;
                    throw DeviceError(std::string("hipMalloc: ") + hipGetErrorString(e));
                }
                s->stats.bytes_allocated += journal;
                ++s->diag.resident_allocations;
            }
        } catch (const InvalidArgument&) {
            o->quota.limit = previous_limit;
            throw;
        }
        o->resident = s.release();
        return 0;
    });
}

int dvgpu_resident_call(void* backend, const DvgpuResidentOp* op, uint64_t* result, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(op && result && o->resident, "invalid resident call");
        *result = o->resident->execute(*op);
        return 0;
    });
}

int dvgpu_resident_call_profiled(void* backend, const DvgpuResidentOp* op, uint64_t* result,
                                 DvgpuResidentProfile* profile, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(op && result && profile && o->resident, "invalid profiled resident call");
        require(op->opcode == DVGPU_RESIDENT_WRITE || op->opcode == DVGPU_RESIDENT_READ ||
                    op->opcode == DVGPU_RESIDENT_TRANSPOSE || op->opcode == DVGPU_RESIDENT_TRANSFORM ||
                    op->opcode == DVGPU_RESIDENT_TRANSFORM_PACKET,
                "unsupported profiled resident operation");
        ResidentState* s = o->resident;
        *profile = {};
        struct Guard {
            ResidentState* s;
            ~Guard() {
                s->profile = nullptr;
                s->op_active = false;
            }
        } guard{s};
        const uint64_t started = now_ns();
        s->profile = profile;
        const DvgpuResidentDiagnostics before = s->diag;
        *result = s->execute(*op);
        profile->h2d_bytes = s->diag.physical_h2d_bytes - before.physical_h2d_bytes;
        profile->d2h_bytes = s->diag.physical_d2h_bytes - before.physical_d2h_bytes;
        profile->d2d_bytes = s->diag.explicit_d2d_bytes - before.explicit_d2d_bytes;
        profile->native_wall_ns = now_ns() - started;
        return 0;
    });
}

void dvgpu_resident_seal(void* backend) {
    Owner* o = as_owner(backend);
    if (o && o->resident) o->resident->publish(true);
}

void dvgpu_resident_commit(void* backend) {
    Owner* o = as_owner(backend);
    if (o && o->resident) o->resident->publish(false);
}

int dvgpu_resident_stats(void* backend, DvgpuResidentStats* out) {
    Owner* o = as_owner(backend);
    if (!o || !out) return -1;
    *out = o->resident ? o->resident->stats : o->retired_resident_stats;
    return 0;
}

int dvgpu_resident_diagnostics(void* backend, DvgpuResidentDiagnostics* out) {
    Owner* o = as_owner(backend);
    if (!o || !out) return -1;
    *out = o->resident ? o->resident->diag : o->retired_resident_diagnostics;
    return 0;
}

}  // extern "C"
