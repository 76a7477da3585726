// Resident operator engine: owner-tagged device buffers and the operators
// the incremental rebuilder runs on them (semantics of
// crates/protocol/src/incremental_rebuild/operators.rs:387-590 over the ABI
// of native/src/icicle_resident.cu / dynaverse_gpu.h:99-158).
//
// Storage.  Fr buffers hold Montgomery words (32 B), G1 buffers XYZZ
// Montgomery words (128 B), packed bases affine Montgomery words (64 B, the
// MSM layout, immutable).  All-zero words are Fr 0 and the point at
// infinity, so a new buffer is one fill.  Values cross the boundary in
// canonical form (Fr 32 B, affine G1 64 B) through pinned host staging and
// one conversion kernel.
//
// Operations.  Each call resolves and checks every buffer range first, then
// plans its journal saves (resident_journal.hpp: all or nothing), and only
// then issues stream work: the saves, the operation's kernels, and one
// synchronisation that books the device time.  Refusals (-1) therefore
// leave buffers, journal and owner untouched.
//
// Admission.  Every allocation is checked before it is made against the
// owner's quota and the device's free memory, together with the transient
// workspace an operation over the longest buffer can leave allocated (MSM
// plan buffers, staging, NTT tables), computed from the same plans the
// operations use.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstring>
#include <vector>

#include "g1mem.hpp"
#include "gfft.hpp"
#include "owner.hpp"
#include "resident_journal.hpp"
#include "setup_impl.hpp"

namespace dvgpu {

namespace {


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

__global__ void __launch_bounds__(64) k_transpose_g1(const uint32_t* __restrict__ in, uint32_t* __restrict__ out, size_t rows,
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

__global__ void __launch_bounds__(64) k_scale(const uint32_t* __restrict__ pts, const uint32_t* __restrict__ sc, uint32_t* __restrict__ out,
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

using namespace resident;

std::atomic<uint32_t> owner_tags{1};
constexpr size_t MAX_BUFFERS = 1024;
constexpr uint32_t OPCODES = DVGPU_RESIDENT_TRANSFORM_PACKET + 1;

uint64_t since_ns(std::chrono::steady_clock::time_point t0) {
    return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - t0).count());
}

int log2_ceil(uint64_t n) {
    int l = 0;
    while (l < 63 && (uint64_t(1) << l) < n) ++l;
    return l;
}

bool power_of_two(uint64_t n) { return n && !(n & (n - 1)); }

const char* kind_name(uint32_t kind) { return kind == KIND_FR ? "Fr" : (kind == KIND_G1 ? "G1" : "packed bases"); }

// One resident buffer.
struct Slot {
    void* ptr = nullptr;
    uint64_t length = 0;  // elements
    uint32_t kind = KIND_FR;
    bool writable_sealed = false;  // stays writable after seal
    uint32_t width() const { return stored_width(kind); }
    uint8_t* at(uint64_t i) const { return static_cast<uint8_t*>(ptr) + i * width(); }
};

// A checked element range of one buffer.
struct View {
    Slot* slot;
    uint32_t index;
    uint64_t off, n;
    uint32_t* words() const { return reinterpret_cast<uint32_t*>(slot->at(off)); }
    WriteRange write() const { return {index, off, off + n, slot->width()}; }
};

// Two ranges of one buffer may not overlap, except (when `same_ok`) exactly.
void check_alias(const View& x, const View& y, bool same_ok) {
    if (x.index != y.index || !x.n || !y.n) return;
    const bool overlap = x.off < y.off + y.n && y.off < x.off + x.n;
    require(!overlap || (same_ok && x.off == y.off && x.n == y.n), "resident operation reads and writes overlapping ranges");
}

// Stream work between two events plus the host time spent issuing it.
struct Span {
    hipEvent_t begin = nullptr, end = nullptr;
    bool recorded = false;
    uint64_t host_ns = 0;
};

}  // namespace

struct ResidentState {
    Owner* owner;
    uint32_t tag;
    uint64_t limit;
    std::vector<Slot> slots;
    Journal journal;
    uint8_t* arena = nullptr;  // journal arena, journal.capacity() bytes
    bool sealed = false;
    uint64_t longest = 0;         // longest buffer allocated (elements)
    uint64_t transient_seen = 0;  // transient bytes after the previous operation
    DvgpuResidentStats stats{};
    DvgpuResidentDiagnostics diag{};
    NttDomain domain;
    NttWorkspace ntt;
    DeviceBuffer affine, canon_scalars, g2_points, boundary;
    HostBuffer pinned;
    Span whole, saves_span, work_span;
    DvgpuResidentProfile* profile = nullptr;

    ResidentState(Owner* o, uint64_t quota_limit, uint64_t journal_bytes)
        : owner(o), tag(owner_tags.fetch_add(1)), limit(quota_limit), journal(journal_bytes) {
        require(tag != 0, "resident owner tags exhausted");
        domain.set_quota(&o->quota);
        ntt.attach(&o->quota);
        for (DeviceBuffer* b : {&affine, &canon_scalars, &g2_points, &boundary}) b->quota = &o->quota;
        slots.reserve(MAX_BUFFERS);
    }

    ~ResidentState() {
        if (owner->stream) (void)hipStreamSynchronize(owner->stream);
        for (Slot& s : slots)
            if (s.ptr) {
                (void)hipFree(s.ptr);
                owner->quota.release(s.length * s.width());
            }
        if (arena) {
            (void)hipFree(arena);
            owner->quota.release(journal.capacity());
        }
        for (Span* s : {&whole, &saves_span, &work_span})
            for (hipEvent_t e : {s->begin, s->end})
                if (e) (void)hipEventDestroy(e);
    }

    hipStream_t stream() const { return owner->stream; }

    // ---- admission -----------------------------------------------------------
    // Transient bytes an operation over `len` elements may leave allocated:
    // the owner's MSM workspace (a G2 MSM is the largest), canonical scalars,
    // affine points, G2 staging (wire and Montgomery), boundary staging, the
    // NTT work buffers (G1 rows, Fr rows above one pass) and the first-pass
    // NTT tables of every length up to len.
    uint64_t transient_need(uint64_t len) const {
        const uint64_t m = std::min<uint64_t>(len, DVGPU_MAX_MSM_SIZE);
        const uint64_t t = std::min<uint64_t>(len, DVGPU_MAX_TRANSFORM_MATRIX_SIZE);
        uint64_t need = msm_run_workspace_bytes(size_t(m), owner->msm_config, true);
        for (uint64_t bytes : {m * 32, m * 64, m * 256, mul_or_throw(len, 64), t * 128, t * 32})
            if (bytes) need += DeviceBuffer::grown(bytes);
        const int L = std::min(log2_ceil(len), NTT_MAX_LOG);
        need += ntt_domain_bytes(L) - ntt_domain_bytes(0);
        return need;
    }

    uint64_t transient_now() const {
        uint64_t b = affine.bytes + canon_scalars.bytes + g2_points.bytes + boundary.bytes + ntt.a.bytes + ntt.b.bytes;
        for (const DeviceBuffer* x : owner->msm.all) b += x->bytes;
        for (const NttTables* t : {&domain.fwd, &domain.inv})
            for (const DeviceBuffer& p : t->pass1) b += p.bytes;
        return b;
    }

    // Refuse (status -1) unless `bytes` more plus the transient workspace of
    // operations over `len` elements fit the quota and the free device memory.
    void admit(uint64_t bytes, uint64_t len) {
        const uint64_t reserve = transient_need(len), used = owner->quota.used;
        const bool fits = used <= limit && bytes <= limit - used && reserve <= limit - used - bytes;
        require(fits, "resident memory quota exceeded: " + std::to_string(used) + " bytes in use, " +
                          std::to_string(bytes) + " requested, " + std::to_string(reserve) +
                          " kept for operation workspace, limit " + std::to_string(limit));
        size_t free_bytes = 0, total_bytes = 0;
        HIP_CHECK(hipMemGetInfo(&free_bytes, &total_bytes));
        require(bytes <= free_bytes && reserve <= free_bytes - bytes,
                "device memory too small for the resident allocation and its operation workspace (quota " +
                    std::to_string(limit) + ", free " + std::to_string(free_bytes) + ")");
    }

    void* device_alloc(uint64_t bytes) {
        owner->quota.charge(bytes);
        void* p = nullptr;
        const hipError_t e = hipMalloc(&p, bytes);
        if (e != hipSuccess) {
            owner->quota.release(bytes);
            throw DeviceError(std::string("hipMalloc: ") + hipGetErrorString(e));
        }
        ++diag.resident_allocations;
        stats.bytes_allocated += bytes;
        return p;
    }

    // Events, the NTT tables and the journal arena, once the configuration is
    // admitted (the destructor releases whatever was made before a failure).
    void build(const uint32_t (&root)[8]) {
        for (Span* s : {&whole, &saves_span, &work_span}) {
            create_event(s->begin);
            create_event(s->end);
        }
        std::memcpy(domain.pass_stages, owner->domain.pass_stages, sizeof(domain.pass_stages));
        ntt_domain_prepare(domain, root, stream());
        diag.resident_allocations += 4;  // power and stage tables, both directions
        if (journal.capacity()) arena = static_cast<uint8_t*>(device_alloc(journal.capacity()));
        // Everything the owner holds counts, retained inputs of earlier calls too.
        stats.bytes_allocated = owner->quota.used;
        diag.journal_metadata_capacity_bytes = journal.record_bound() * sizeof(SavedRun);
    }

    // ---- buffers ---------------------------------------------------------------
    View view(uint64_t handle, uint64_t off, uint64_t n) {
        const uint32_t i = slot_of(handle, tag, slots.size());
        Slot& s = slots[i];
        require(off <= s.length && n <= s.length - off, "resident range [" + std::to_string(off) + ", +" +
                                                             std::to_string(n) + ") outside a buffer of " +
                                                             std::to_string(s.length) + " elements");
        return {&s, i, off, n};
    }
    View output(uint64_t handle, uint64_t off, uint64_t n) {
        const View v = view(handle, off, n);
        require(v.slot->kind != KIND_BASES, "packed MSM bases are immutable");
        require(!sealed || v.slot->writable_sealed, "resident buffer is immutable after seal");
        return v;
    }
    static void expect(const View& v, uint32_t kind) {
        if (v.slot->kind != kind)
            throw InvalidArgument(std::string("resident buffer kind mismatch: holds ") + kind_name(v.slot->kind) +
                                  ", operation needs " + kind_name(kind));
    }

    // ---- stream work of one operation ---------------------------------------
    void open_span(Span& s) {
        HIP_CHECK(hipEventRecord(s.begin, stream()));
        s.recorded = true;
    }
    void close_span(Span& s) { HIP_CHECK(hipEventRecord(s.end, stream())); }
    static uint64_t span_ns(const Span& s) {
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, s.begin, s.end));
        return uint64_t(double(ms) * 1e6);
    }

    void start() {
        saves_span.recorded = work_span.recorded = false;
        saves_span.host_ns = work_span.host_ns = 0;
        open_span(whole);
    }

    // Copy the old values of planned saves into the arena, then record them.
    void save(const std::vector<SavedRun>& saves) {
        if (saves.empty()) return;
        const auto t0 = std::chrono::steady_clock::now();
        if (profile) open_span(saves_span);
        for (const SavedRun& s : saves) {
            HIP_CHECK(hipMemcpyAsync(arena + s.arena, slots[s.slot].at(s.lo), s.bytes(), hipMemcpyDeviceToDevice,
                                     stream()));
            diag.journal_saved_bytes += s.bytes();
            diag.explicit_d2d_bytes += s.bytes();
        }
        diag.journal_saved_ranges += saves.size();
        journal.record(saves);
        diag.journal_peak_bytes = std::max(diag.journal_peak_bytes, journal.used());
        if (profile) close_span(saves_span);
        saves_span.host_ns += since_ns(t0);
    }

    // The operation's own kernels and copies.
    template <class F>
    void issue(F&& f) {
        const auto t0 = std::chrono::steady_clock::now();
        if (profile) open_span(work_span);
        f();
        HIP_CHECK_LAUNCH();
        if (profile) close_span(work_span);
        work_span.host_ns += since_ns(t0);
    }

    // Wait for the operation and book its device time.
    void finish() {
        close_span(whole);
        const auto t0 = std::chrono::steady_clock::now();
        const hipError_t waited = hipStreamSynchronize(stream());
        const uint64_t wait_ns = since_ns(t0);
        hip_check(waited, "resident operation");
        const uint64_t total = span_ns(whole);
        diag.device_operation_ns += total;
        ++diag.device_operations;
        const uint64_t transient = transient_now();
        if (transient > transient_seen) {
            ++diag.temporary_allocations;
            transient_seen = transient;
        }
        diag.temporary_peak_bytes = std::max(diag.temporary_peak_bytes, transient);
        if (!profile) return;
        profile->wait_host_ns += wait_ns;
        profile->journal_host_ns += saves_span.host_ns;
        profile->operation_host_ns += work_span.host_ns;
        profile->event_total_ns = total;
        profile->event_mask |= 4;
        if (saves_span.recorded) {
            profile->event_journal_ns = span_ns(saves_span);
            profile->event_mask |= 1;
        }
        if (work_span.recorded) {
            profile->event_operation_ns = span_ns(work_span);
            profile->event_mask |= 2;
        }
    }

    void host_copy(void* dst, const void* src, size_t bytes) {
        const auto t0 = std::chrono::steady_clock::now();
        std::memcpy(dst, src, bytes);
        if (profile) profile->conversion_host_ns += since_ns(t0);
    }

    // ---- transforms ------------------------------------------------------------
    void fr_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse) {
        ntt_rows_mont(domain, in, out, L, rows, inverse, ntt, stream());
    }

    // Natural order in and out: the G1 transform of gfft_rows (wave-shared
    // twiddles, GLV products) with one permutation pass, and for the inverse
    // the 1/n scaling as one shared-scalar product per point.
    void g1_rows(const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse) {
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
        const Fr ninv = ntt_inverse_scale(L);
        k_g1_mul_const<<<grid_for(total, 64), 64, 0, stream()>>>(work, out, total, ninv);
    }

    void rows_of(uint32_t kind, const uint32_t* in, uint32_t* out, int L, size_t rows, bool inverse) {
        if (kind == KIND_FR) fr_rows(in, out, L, rows, inverse);
        else g1_rows(in, out, L, rows, inverse);
    }

    void transpose_of(uint32_t kind, const uint32_t* in, uint32_t* out, size_t rows, size_t cols, size_t start,
                      size_t n, bool twiddle, const Fr& root) {
        if (kind == KIND_FR)
            k_transpose_fr<<<grid_for(n, 128), 128, 0, stream()>>>(in, out, rows, cols, start, n, twiddle, root);
        else
            k_transpose_g1<<<grid_for(n, 64), 64, 0, stream()>>>(in, out, rows, cols, start, n, twiddle, root);
        HIP_CHECK_LAUNCH();
    }

    // ---- operations ------------------------------------------------------------
    uint64_t op_alloc(const DvgpuResidentOp& op);
    uint64_t op_length(const DvgpuResidentOp& op) { return view(op.a, 0, 0).slot->length; }
    uint64_t op_kind(const DvgpuResidentOp& op) { return view(op.a, 0, 0).slot->kind; }
    uint64_t op_begin(const DvgpuResidentOp&) {
        journal.begin();
        return 0;
    }
    uint64_t op_rollback(const DvgpuResidentOp& op);
    uint64_t op_transfer(const DvgpuResidentOp& op);
    uint64_t op_copy(const DvgpuResidentOp& op);
    uint64_t op_transpose(const DvgpuResidentOp& op);
    uint64_t op_transform(const DvgpuResidentOp& op);
    uint64_t op_packet(const DvgpuResidentOp& op);
    uint64_t op_elementwise(const DvgpuResidentOp& op);
    uint64_t op_msm_g1(const DvgpuResidentOp& op);
    uint64_t op_msm_g2(const DvgpuResidentOp& op);

    uint64_t call(const DvgpuResidentOp& op);

    void seal() { sealed = true; }
    void commit() { journal.close(); }
};


uint64_t ResidentState::call(const DvgpuResidentOp& op) {
    using Handler = uint64_t (ResidentState::*)(const DvgpuResidentOp&);
    // One handler per opcode (dynaverse_gpu.h numbering).
    static const Handler handlers[OPCODES] = {
        nullptr,
        &ResidentState::op_alloc,        // ALLOC
        &ResidentState::op_length,       // LENGTH
        &ResidentState::op_kind,         // KIND
        &ResidentState::op_begin,        // BEGIN
        &ResidentState::op_rollback,     // ROLLBACK
        &ResidentState::op_transfer,     // WRITE
        &ResidentState::op_transfer,     // READ
        &ResidentState::op_copy,         // COPY
        &ResidentState::op_copy,         // REVERSE
        &ResidentState::op_transpose,    // TRANSPOSE
        &ResidentState::op_transform,    // TRANSFORM
        &ResidentState::op_elementwise,  // ZIP
        &ResidentState::op_elementwise,  // SCALE
        &ResidentState::op_msm_g1,       // MSM_G1
        &ResidentState::op_msm_g2,       // MSM_G2
        &ResidentState::op_alloc,        // ALLOC_BASES
        &ResidentState::op_packet,       // TRANSFORM_PACKET
    };
    ++stats.calls;
    require(op.opcode < OPCODES && handlers[op.opcode], "unknown resident opcode " + std::to_string(op.opcode));
    return (this->*handlers[op.opcode])(op);
}

void resident_destroy(ResidentState* s) noexcept { delete s; }
DvgpuResidentStats resident_stats(const ResidentState* s) { return s ? s->stats : DvgpuResidentStats{}; }
DvgpuResidentDiagnostics resident_diagnostics(const ResidentState* s) { return s ? s->diag : DvgpuResidentDiagnostics{}; }

// ALLOC (kind Fr or G1, flags = writable after seal) and ALLOC_BASES (input:
// canonical affine points, packed and immutable).  Returns the handle.
uint64_t ResidentState::op_alloc(const DvgpuResidentOp& op) {
    const bool bases = op.opcode == DVGPU_RESIDENT_ALLOC_BASES;
    const uint64_t n = op.length;
    require(!sealed, "resident buffers cannot be allocated after seal");
    require(!journal.open(), "resident buffers cannot be allocated inside a transaction");
    require(bases || op.kind <= KIND_G1, "resident buffer kind must be 0 (Fr) or 1 (G1)");
    require(slots.size() < MAX_BUFFERS, "resident buffer table is full");
    require(!bases || !n || op.input, "resident bases need their points");
    Slot s;
    s.kind = bases ? uint32_t(KIND_BASES) : op.kind;
    s.length = n;
    s.writable_sealed = !bases && op.flags != 0;
    const uint64_t bytes = mul_or_throw(n, s.width());
    admit(bytes, std::max(longest, n));
    if (bytes) s.ptr = device_alloc(bytes);
    slots.push_back(s);
    longest = std::max(longest, n);
    const Slot& b = slots.back();
    start();
    if (bases && n) {
        uint8_t* pin = pinned.as<uint8_t>(n * 64);
        host_copy(pin, op.input, n * 64);
        uint32_t* stage = boundary.as<uint32_t>(n * 16);
        issue([&] {
            HIP_CHECK(hipMemcpyAsync(stage, pin, n * 64, hipMemcpyHostToDevice, stream()));
            k_bases_in<<<grid_for(n, 128), 128, 0, stream()>>>(stage, static_cast<uint32_t*>(b.ptr), n);
        });
        stats.h2d_bytes += n * 64;
        diag.physical_h2d_bytes += n * 64;
    } else if (bytes) {
        issue([&] { device_fill_words(static_cast<uint32_t*>(b.ptr), bytes / 4, 0u, stream()); });
    }
    finish();
    return handle_of(tag, uint32_t(slots.size() - 1));
}

// Restore every saved run (newest first) and close the transaction.
uint64_t ResidentState::op_rollback(const DvgpuResidentOp&) {
    require(journal.open(), "no resident transaction is open");
    start();
    issue([&] {
        const auto& log = journal.log();
        for (auto r = log.rbegin(); r != log.rend(); ++r) {
            HIP_CHECK(hipMemcpyAsync(slots[r->slot].at(r->lo), arena + r->arena, r->bytes(), hipMemcpyDeviceToDevice,
                                     stream()));
            diag.explicit_d2d_bytes += r->bytes();
        }
    });
    finish();
    journal.close();
    return 0;
}

// WRITE / READ: canonical host values to or from a range (kind Fr, or G1 for
// G1 buffers and, read-only, packed bases; flags DVGPU_READ_DEVICE_AFFINE
// accepted for G1 reads, whose points are always normalised on the device).
uint64_t ResidentState::op_transfer(const DvgpuResidentOp& op) {
    const bool write = op.opcode == DVGPU_RESIDENT_WRITE;
    const uint64_t n = op.length;
    const View v = write ? output(op.a, op.a_offset, n) : view(op.a, op.a_offset, n);
    const uint32_t kind = v.slot->kind;
    require(kind == op.kind || (kind == KIND_BASES && op.kind == KIND_G1), "resident buffer kind mismatch");
    require(write || op.flags == 0 || (op.flags == DVGPU_READ_DEVICE_AFFINE && op.kind == KIND_G1),
            "invalid resident read flags");
    require(!n || (write ? op.input != nullptr : op.output != nullptr), "null resident host buffer");
    const std::vector<SavedRun> saves = write ? journal.plan({v.write()}) : std::vector<SavedRun>{};
    const uint32_t bw = boundary_width(kind);
    start();
    uint8_t* pin = nullptr;
    if (n) {
        save(saves);
        pin = pinned.as<uint8_t>(n * bw);
        uint32_t* stage = boundary.as<uint32_t>(n * bw / 4);
        if (write) {
            host_copy(pin, op.input, n * bw);
            issue([&] {
                HIP_CHECK(hipMemcpyAsync(stage, pin, n * bw, hipMemcpyHostToDevice, stream()));
                if (kind == KIND_FR) k_fr_in<<<grid_for(n, 128), 128, 0, stream()>>>(stage, v.words(), n);
                else k_g1_in<<<grid_for(n, 128), 128, 0, stream()>>>(stage, v.words(), n);
            });
        } else {
            issue([&] {
                if (kind == KIND_FR) k_fr_out<<<grid_for(n, 128), 128, 0, stream()>>>(v.words(), stage, n);
                else if (kind == KIND_G1) k_g1_out<true><<<grid_for(n, 128), 128, 0, stream()>>>(v.words(), stage, n);
                else k_bases_out<<<grid_for(n, 128), 128, 0, stream()>>>(v.words(), stage, n);
                HIP_CHECK_LAUNCH();
                HIP_CHECK(hipMemcpyAsync(pin, stage, n * bw, hipMemcpyDeviceToHost, stream()));
            });
        }
    }
    finish();
    if (write) {
        stats.h2d_bytes += n * bw;
        diag.physical_h2d_bytes += n * bw;
    } else {
        if (n) host_copy(op.output, pin, n * bw);
        stats.d2h_bytes += n * bw;
        diag.physical_d2h_bytes += n * bw;
    }
    return 0;
}

// COPY a -> c, or REVERSE (c[i] = a[n - 1 - i]); Fr or G1 buffers.
uint64_t ResidentState::op_copy(const DvgpuResidentOp& op) {
    const bool reverse = op.opcode == DVGPU_RESIDENT_REVERSE;
    const uint64_t n = op.length;
    const View src = view(op.a, op.a_offset, n);
    require(src.slot->kind != KIND_BASES, "packed MSM bases cannot be copied or reversed");
    if (reverse) require(src.slot->kind == op.kind, "resident reverse kind mismatch");
    const View dst = output(op.c, op.c_offset, n);
    expect(dst, src.slot->kind);
    check_alias(src, dst, !reverse || n <= 1);
    const bool nothing = !n || (!reverse && src.index == dst.index && src.off == dst.off);
    const std::vector<SavedRun> saves = nothing ? std::vector<SavedRun>{} : journal.plan({dst.write()});
    start();
    if (!nothing) {
        save(saves);
        const uint32_t w = src.slot->width();
        issue([&] {
            if (reverse) {
                k_reverse<<<grid_for(n, 128), 128, 0, stream()>>>(src.words(), dst.words(), n, int(w / 4));
            } else {
                HIP_CHECK(hipMemcpyAsync(dst.words(), src.words(), n * w, hipMemcpyDeviceToDevice, stream()));
                diag.explicit_d2d_bytes += n * w;
            }
        });
    }
    finish();
    return 0;
}

// TRANSPOSE elements [a_offset, a_offset + length) of the rows x cols
// matrix a into c (element r * cols + col -> col * rows + r), each value
// optionally times root^(r * col) (flags != 0, root canonical at input).
uint64_t ResidentState::op_transpose(const DvgpuResidentOp& op) {
    require(op.kind <= KIND_G1, "resident transpose kind must be 0 (Fr) or 1 (G1)");
    const uint64_t rows = op.rows, cols = op.cols, start_at = op.a_offset, n = op.length;
    const uint64_t total = mul_or_throw(rows, cols);
    require(rows && cols && total <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE && start_at <= total && n <= total - start_at &&
                op.a != op.c,
            "invalid resident transpose shape or alias");
    const View src = view(op.a, 0, total), dst = output(op.c, 0, total);
    expect(src, op.kind);
    expect(dst, op.kind);
    Fr root = fp_one<FrParams>();
    if (op.flags) {
        require(op.input != nullptr, "null resident transpose root");
        uint32_t w[8];
        std::memcpy(w, op.input, 32);
        require(is_canonical<FrParams>(w), "resident transpose root is not a canonical Fr element");
        root = canon(to_mont<FrParams>(w));
    }
    std::vector<Run> runs;
    transpose_destinations(rows, cols, start_at, n, runs);
    std::vector<WriteRange> writes;
    writes.reserve(runs.size());
    for (const Run& r : runs) writes.push_back({dst.index, r.lo, r.hi, dst.slot->width()});
    const std::vector<SavedRun> saves = journal.plan(writes);
    start();
    if (n) {
        save(saves);
        issue([&] { transpose_of(op.kind, src.words(), dst.words(), rows, cols, start_at, n, op.flags != 0, root); });
    }
    finish();
    return 0;
}

// TRANSFORM rows [a_offset, a_offset + rows) of length cols (a power of two)
// from a into the same rows of c; flags != 0 for the inverse.
uint64_t ResidentState::op_transform(const DvgpuResidentOp& op) {
    require(op.kind <= KIND_G1, "resident transform kind must be 0 (Fr) or 1 (G1)");
    const uint64_t len = op.cols, count = op.rows;
    const uint64_t first = mul_or_throw(op.a_offset, len), total = mul_or_throw(len, count);
    require(power_of_two(len) && len <= DVGPU_MAX_NTT_SIZE && count <= INT_MAX &&
                total <= DVGPU_MAX_TRANSFORM_MATRIX_SIZE && op.a != op.c,
            "invalid resident transform shape or alias");
    const View src = view(op.a, first, total), dst = output(op.c, first, total);
    expect(src, op.kind);
    expect(dst, op.kind);
    const std::vector<SavedRun> saves = journal.plan({dst.write()});
    start();
    if (total) {
        save(saves);
        issue([&] { rows_of(op.kind, src.words(), dst.words(), log2_ceil(len), size_t(count), op.flags != 0); });
        (op.kind == KIND_FR ? stats.fr_transform_elements : stats.g1_transform_elements) += total;
    }
    finish();
    return 0;
}

// TRANSFORM_PACKET: a length-n (n <= 4096) natural-order transform of a into
// c as a four-step over an a1 x a2 grid (a1 = 2^(L/2)), with b as scratch;
// b and c are journaled as a whole before either is touched.
uint64_t ResidentState::op_packet(const DvgpuResidentOp& op) {
    const uint64_t n = op.length;
    require(op.kind <= KIND_G1 && op.flags <= 1 && power_of_two(n) && n <= DVGPU_MAX_TRANSFORM_PACKET_SIZE &&
                op.a != op.b && op.a != op.c && op.b != op.c,
            "invalid resident packet shape, kind or alias");
    const View src = view(op.a, 0, n), scratch = output(op.b, 0, n), dst = output(op.c, 0, n);
    for (const View* v : {&src, &scratch, &dst}) expect(*v, op.kind);
    const std::vector<SavedRun> saves = journal.plan({scratch.write(), dst.write()});
    const int L = log2_ceil(n);
    const bool inverse = op.flags != 0;
    // omega_n (or its inverse) from the domain's 2^20-th root.
    Fr w = to_mont<FrParams>(domain.root_words);
    for (int k = L; k < NTT_MAX_LOG; ++k) w = sqr(w);
    if (inverse) w = inv(w);
    w = canon(w);
    const size_t a1 = size_t(1) << (L / 2), a2 = size_t(n) / a1;
    start();
    save(saves);
    issue([&] {
        uint32_t* tmp = scratch.words();
        uint32_t* out = dst.words();
        transpose_of(op.kind, src.words(), out, a1, a2, 0, n, false, w);
        rows_of(op.kind, out, tmp, log2_ceil(a1), a2, inverse);
        transpose_of(op.kind, tmp, out, a2, a1, 0, n, true, w);
        rows_of(op.kind, out, tmp, log2_ceil(a2), a1, inverse);
        transpose_of(op.kind, tmp, out, a1, a2, 0, n, false, w);
    });
    (op.kind == KIND_FR ? stats.fr_transform_elements : stats.g1_transform_elements) += 2 * n;
    finish();
    return 0;
}

// ZIP: c = a (+, -, *) b over Fr (flags 0, 1, 2); SCALE: c = [b] a over G1.
uint64_t ResidentState::op_elementwise(const DvgpuResidentOp& op) {
    const bool scale = op.opcode == DVGPU_RESIDENT_SCALE;
    const uint32_t kind = scale ? uint32_t(KIND_G1) : uint32_t(KIND_FR);
    const uint64_t n = op.length;
    const View left = view(op.a, op.a_offset, n), right = view(op.b, op.b_offset, n);
    const View dst = output(op.c, op.c_offset, n);
    expect(left, kind);
    expect(right, KIND_FR);
    expect(dst, kind);
    require(scale || op.flags <= 2, "invalid resident field operation (0 add, 1 subtract, 2 multiply)");
    check_alias(left, dst, true);
    check_alias(right, dst, true);
    const std::vector<SavedRun> saves = journal.plan({dst.write()});
    start();
    if (n) {
        save(saves);
        issue([&] {
            if (scale) k_scale<<<grid_for(n, 64), 64, 0, stream()>>>(left.words(), right.words(), dst.words(), n);
            else k_zip<<<grid_for(n, 128), 128, 0, stream()>>>(left.words(), right.words(), dst.words(), n, op.flags);
        });
    }
    finish();
    return 0;
}

// MSM_G1: sum b[i] a[i] over G1 or packed-bases points; 64-byte canonical
// affine result at output.
uint64_t ResidentState::op_msm_g1(const DvgpuResidentOp& op) {
    const uint64_t n = op.length;
    const View sc = view(op.b, op.b_offset, n);
    expect(sc, KIND_FR);
    require(n <= DVGPU_MAX_MSM_SIZE && op.output, "invalid resident MSM length or output");
    const View pts = view(op.a, op.a_offset, n);
    require(pts.slot->kind != KIND_FR, "resident MSM points must be G1 or packed bases");
    uint8_t result[64] = {0};
    start();
    if (n) {
        issue([&] {
            const uint32_t* bases = pts.words();
            if (pts.slot->kind == KIND_G1) {
                uint32_t* aff = affine.as<uint32_t>(n * 16);
                k_g1_out<false><<<grid_for(n, 128), 128, 0, stream()>>>(bases, aff, n);
                bases = aff;
            }
            uint32_t* s = canon_scalars.as<uint32_t>(n * 8);
            k_fr_out<<<grid_for(n, 128), 128, 0, stream()>>>(sc.words(), s, n);
            HIP_CHECK_LAUNCH();
            affine_to_bytes<FqOps>(msm_run<FqOps>(bases, s, n, owner->msm, stream(), owner->msm_config), result);
        });
        diag.physical_d2h_bytes += 64;
    }
    finish();
    std::memcpy(op.output, result, 64);
    stats.msm_g1_points += n;
    stats.d2h_bytes += 64;
    return 0;
}

// MSM_G2: sum b[i] Q_i over n canonical G2 points Q_i given at input (128 B
// each); 128-byte canonical affine result at output.
uint64_t ResidentState::op_msm_g2(const DvgpuResidentOp& op) {
    const uint64_t n = op.length;
    const View sc = view(op.b, op.b_offset, n);
    expect(sc, KIND_FR);
    require(n <= DVGPU_MAX_MSM_SIZE && op.output, "invalid resident MSM length or output");
    require(!n || op.input, "null resident G2 points");
    uint8_t result[128] = {0};
    start();
    if (n) {
        uint8_t* pin = pinned.as<uint8_t>(n * 128);
        host_copy(pin, op.input, n * 128);
        issue([&] {
            uint8_t* wire = g2_points.as<uint8_t>(n * 256);
            uint32_t* mont = reinterpret_cast<uint32_t*>(wire + n * 128);
            HIP_CHECK(hipMemcpyAsync(wire, pin, n * 128, hipMemcpyHostToDevice, stream()));
            points_to_mont<Fq2Ops>(wire, mont, n, stream());
            uint32_t* s = canon_scalars.as<uint32_t>(n * 8);
            k_fr_out<<<grid_for(n, 128), 128, 0, stream()>>>(sc.words(), s, n);
            HIP_CHECK_LAUNCH();
            affine_to_bytes<Fq2Ops>(msm_run<Fq2Ops>(mont, s, n, owner->msm, stream(), owner->msm_config), result);
        });
        diag.physical_h2d_bytes += n * 128;
        diag.physical_d2h_bytes += 128;
    }
    finish();
    std::memcpy(op.output, result, 128);
    stats.msm_g2_points += n;
    stats.h2d_bytes += n * 128;
    stats.d2h_bytes += 128;
    return 0;
}

}  // namespace dvgpu

using namespace dvgpu;

namespace {
Owner* as_owner(void* p) { return static_cast<Owner*>(p); }
}  // namespace

extern "C" {

// Reference: dynaverse_gpu.h:99-158 and native/src/icicle_resident.cu (the
// configure / call / seal / commit / stats ABI).  The quota covers every
// device byte the owner holds from here on, including inputs retained by
// earlier calls; the configuration is refused (-1, owner healthy, retry
// allowed) unless the NTT tables, the journal arena and the transient
// workspace of empty operations fit.
int dvgpu_resident_configure(void* backend, uint64_t limit, uint64_t journal, const uint8_t* root, char* error,
                             size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(root != nullptr, "resident configuration needs the 2^20-th root of unity");
        require(!o->resident, "resident engine already configured for this owner");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "resident NTT root must have order 2^20");
        auto s = std::make_unique<ResidentState>(o, limit, journal);
        const uint64_t fixed = ntt_domain_bytes(0);
        require(fixed <= UINT64_MAX - journal, "resident journal size overflows 64 bits");
        s->admit(fixed + journal, 0);
        const uint64_t previous = o->quota.limit;
        o->quota.limit = limit;
        try {
            s->build(rw);
        } catch (...) {
            s.reset();
            o->quota.limit = previous;
            throw;
        }
        o->resident = s.release();
        return 0;
    });
}

int dvgpu_resident_call(void* backend, const DvgpuResidentOp* op, uint64_t* result, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(op && result, "null resident operation or result");
        require(o->resident != nullptr, "resident engine not configured");
        *result = o->resident->call(*op);
        return 0;
    });
}

// Reference: crates/gpu/src/resident_profile.rs (host and event split of one
// call).  Profiles the operations that move or transform data.
int dvgpu_resident_call_profiled(void* backend, const DvgpuResidentOp* op, uint64_t* result,
                                 DvgpuResidentProfile* profile, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(op && result && profile, "null resident operation, result or profile");
        require(o->resident != nullptr, "resident engine not configured");
        const uint32_t c = op->opcode;
        require(c == DVGPU_RESIDENT_WRITE || c == DVGPU_RESIDENT_READ || c == DVGPU_RESIDENT_TRANSPOSE ||
                    c == DVGPU_RESIDENT_TRANSFORM || c == DVGPU_RESIDENT_TRANSFORM_PACKET,
                "operation cannot be profiled (write, read, transpose, transform or packet only)");
        ResidentState* s = o->resident;
        *profile = {};
        const uint64_t t0 = now_ns();
        const DvgpuResidentDiagnostics before = s->diag;
        s->profile = profile;
        try {
            *result = s->call(*op);
        } catch (...) {
            s->profile = nullptr;
            throw;
        }
        s->profile = nullptr;
        profile->h2d_bytes = s->diag.physical_h2d_bytes - before.physical_h2d_bytes;
        profile->d2h_bytes = s->diag.physical_d2h_bytes - before.physical_d2h_bytes;
        profile->d2d_bytes = s->diag.explicit_d2d_bytes - before.explicit_d2d_bytes;
        profile->native_wall_ns = now_ns() - t0;
        return 0;
    });
}

void dvgpu_resident_seal(void* backend) {
    Owner* o = as_owner(backend);
    if (o && o->resident) o->resident->seal();
}

void dvgpu_resident_commit(void* backend) {
    Owner* o = as_owner(backend);
    if (o && o->resident) o->resident->commit();
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
