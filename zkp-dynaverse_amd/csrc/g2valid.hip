// G2 subgroup validation kernels and the setup / G2 entry points of the ABI.
//
// Wire rules (native/include/icicle_g2_wire.cuh:9-35): 128 bytes
// x.c0 || x.c1 || y.c0 || y.c1 little endian; all zero = infinity; byte 127
// carries two flag bits: 0xC0 is an encoding error, 0x40 marks infinity (the
// coordinates must still be canonical); the flag bits are masked off before
// the canonical (< q) check.  Masks: 1 encoding, 2 curve, 4 subgroup.
// Subgroup predicates (native/src/icicle_g2_validation.cu:20-71,
// crates/gpu/src/g2_validation.rs:250-299):
//   1: [r] P = O               (r as a 254-bit integer)
//   2: psi(P) = [6 x^2] P      (127-bit multiplier)
//   3: [x+1]P + psi([x]P) + psi^2([x]P) = 2 psi^3([x]P)   (63-bit seed x)
#include <algorithm>
#include <chrono>
#include <cstring>

#include "owner.hpp"
#include "setup_impl.hpp"

namespace dvgpu {

namespace {

using G2 = Xyzz<Fq2Ops>;

// Status of one 128-byte record given as 32 LE words (0 valid or infinity,
// 1 encoding error); p gets the point in Montgomery form.
__device__ uint32_t parse_words(uint32_t (&w)[32], Affine<Fq2Ops>& p) {
    uint32_t any = 0;
#pragma unroll
    for (int k = 0; k < 32; ++k) any |= w[k];
    p.inf = true;
    if (!any) return 0;
    const uint32_t flags = (w[31] >> 24) & 0xC0u;
    if (flags == 0xC0u) return 1;
    w[31] &= 0x3FFFFFFFu;
    uint32_t f[4][8];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int k = 0; k < 8; ++k) f[j][k] = w[8 * j + k];
        if (!is_canonical<FqParams>(f[j])) return 1;
    }
    if (flags == 0x40u) return 0;
    p.x = {to_mont<FqParams>(f[0]), to_mont<FqParams>(f[1])};
    p.y = {to_mont<FqParams>(f[2]), to_mont<FqParams>(f[3])};
    p.inf = false;
    return 0;
}

__device__ uint32_t parse_wire(const uint8_t* wire, size_t index, Affine<Fq2Ops>& p) {
    const uint4* q = reinterpret_cast<const uint4*>(wire + index * 128);
    uint32_t w[32];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const uint4 v = q[k];
        w[4 * k] = v.x;
        w[4 * k + 1] = v.y;
        w[4 * k + 2] = v.z;
        w[4 * k + 3] = v.w;
    }
    return parse_words(w, p);
}

// Non-adjacent form of the BN seed x, least significant digit first: 24
// nonzero digits where the binary form has 28, so the [x]P ladder below
// does 24 mixed additions (P or -P, both affine) after its doublings.
struct SeedNaf {
    int8_t d[66];
    int len;
};

constexpr SeedNaf seed_naf() {
    SeedNaf n{};
    unsigned __int128 kk = BN_SEED;
    while (kk) {
        int8_t v = 0;
        if (kk & 1) {
            v = (kk & 3) == 1 ? 1 : -1;
            kk = v > 0 ? kk - 1 : kk + 1;
        }
        n.d[n.len++] = v;
        kk >>= 1;
    }
    return n;
}

constexpr SeedNaf SEED_NAF = seed_naf();

// ---- lane teams --------------------------------------------------------------
//
// One subgroup check is a serial chain of about 90 G2 operations, so with
// one lane per point a call of 4096 points keeps 64 waves busy and the chain
// sets the time.  TeamFq2Ops runs a point's chain on a team of four
// consecutive lanes: every lane holds the same operands and follows the same
// control flow, and the three Fq products of a Karatsuba Fq2 product (two
// for a square) run on different lanes at once and are exchanged by lane
// shuffles.  Additions are computed redundantly on every lane of the team.
constexpr uint32_t TEAM = 4;

__device__ __forceinline__ Fq team_take(const Fq& v, int lane) {
    Fq r;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) r.l[k] = __shfl(v.l[k], lane, 64);
    return r;
}

struct TeamFq2Ops : Fq2Ops {
    static __device__ E mul(const E& a, const E& b) {
        const uint32_t role = threadIdx.x & (TEAM - 1);
        const int base = int(threadIdx.x & 63u & ~(TEAM - 1));
        // role 0: a0 b0, 1: a1 b1, 2 (and 3, unused): (a0 + a1)(b0 + b1)
        Fq x = add_w(a.c0, a.c1), y = norm(add_w(b.c0, b.c1));
        if (role == 0) {
            x = a.c0;
            y = b.c0;
        } else if (role == 1) {
            x = a.c1;
            y = b.c1;
        }
        const Fq p = bn254::mul(x, y);
        const Fq t0 = team_take(p, base), t1 = team_take(p, base + 1), t2 = team_take(p, base + 2);
        return {bn254::sub(t0, t1), reduce(sub2_n(t2, t0, t1))};
    }
    static __device__ E sqr(const E& a) {
        const uint32_t role = threadIdx.x & (TEAM - 1);
        const int base = int(threadIdx.x & 63u & ~(TEAM - 1));
        // role 0: (a0 + a1)(a0 - a1), others: a0 a1
        Fq x = a.c0, y = a.c1;
        if (role == 0) {
            x = add_w(a.c0, a.c1);
            y = norm(sub_w(a.c0, a.c1));
        }
        const Fq p = bn254::mul(x, y);
        const Fq s = team_take(p, base), t = team_take(p, base + 1);
        return {s, bn254::add(t, t)};
    }
};

template <class F>
__device__ Xyzz<F> psi_of(const Xyzz<F>& p) {
    return {F::mul(f2_conj(p.X), psi_cx()), F::mul(f2_conj(p.Y), psi_cy()), f2_conj(p.ZZ), f2_conj(p.ZZZ)};
}

// [x] P for an affine P by a left-to-right ladder over the NAF digits.
template <class F>
__device__ Xyzz<F> mul_seed(const Fq2& x, const Fq2& y) {
    const Fq2 ny = f2_neg(y);
    Xyzz<F> r = xyzz_inf<F>();
    for (int i = SEED_NAF.len - 1; i >= 0; --i) {
        r = xyzz_dbl<F>(r);
        const int8_t v = SEED_NAF.d[i];
        if (v) r = xyzz_madd<F>(r, x, v > 0 ? y : ny);
    }
    return r;
}

// compact: the schedule that keeps one running sum and one image live
// (the reference's second schedule) instead of [x]P and both images.
template <class F>
__device__ bool in_subgroup(const Affine<Fq2Ops>& a, uint32_t algorithm, bool compact) {
    const Xyzz<F> P = {a.x, a.y, f2_one(), f2_one()};
    if (algorithm == 1) {
        uint32_t r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) r[k] = FR_MOD32[k];
        return xyzz_is_inf<F>(xyzz_mul_words<F, 8>(P, r, 254));
    }
    if (algorithm == 2) {
        uint32_t lam[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) lam[k] = G2_LAMBDA[k];
        return xyzz_eq<F>(psi_of<F>(P), xyzz_mul_words<F, 4>(P, lam, 127));
    }
    if (compact) {
        Xyzz<F> t = mul_seed<F>(a.x, a.y);
        Xyzz<F> sum = xyzz_madd<F>(t, a.x, a.y);
        t = psi_of<F>(t);
        sum = xyzz_add<F>(sum, t);
        t = psi_of<F>(t);
        sum = xyzz_add<F>(sum, t);
        return xyzz_eq<F>(sum, xyzz_dbl<F>(psi_of<F>(t)));
    }
    const Xyzz<F> A = mul_seed<F>(a.x, a.y);
    const Xyzz<F> Bp = psi_of<F>(A);
    const Xyzz<F> C = psi_of<F>(Bp);
    const Xyzz<F> lhs = xyzz_add<F>(xyzz_add<F>(xyzz_madd<F>(A, a.x, a.y), Bp), C);
    return xyzz_eq<F>(lhs, xyzz_dbl<F>(psi_of<F>(C)));
}

// Point i from the wire (128-byte records) or from word planes (word w of
// point i at planes[w n + i], the repacked layout).
__device__ uint32_t parse_point(const uint8_t* input, size_t n, size_t i, bool soa, Affine<Fq2Ops>& p) {
    if (!soa) return parse_wire(input, i, p);
    const uint32_t* planes = reinterpret_cast<const uint32_t*>(input);
    uint32_t w[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) w[k] = planes[size_t(k) * n + i];
    return parse_words(w, p);
}

__global__ void k_parse_curve(const uint8_t* __restrict__ wire, size_t n, uint32_t* __restrict__ masks) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Affine<Fq2Ops> p;
    const uint32_t status = parse_wire(wire, i, p);
    masks[i] = status ? status : (affine_on_curve<Fq2Ops>(p) ? 0u : 2u);
}

// One team of TEAM lanes per point (i = thread / TEAM); fused: the encoding
// and curve checks run here too (no k_parse_curve pass).
// (At most 256 threads per block, the largest configuration: the bound
// lets a point's chain keep its operands in registers.)
template <bool FUSED>
__global__ void __launch_bounds__(256) k_subgroup(const uint8_t* __restrict__ input, size_t n, uint32_t algorithm, bool soa, bool compact,
                           uint32_t* __restrict__ masks) {
    const size_t i = (size_t(blockIdx.x) * blockDim.x + threadIdx.x) / TEAM;
    if (i >= n) return;
    const bool writer = (threadIdx.x & (TEAM - 1)) == 0;
    if (!FUSED && masks[i]) return;
    Affine<Fq2Ops> p;
    const uint32_t status = parse_point(input, n, i, soa, p);
    if (FUSED) {
        const uint32_t m = status ? status : (affine_on_curve<Fq2Ops>(p) ? 0u : 2u);
        if (writer) masks[i] = m;
        if (m) return;
    }
    if (p.inf) return;
    const bool ok = in_subgroup<TeamFq2Ops>(p, algorithm, compact);
    if (writer) masks[i] = ok ? 0u : 4u;
}

// Word planes of n 128-byte records: planes[w n + i] = word w of record i.
__global__ void k_repack(const uint32_t* __restrict__ wire, uint32_t* __restrict__ planes, size_t n) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t < n * 32) planes[(t % 32) * n + t / 32] = wire[t];
}

__global__ void k_reduce_status(const uint32_t* __restrict__ masks, size_t n, uint32_t algorithm,
                                DvgpuG2ValidationReport* __restrict__ report) {
    __shared__ uint64_t counts[256], first[256];
    __shared__ uint32_t bits[256];
    const unsigned t = threadIdx.x;
    uint64_t count = 0, idx = UINT64_MAX;
    uint32_t mask = 0;
    for (size_t i = t; i < n; i += 256) {
        const uint32_t m = masks[i];
        if (m) {
            ++count;
            mask |= m;
            if (i < idx) idx = i;
        }
    }
    counts[t] = count;
    first[t] = idx;
    bits[t] = mask;
    __syncthreads();
    for (unsigned s = 128; s; s >>= 1) {
        if (t < s) {
            counts[t] += counts[t + s];
            bits[t] |= bits[t + s];
            first[t] = min(first[t], first[t + s]);
        }
        __syncthreads();
    }
    if (t == 0) {
        DvgpuG2ValidationReport r;
        r.checked_points = n;
        r.invalid_points = counts[0];
        r.first_invalid = first[0];
        r.failure_mask = bits[0];
        r.algorithm = algorithm;
        *report = r;
    }
}

__global__ void k_corrupt(uint8_t* wire, size_t index) { wire[index * 128] ^= 1u; }

// psi images of homogeneous (X : Y : Z) points, canonical 192-byte records.
__global__ void __launch_bounds__(64) k_psi_probe(const uint8_t* __restrict__ in, size_t n, uint8_t* __restrict__ out,
                            uint32_t* __restrict__ masks) {
    const size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t* w = reinterpret_cast<const uint32_t*>(in + i * 192);
    uint32_t f[6][8];
    bool valid = true;
    for (int j = 0; j < 6; ++j) {
        for (int k = 0; k < 8; ++k) f[j][k] = w[8 * j + k];
        valid &= is_canonical<FqParams>(f[j]);
    }
    uint32_t* o = reinterpret_cast<uint32_t*>(out + i * 576);
    Fq2 X{}, Y{}, Z{};
    if (valid) {
        X = {to_mont<FqParams>(f[0]), to_mont<FqParams>(f[1])};
        Y = {to_mont<FqParams>(f[2]), to_mont<FqParams>(f[3])};
        Z = {to_mont<FqParams>(f[4]), to_mont<FqParams>(f[5])};
        // Y^2 Z = X^3 + b Z^3
        const Fq2 lhs = f2_mul(f2_sqr(Y), Z);
        const Fq2 rhs = f2_add(f2_mul(f2_sqr(X), X), f2_mul(Fq2Ops::b(), f2_mul(f2_sqr(Z), Z)));
        valid = f2_eq(lhs, rhs);
    }
    masks[i] = valid ? 0u : 1u;
    for (int j = 0; j < 3; ++j) {
        uint32_t* r = o + j * 48;
        if (!valid) {
            for (int k = 0; k < 48; ++k) r[k] = 0;
            continue;
        }
        X = f2_mul(f2_conj(X), psi_cx());
        Y = f2_mul(f2_conj(Y), psi_cy());
        Z = f2_conj(Z);
        put_canon(r, X);
        put_canon(r + 16, Y);
        put_canon(r + 32, Z);
    }
}

}  // namespace

void g2_validate_enqueue(const uint8_t* wire, size_t n, uint32_t algorithm, uint32_t* masks,
                         DvgpuG2ValidationReport* report, const G2CheckLaunch& l, hipStream_t stream) {
    const unsigned threads = l.config.threads;
    const bool fused = l.config.fused != 0, soa = l.config.layout != 0 && l.scratch;
    if (!fused) {
        k_parse_curve<<<grid_for(n, threads), threads, 0, stream>>>(wire, n, masks);
        HIP_CHECK_LAUNCH();
    }
    if (l.parsed) HIP_CHECK(hipEventRecord(l.parsed, stream));
    const uint8_t* input = wire;
    if (soa) {
        k_repack<<<grid_for(n * 32, 256), 256, 0, stream>>>(reinterpret_cast<const uint32_t*>(wire),
                                                          reinterpret_cast<uint32_t*>(l.scratch), n);
        HIP_CHECK_LAUNCH();
        input = l.scratch;
    }
    if (l.repacked) HIP_CHECK(hipEventRecord(l.repacked, stream));
    const bool compact = l.config.schedule != 0;
    if (fused)
        k_subgroup<true><<<grid_for(n * TEAM, threads), threads, 0, stream>>>(input, n, algorithm, soa, compact, masks);
    else
        k_subgroup<false><<<grid_for(n * TEAM, threads), threads, 0, stream>>>(input, n, algorithm, soa, compact, masks);
    HIP_CHECK_LAUNCH();
    if (l.subgroup_done) HIP_CHECK(hipEventRecord(l.subgroup_done, stream));
    k_reduce_status<<<1, 256, 0, stream>>>(masks, n, algorithm, report);
    HIP_CHECK_LAUNCH();
}

void g2_corrupt_point(uint8_t* wire, size_t index, hipStream_t stream) {
    k_corrupt<<<1, 1, 0, stream>>>(wire, index);
    HIP_CHECK_LAUNCH();
}

void g2_psi_probe_enqueue(const uint8_t* points, size_t n, uint8_t* out, uint32_t* masks, hipStream_t stream) {
    k_psi_probe<<<grid_for(n, 64), 64, 0, stream>>>(points, n, out, masks);
    HIP_CHECK_LAUNCH();
}

void g2_kernel_attributes(uint32_t, unsigned threads, DvgpuG2KernelResources& out) {
    hipFuncAttributes a{};
    HIP_CHECK(hipFuncGetAttributes(&a, reinterpret_cast<const void*>(k_subgroup<false>)));
    int blocks = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, reinterpret_cast<const void*>(k_subgroup<false>),
                                                           int(threads), 0));
    int device = 0, maximum = 0;
    HIP_CHECK(hipGetDevice(&device));
    HIP_CHECK(hipDeviceGetAttribute(&maximum, hipDeviceAttributeMaxThreadsPerMultiProcessor, device));
    out.registers = uint32_t(a.numRegs);
    out.block_threads = threads;
    out.active_blocks = uint32_t(blocks);
    out.max_threads_per_sm = uint32_t(maximum);
    out.local_bytes = a.localSizeBytes;
    out.shared_bytes = a.sharedSizeBytes;
    hipFuncAttributes b{};
    HIP_CHECK(hipFuncGetAttributes(&b, reinterpret_cast<const void*>(k_fixed_base<Fq2Ops, true>)));
    int nblocks = 0;
    HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &nblocks, reinterpret_cast<const void*>(k_fixed_base<Fq2Ops, true>), 256, 0));
    out.normalize_registers = uint32_t(b.numRegs);
    out.normalize_active_blocks = uint32_t(nblocks);
    out.normalize_local_bytes = b.localSizeBytes;
    out.normalize_shared_bytes = b.sharedSizeBytes;
}

void setup_destroy(SetupState* s) noexcept { delete s; }

}  // namespace dvgpu

using namespace dvgpu;

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

SetupState& setup_of(Owner* o) {
    if (!o->setup) {
        o->setup = new SetupState();
        o->setup->attach(&o->quota);
    }
    return *o->setup;
}

using EventSet = ScopedEvents<9>;  // 8: after the repack of a layout-1 validation

// [s_i] G into host `output` (canonical affine); events 0..5 are the
// table/H2D/generate/normalize/D2H boundaries of DvgpuSetupStats.device_ns.
template <class F>
void run_fixed_base(Owner* o, const uint8_t* scalars, size_t n, uint8_t* output, EventSet& ev,
                    int64_t corrupt, uint32_t algorithm, DvgpuG2ValidationReport* report_out,
                    uint32_t* dev_masks_out) {
    constexpr int PB = PointLayout<F>::CANON_BYTES;
    SetupState& st = setup_of(o);
    HIP_CHECK(hipEventRecord(ev.e[0], o->stream));
    if (st.mode) {
        // Building the table is part of the first call's "table" span.
        (void)setup_table<F>(st, o->stream);
    }
    HIP_CHECK(hipEventRecord(ev.e[1], o->stream));
    uint32_t* dsc = st.scalars.as<uint32_t>(n * 8);
    uint8_t* dpts = st.points.as<uint8_t>(n * PB);
    HIP_CHECK(hipMemcpyAsync(dsc, scalars, n * 32, hipMemcpyHostToDevice, o->stream));
    HIP_CHECK(hipEventRecord(ev.e[2], o->stream));
    fixed_base_enqueue<F>(st, dsc, n, dpts, o->stream);
    HIP_CHECK(hipEventRecord(ev.e[3], o->stream));
    // Generation and normalization run fused in one kernel; the normalize span
    // is recorded as empty.
    HIP_CHECK(hipEventRecord(ev.e[4], o->stream));
    if (report_out) {
        if (corrupt >= 0) g2_corrupt_point(dpts, size_t(corrupt), o->stream);
        uint32_t* masks = st.masks.as<uint32_t>(n);
        auto* rep = st.report.as<DvgpuG2ValidationReport>(1);
        G2CheckLaunch launch;
        launch.config = o->g2_kernel;
        if (launch.config.layout) launch.scratch = st.planes.as<uint8_t>(n * 128);
        launch.parsed = ev.e[6];
        launch.repacked = ev.e[8];
        launch.subgroup_done = ev.e[7];
        g2_validate_enqueue(dpts, n, algorithm, masks, rep, launch, o->stream);
        HIP_CHECK(hipMemcpyAsync(report_out, rep, sizeof(*rep), hipMemcpyDeviceToHost, o->stream));
        (void)dev_masks_out;
    }
    HIP_CHECK(hipMemcpyAsync(output, dpts, n * PB, hipMemcpyDeviceToHost, o->stream));
    HIP_CHECK(hipEventRecord(ev.e[5], o->stream));
    HIP_CHECK(hipStreamSynchronize(o->stream));
    const int g2 = PB == 128 ? 1 : 0;
    for (int k = 0; k < 5; ++k) st.stats.device_ns[k] += ev.ns(k, k + 1);
    ++st.stats.calls[g2];
    st.stats.points[g2] += n;
}

}  // namespace

extern "C" {

// icicle_msm.cu:523-532
int dvgpu_setup_mode(void* backend, int mode) {
    Owner* o = as_owner(backend);
    if (!o || o->retired || mode < 0 || mode > 2) return -1;
    setup_of(o).mode = mode;
    return 0;
}

int dvgpu_setup_stats(void* backend, DvgpuSetupStats* stats) {
    Owner* o = as_owner(backend);
    if (!o || !stats) return -1;
    *stats = o->setup ? o->setup->stats : DvgpuSetupStats{};
    return 0;
}

// icicle_msm.cu:494-521
int dvgpu_bn254_fixed_base(void* backend, const uint8_t* scalars, size_t length, int g2, uint8_t* output,
                           char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(length <= 4096 && (g2 == 0 || g2 == 1) && (!length || (scalars && output)),
                "invalid fixed-base buffers or chunk length");
        if (!length) return 0;
        const uint64_t t0 = now_ns();
        EventSet ev;
        uint8_t* staged = o->pinned.as<uint8_t>(length * (g2 ? 128 : 64));
        if (g2)
            run_fixed_base<Fq2Ops>(o, scalars, length, staged, ev, -1, 0, nullptr, nullptr);
        else
            run_fixed_base<FqOps>(o, scalars, length, staged, ev, -1, 0, nullptr, nullptr);
        std::memcpy(output, staged, length * (g2 ? 128 : 64));
        o->setup->stats.native_wall_ns += now_ns() - t0;
        if (g2) {
            o->g2_traffic.scalar_h2d_bytes += length * 32;
            o->g2_traffic.point_d2h_bytes += length * 128;
        }
        return 0;
    });
}

// icicle_msm.cu:583-618: generate, validate on the device, publish only if
// every point is accepted; -5 publishes the report alone.
int dvgpu_bn254_fixed_base_g2_checked(void* backend, const uint8_t* scalars, size_t n, uint32_t algorithm,
                                      uint8_t* output, DvgpuG2ValidationReport* report, char* error,
                                      size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(n <= 4096 && algorithm >= 1 && algorithm <= 3 && report && (!n || (scalars && output)),
                "invalid checked G2 buffers or algorithm");
        if (!n) {
            *report = {0, 0, UINT64_MAX, 0, algorithm};
            return 0;
        }
        const int64_t corrupt = o->g2_corrupt_next;
        o->g2_corrupt_next = -1;
        require(corrupt < int64_t(n), "test corruption index outside chunk");
        const uint64_t t0 = now_ns();
        EventSet ev;
        uint8_t* staged = o->pinned.as<uint8_t>(n * 128 + sizeof(DvgpuG2ValidationReport));
        auto* rep = reinterpret_cast<DvgpuG2ValidationReport*>(staged + n * 128);
        run_fixed_base<Fq2Ops>(o, scalars, n, staged, ev, corrupt, algorithm, rep, nullptr);
        const uint64_t wall = now_ns() - t0;
        o->setup->stats.native_wall_ns += wall;
        auto& s = o->g2_stats;
        s.parse_curve_device_ns += ev.ns(4, 6);
        s.subgroup_device_ns += ev.ns(6, 7);
        o->g2_experiment.repack_device_ns += ev.ns(6, 8);
        o->g2_experiment.subgroup_kernel_ns += ev.ns(8, 7);
        o->g2_experiment.logical_repack_bytes += o->g2_kernel.layout ? uint64_t(n) * 256 : 0;
        s.d2h_device_ns += ev.ns(7, 5);
        s.native_wall_ns += wall;
        s.checked_points += n;
        o->g2_traffic.scalar_h2d_bytes += n * 32;
        o->g2_traffic.point_d2h_bytes += n * 128;
        o->g2_traffic.status_d2h_bytes += sizeof(DvgpuG2ValidationReport);
        *report = *rep;
        if (rep->invalid_points) return fail(-5, "G2 content validation rejected", error, capacity);
        s.accepted_points += n;
        std::memcpy(output, staged, n * 128);
        return 0;
    });
}

// icicle_msm.cu:619-654
int dvgpu_bn254_g2_validation_inspect(void* backend, const uint8_t* wire, size_t n, uint32_t algorithm,
                                      uint32_t* masks, DvgpuG2ValidationReport* report, char* error,
                                      size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(n <= 4096 && algorithm >= 1 && algorithm <= 3 && report && (!n || (wire && masks)),
                "invalid G2 validation buffers, algorithm or chunk length");
        if (!n) {
            *report = {0, 0, UINT64_MAX, 0, algorithm};
            return 0;
        }
        const uint64_t t0 = now_ns();
        SetupState& st = setup_of(o);
        EventSet ev;
        uint8_t* dwire = st.points.as<uint8_t>(n * 128);
        uint32_t* dmasks = st.masks.as<uint32_t>(n);
        auto* drep = st.report.as<DvgpuG2ValidationReport>(1);
        uint8_t* staged = o->pinned.as<uint8_t>(n * 4 + sizeof(DvgpuG2ValidationReport));
        HIP_CHECK(hipEventRecord(ev.e[0], o->stream));
        HIP_CHECK(hipMemcpyAsync(dwire, wire, n * 128, hipMemcpyHostToDevice, o->stream));
        HIP_CHECK(hipEventRecord(ev.e[1], o->stream));
        G2CheckLaunch launch;
        launch.config = o->g2_kernel;
        if (launch.config.layout) launch.scratch = st.planes.as<uint8_t>(n * 128);
        launch.parsed = ev.e[2];
        launch.repacked = ev.e[6];
        launch.subgroup_done = ev.e[3];
        g2_validate_enqueue(dwire, n, algorithm, dmasks, drep, launch, o->stream);
        HIP_CHECK(hipEventRecord(ev.e[4], o->stream));
        HIP_CHECK(hipMemcpyAsync(staged, dmasks, n * 4, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipMemcpyAsync(staged + n * 4, drep, sizeof(DvgpuG2ValidationReport), hipMemcpyDeviceToHost,
                                 o->stream));
        HIP_CHECK(hipEventRecord(ev.e[5], o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        auto& s = o->g2_stats;
        s.h2d_device_ns += ev.ns(0, 1);
        s.parse_curve_device_ns += ev.ns(1, 2);
        s.subgroup_device_ns += ev.ns(2, 3);
        s.reduce_device_ns += ev.ns(3, 4);
        s.d2h_device_ns += ev.ns(4, 5);
        s.native_wall_ns += now_ns() - t0;
        s.checked_points += n;
        o->g2_experiment.repack_device_ns += ev.ns(2, 6);
        o->g2_experiment.subgroup_kernel_ns += ev.ns(6, 3);
        o->g2_experiment.logical_repack_bytes += o->g2_kernel.layout ? uint64_t(n) * 256 : 0;
        o->g2_traffic.point_h2d_bytes += n * 128;
        o->g2_traffic.status_d2h_bytes += n * 4 + sizeof(DvgpuG2ValidationReport);
        std::memcpy(masks, staged, n * 4);
        std::memcpy(report, staged + n * 4, sizeof(DvgpuG2ValidationReport));
        return 0;
    });
}

int dvgpu_g2_validation_stats(void* backend, DvgpuG2ValidationStats* stats) {
    Owner* o = as_owner(backend);
    if (!o || !stats) return -1;
    *stats = o->g2_stats;
    return 0;
}

int dvgpu_g2_validation_traffic(void* backend, DvgpuG2ValidationTraffic* stats) {
    Owner* o = as_owner(backend);
    if (!o || !stats) return -1;
    *stats = o->g2_traffic;
    return 0;
}

// icicle_msm.cu:543-556
int dvgpu_g2_kernel_config_set(void* backend, const DvgpuG2KernelConfig* config) {
    Owner* o = as_owner(backend);
    if (!o || o->retired) return -4;
    if (!config || (config->threads != 64 && config->threads != 128 && config->threads != 256) ||
        config->schedule > 1 || config->layout > 1 || config->fused > 1)
        return -1;
    o->g2_kernel = *config;
    return 0;
}

int dvgpu_g2_kernel_config_get(void* backend, DvgpuG2KernelConfig* config) {
    Owner* o = as_owner(backend);
    if (!o || !config) return -1;
    *config = o->g2_kernel;
    return 0;
}

int dvgpu_g2_experiment_stats(void* backend, DvgpuG2ExperimentStats* stats) {
    Owner* o = as_owner(backend);
    if (!o || !stats) return -1;
    *stats = o->g2_experiment;
    return 0;
}

// icicle_msm.cu:557-576
int dvgpu_g2_kernel_resources(void* backend, uint32_t algorithm, DvgpuG2KernelResources* output) {
    Owner* o = as_owner(backend);
    if (!o || o->retired) return -4;
    if (!output || algorithm < 1 || algorithm > 3) return -1;
    try {
        DvgpuG2KernelResources r{};
        g2_kernel_attributes(algorithm, o->g2_kernel.threads, r);
        *output = r;
        return 0;
    } catch (...) {
        o->retire();
        return -2;
    }
}

// icicle_msm.cu:577-582
int dvgpu_g2_test_corrupt_next(void* backend, uint64_t index) {
    Owner* o = as_owner(backend);
    if (!o || o->retired || index >= 4096) return -1;
    o->g2_corrupt_next = int64_t(index);
    return 0;
}

// dynaverse_hip_ext.h: window tables of resident bases.  Entry (i, w) of the
// scalar table is s_i 2^(c w) mod r; 2^(c w) < 2^253 < r for every c <= 20.
}  // extern "C"

namespace {
__global__ void k_table_scalars(const uint32_t* __restrict__ scalars, size_t n, int c, uint32_t windows,
                                uint32_t* __restrict__ out) {
    const size_t t = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= n * windows) return;
    const size_t i = t / windows;
    const uint32_t w = uint32_t(t % windows);
    uint32_t sw[8], pw[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < 8; ++k) sw[k] = scalars[i * 8 + k];
    const uint32_t bit = uint32_t(c) * w;
    pw[bit / 32] = 1u << (bit % 32);
    // (s R) * 2^(cw) * R^-1 = s 2^(cw): the canonical product.
    const Fr prod = canon(mul(to_mont<FrParams>(sw), unpack<FrParams>(pw)));
    uint32_t o[8];
    pack(prod, o);
    for (int k = 0; k < 8; ++k) out[t * 8 + k] = o[k];
}
}  // namespace

extern "C" {

int dvgpu_ext_window_table_device(void* backend, const void* scalars, size_t n, int window_bits, void* rows,
                                  char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(window_bits >= 4 && window_bits <= 20 && (!n || (scalars && rows)), "invalid window-table arguments");
        if (!n) return 0;
        const uint32_t windows = uint32_t((255 + window_bits - 1) / window_bits);
        require(n * windows <= (size_t(1) << 30), "window table too large");
        SetupState& st = setup_of(o);
        uint32_t* tab = o->stage_c.as<uint32_t>(n * windows * 8);
        k_table_scalars<<<grid_for(n * windows, 256), 256, 0, o->stream>>>(static_cast<const uint32_t*>(scalars), n,
                                                                          window_bits, windows, tab);
        HIP_CHECK_LAUNCH();
        fixed_base_mont_enqueue<FqOps>(st, tab, n * windows, static_cast<uint32_t*>(rows), o->stream);
        HIP_CHECK(hipStreamSynchronize(o->stream));
        return 0;
    });
}

// dynaverse_hip_ext.h: device-resident SRS generation.
int dvgpu_ext_fixed_base_device(void* backend, const void* scalars, size_t n, int g2, void* points,
                                char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require((g2 == 0 || g2 == 1) && (!n || (scalars && points)), "invalid fixed-base arguments");
        SetupState& st = setup_of(o);
        if (g2)
            fixed_base_mont_enqueue<Fq2Ops>(st, static_cast<const uint32_t*>(scalars), n,
                                            static_cast<uint32_t*>(points), o->stream);
        else
            fixed_base_mont_enqueue<FqOps>(st, static_cast<const uint32_t*>(scalars), n,
                                           static_cast<uint32_t*>(points), o->stream);
        HIP_CHECK(hipStreamSynchronize(o->stream));
        return 0;
    });
}

// icicle_msm.cu:750-765
int dvgpu_g2_psi_probe(void* backend, const uint8_t* points, size_t n, uint8_t* output, char* error,
                       size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(n <= 129 && (!n || (points && output)), "invalid psi probe buffers");
        if (!n) return 0;
        SetupState& st = setup_of(o);
        uint8_t* din = st.points.as<uint8_t>(n * 192 + n * 576);
        uint32_t* dmasks = st.masks.as<uint32_t>(n);
        uint8_t* staged = o->pinned.as<uint8_t>(n * 576 + n * 4);
        HIP_CHECK(hipMemcpyAsync(din, points, n * 192, hipMemcpyHostToDevice, o->stream));
        g2_psi_probe_enqueue(din, n, din + n * 192, dmasks, o->stream);
        HIP_CHECK(hipMemcpyAsync(staged, din + n * 192, n * 576, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipMemcpyAsync(staged + n * 576, dmasks, n * 4, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        const uint32_t* m = reinterpret_cast<const uint32_t*>(staged + n * 576);
        for (size_t i = 0; i < n; ++i)
            if (m[i]) throw InvalidArgument("invalid homogeneous psi input");
        std::memcpy(output, staged, n * 576);
        return 0;
    });
}

}  // extern "C"
