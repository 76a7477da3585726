// Opening-cache precompute: every single-point KZG opening of a polynomial
// over a radix-2 domain (reference: crates/protocol/src/kzg.rs:217-300,
// `precompute_single_openings` and `prepared_opening_basis`).
//
// For f of degree d the openings are the group DFT of H_j = sum_{l>j} f_l
// [tau^(l-1-j)], a Toeplitz product computed as a cyclic convolution of
// length M = next_pow2(2d - 1):
//   s   = DFT_M(f_d, ..., f_1, 0, ...)                (Fr NTT)
//   B   = DFT_M([tau^0], ..., [tau^(d-1)], 0, ...)    (G1 transform, retained)
//   c   = IDFT_M(B * s)                               (G1 transform)
//   H_j = c_(d-1-j), j < d                             (truncate + reverse)
//   out = DFT_n(H) (times offset^j first on a coset)   (G1 transform)
// and every output point is normalised to affine.
//
// G1 transforms: radix-2 with the butterfly scalar (the twiddle) applied by
// GLV double-and-add.  A transform of 2^L points is split as 2^k x 2^(L-k)
// (k = 6 from L = 12): the stages over the high k index bits and the stages
// over the low L-k bits are separated by one diagonal of twiddles
// omega^(i2 rev(p1)), so that within every stage the 64 lanes of a wave share
// one twiddle and the double-and-add branches on a scalar digit.  DIF runs
// natural order to bit-reversed order, DIT the reverse, so the basis, the
// pointwise product and the two inverse/forward transforms need no
// permutation pass: consumers index the bit-reversed positions directly.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/dynaverse_kzg.h"
#include "g1mem.hpp"
#include "gfft.hpp"
#include "glv.hpp"
#include "owner.hpp"

namespace dvgpu {

struct KzgSrs {
    Owner* owner = nullptr;
    std::mutex mu;               // one call on this SRS at a time
    int users = 0;               // calls registered on it (under the owner's gate lock)
    uint64_t count = 0;          // SRS powers
    DeviceBuffer powers;         // canonical affine, 16 words each
    NttDomain domain;            // Fr twiddles omega_(2^20)^x, x < 2^19
    NttWorkspace ntt;
    uint64_t basis_degree = 0;   // degree of the retained basis (0 = none)
    uint64_t basis_builds = 0;
    DeviceBuffer basis;          // M XYZZ points, bit-reversed DFT order
    DeviceBuffer coeffs, scal, conv, h, out;
    void attach(Quota* q) {
        powers.quota = basis.quota = coeffs.quota = scal.quota = conv.quota = h.quota = out.quota = q;
        domain.set_quota(q);
        ntt.attach(q);
    }
};

void kzg_destroy(KzgSrs* s) noexcept { delete s; }

namespace {

constexpr unsigned GFFT_THREADS = 64;  // one wave per block: twiddles are wave-uniform

__device__ __forceinline__ uint32_t rev_bits(uint32_t x, int bits) { return bits ? __brev(x) >> (32 - bits) : 0u; }

__device__ __forceinline__ G1 twiddle_mul(const G1& p, const uint32_t* __restrict__ tw, uint32_t x, bool uniform) {
    uint32_t k[8];
    from_mont(ldfr(tw + size_t(x) * 8), k);
    return g1_mul_scalar(p, k, uniform);
}

// One radix-2 stage over bit s of rows x 2^L points.  The twiddle exponent of
// a butterfly is e = bits [lo, s) of its index, with root omega_(2^(s+1-lo));
// threads run over the other index bits first, so consecutive lanes share e.
template <bool DIT>
__global__ void __launch_bounds__(GFFT_THREADS) k_gfft_stage(uint32_t* __restrict__ data, const uint32_t* __restrict__ tw,
                                                        int L, int s, int lo, uint64_t rows) {
    const uint64_t t = uint64_t(blockIdx.x) * GFFT_THREADS + threadIdx.x;
    if (t >= (rows << (L - 1))) return;
    const int hb = L - 1 - s;
    const uint64_t rest = rows << (lo + hb);
    const uint64_t rid = t % rest, e = t / rest;
    const uint64_t low = rid & ((uint64_t(1) << lo) - 1), high = (rid >> lo) & ((uint64_t(1) << hb) - 1);
    const uint64_t row = rid >> (lo + hb);
    const uint64_t i = (row << L) | low | (e << lo) | (high << (s + 1)), j = i + (uint64_t(1) << s);
    const uint32_t x = uint32_t(e << (NTT_MAX_LOG - 1 - s + lo));  // omega_(2^20)^x
    const bool uniform = __all(x == __builtin_amdgcn_readfirstlane(x));
    if (DIT) {
        G1 b = ldg1(data + j * 32);
        if (x) b = twiddle_mul(b, tw, x, uniform);
        const G1 a = ldg1(data + i * 32);  // loaded after the product: fewer live registers
        stg1(data + i * 32, xyzz_add<FqOps>(a, b));
        stg1(data + j * 32, xyzz_add<FqOps>(a, xyzz_neg<FqOps>(b)));
    } else {
        const G1 a = ldg1(data + i * 32), b = ldg1(data + j * 32);
        const G1 d = xyzz_add<FqOps>(a, xyzz_neg<FqOps>(b));
        stg1(data + i * 32, xyzz_add<FqOps>(a, b));
        stg1(data + j * 32, x ? twiddle_mul(d, tw, x, uniform) : d);
    }
}

// Diagonal between the two stage groups: position p = p1 2^m + i2 (m = L - k)
// times omega_(2^L)^(i2 rev_k(p1)) (tw is the inverse table for DIT).
__global__ void __launch_bounds__(GFFT_THREADS) k_gfft_diag(uint32_t* __restrict__ data, const uint32_t* __restrict__ tw,
                                                       int L, int k, uint64_t rows) {
    const uint64_t t = uint64_t(blockIdx.x) * GFFT_THREADS + threadIdx.x;
    if (t >= (rows << L)) return;
    const int m = L - k;
    const uint32_t p = uint32_t(t & ((uint64_t(1) << L) - 1));
    const uint32_t e = ((p & ((1u << m) - 1)) * rev_bits(p >> m, k)) & ((1u << L) - 1);
    if (!e) return;
    uint32_t x = e << (NTT_MAX_LOG - L);
    const bool negate = x >> (NTT_MAX_LOG - 1);  // omega^(x + 2^19) = -omega^x
    x &= (1u << (NTT_MAX_LOG - 1)) - 1;
    G1 v = ldg1(data + t * 32);
    if (negate) v = xyzz_neg<FqOps>(v);
    if (x) v = twiddle_mul(v, tw, x, false);
    stg1(data + t * 32, v);
}

// (f_d, ..., f_1, 0, ...) per row in Montgomery form; coefficients canonical.
__global__ void k_kzg_reversed(const uint32_t* __restrict__ coeffs, uint64_t count, uint64_t rows, int LM,
                               uint32_t* __restrict__ out) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (rows << LM)) return;
    const uint64_t row = t >> LM, p = t & ((uint64_t(1) << LM) - 1), degree = count - 1;
    Fr v = fp_zero<FrParams>();
    if (p < degree) {
        uint32_t w[8];
        const uint32_t* src = coeffs + (row * count + degree - p) * 8;
        for (int q = 0; q < 8; ++q) w[q] = src[q];
        v = canon(to_mont<FrParams>(w));
    }
    stfr(out + t * 8, v);
}

// ([tau^0], ..., [tau^(d-1)], inf, ...) from the canonical affine powers.
__global__ void k_kzg_basis_in(const uint32_t* __restrict__ powers, uint64_t degree, int LM, uint32_t* __restrict__ out) {
    const uint64_t p = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= (uint64_t(1) << LM)) return;
    G1 v = xyzz_inf<FqOps>();
    if (p < degree) {
        uint32_t x[8], y[8], any = 0;
        for (int q = 0; q < 8; ++q) {
            x[q] = powers[p * 16 + q];
            y[q] = powers[p * 16 + 8 + q];
            any |= x[q] | y[q];
        }
        if (any) v = {canon(to_mont<FqParams>(x)), canon(to_mont<FqParams>(y)), fp_one<FqParams>(), fp_one<FqParams>()};
    }
    stg1(out + p * 32, v);
}

// conv[row][p] = B[p] * (s[row][rev(p)] / M): both operands bit-reversed.
__global__ void __launch_bounds__(GFFT_THREADS) k_kzg_pointwise(const uint32_t* __restrict__ basis,
                                                           const uint32_t* __restrict__ scal, uint64_t rows, int LM,
                                                           Fr minv, uint32_t* __restrict__ out) {
    const uint64_t t = uint64_t(blockIdx.x) * GFFT_THREADS + threadIdx.x;
    if (t >= (rows << LM)) return;
    const uint64_t row = t >> LM;
    const uint32_t p = uint32_t(t & ((uint64_t(1) << LM) - 1));
    uint32_t k[8];
    from_mont(mul(ldfr(scal + ((row << LM) + rev_bits(p, LM)) * 8), minv), k);
    stg1(out + t * 32, g1_mul_scalar(ldg1(basis + size_t(p) * 32), k, false));
}

// H[row][j] = c[row][d - 1 - j] for j < d, infinity above.
__global__ void k_kzg_truncate(const uint32_t* __restrict__ conv, uint64_t rows, int LM, uint64_t degree, int Ln,
                               uint32_t* __restrict__ h) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (rows << Ln)) return;
    const uint64_t row = t >> Ln, j = t & ((uint64_t(1) << Ln) - 1);
    G1 v = xyzz_inf<FqOps>();
    if (j < degree) v = ldg1(conv + ((row << LM) + degree - 1 - j) * 32);
    stg1(h + t * 32, v);
}

// H[row][j] *= offset^j (evaluation on the coset offset * <omega>).
__global__ void __launch_bounds__(GFFT_THREADS) k_kzg_coset(uint32_t* __restrict__ h, uint64_t rows, int Ln, Fr offset) {
    const uint64_t t = uint64_t(blockIdx.x) * GFFT_THREADS + threadIdx.x;
    if (t >= (rows << Ln)) return;
    const uint64_t j = t & ((uint64_t(1) << Ln) - 1);
    if (!j) return;
    uint32_t k[8];
    from_mont(fr_pow64(offset, j), k);
    stg1(h + t * 32, g1_mul_scalar(ldg1(h + t * 32), k, false));
}

// openings[row][i] = affine(H_hat[row][rev(i)]), canonical 64 B (0 = infinity).
__global__ void k_kzg_affine(const uint32_t* __restrict__ h, uint64_t rows, int Ln, uint32_t* __restrict__ out) {
    const uint64_t t = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= (rows << Ln)) return;
    const uint64_t row = t >> Ln;
    const uint32_t i = uint32_t(t & ((uint64_t(1) << Ln) - 1));
    const Affine<FqOps> a = xyzz_to_affine<FqOps>(ldg1(h + ((row << Ln) + rev_bits(i, Ln)) * 32));
    uint32_t* o = out + t * 16;
    uint32_t w[8] = {0};
    if (!a.inf) from_mont(a.x, w);
    for (int q = 0; q < 8; ++q) o[q] = w[q];
    if (!a.inf) from_mont(a.y, w);
    for (int q = 0; q < 8; ++q) o[8 + q] = w[q];
}

// Index bits of the high stage group: 6 keeps every stage wave-uniform from
// 2^12 points on; below that the stages split the bits evenly.
int gfft_split(int L, uint64_t rows) {
    if (rows >= 64) return 0;
    return L >= 12 ? 6 : L / 2;
}

}  // namespace

// rows x 2^L XYZZ points in place.  Forward (DIF): natural -> bit-reversed.
// Inverse (DIT, unscaled, inverse root): bit-reversed -> natural.
void gfft_rows(const NttDomain& d, uint32_t* data, int L, uint64_t rows, bool inverse, hipStream_t st) {
    if (!L || !rows) return;
    const int k = gfft_split(L, rows), m = L - k;
    const auto* tw = static_cast<const uint32_t*>(inverse ? d.inv.pow.ptr : d.fwd.pow.ptr);
    const unsigned stage_grid = grid_for(rows << (L - 1), GFFT_THREADS), diag_grid = grid_for(rows << L, GFFT_THREADS);
    if (!inverse) {
        for (int s = L - 1; s >= m; --s) {
            k_gfft_stage<false><<<stage_grid, GFFT_THREADS, 0, st>>>(data, tw, L, s, m, rows);
            HIP_CHECK_LAUNCH();
        }
        if (k && m) {
            k_gfft_diag<<<diag_grid, GFFT_THREADS, 0, st>>>(data, tw, L, k, rows);
            HIP_CHECK_LAUNCH();
        }
        for (int s = m - 1; s >= 0; --s) {
            k_gfft_stage<false><<<stage_grid, GFFT_THREADS, 0, st>>>(data, tw, L, s, 0, rows);
            HIP_CHECK_LAUNCH();
        }
    } else {
        for (int s = 0; s < m; ++s) {
            k_gfft_stage<true><<<stage_grid, GFFT_THREADS, 0, st>>>(data, tw, L, s, 0, rows);
            HIP_CHECK_LAUNCH();
        }
        if (k && m) {
            k_gfft_diag<<<diag_grid, GFFT_THREADS, 0, st>>>(data, tw, L, k, rows);
            HIP_CHECK_LAUNCH();
        }
        for (int s = m; s < L; ++s) {
            k_gfft_stage<true><<<stage_grid, GFFT_THREADS, 0, st>>>(data, tw, L, s, m, rows);
            HIP_CHECK_LAUNCH();
        }
    }
}

namespace {

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

int log2_pow2(uint64_t n) {
    int l = 0;
    while ((uint64_t(1) << l) < n) ++l;
    return l;
}

bool fr_canonical(const uint8_t* b) {
    uint32_t w[8];
    std::memcpy(w, b, 32);
    return is_canonical<FrParams>(w);
}

KzgSrs* registered(Owner* o, void* srs) {
    auto it = std::find(o->kzg.begin(), o->kzg.end(), static_cast<KzgSrs*>(srs));
    return it == o->kzg.end() ? nullptr : *it;
}

// One KZG call on one SRS.  It registers with the owner's gate (a retirement
// on another thread then waits for it to leave) as a user of its SRS
// (destroying the SRS waits for its users), and holds the SRS's lock.  A
// call that gets the lock only after another thread retired the owner
// leaves with status -4 without touching the device.
struct KzgRegistration {
    Owner* o;
    KzgSrs* z = nullptr;
    bool entered = false;
    KzgRegistration(Owner* owner, void* handle) : o(owner) {
        entered = o->gate.enter([&] {
            z = registered(o, handle);
            if (z) ++z->users;
        });
    }
    ~KzgRegistration() {
        if (entered) o->gate.leave([&] {
            if (z) --z->users;
        });
    }
    KzgRegistration(const KzgRegistration&) = delete;
    KzgRegistration& operator=(const KzgRegistration&) = delete;
};

class KzgCall {
  public:
    KzgCall(Owner* o, void* handle) : reg_(o, handle) {
        if (!reg_.entered) throw OwnerRetired();
        require(reg_.z != nullptr, "unknown KZG SRS handle for this backend");
        lock_ = std::unique_lock<std::mutex>(reg_.z->mu);
        if (o->retired) throw OwnerRetired();
    }
    KzgSrs& srs() const { return *reg_.z; }

  private:
    KzgRegistration reg_;  // declared first: released last
    std::unique_lock<std::mutex> lock_;
};

void precompute(KzgSrs& z, const uint8_t* coefficients, uint64_t coeff_count, uint64_t rows, uint64_t domain_size,
                const uint8_t* offset, uint8_t* openings, DvgpuKzgTimings* timings) {
    // Argument order follows kzg.rs:222-228: domain bound, SRS bound, constants.
    require(domain_size && !(domain_size & (domain_size - 1)) && domain_size <= DVGPU_MAX_NTT_SIZE,
            "opening domain size must be a power of two up to 2^20");
    require(coeff_count <= domain_size, "cached polynomial degree must be below domain size");
    require(coeff_count <= z.count, "polynomial exceeds KZG SRS degree bound");
    require(!rows || openings, "null openings buffer");
    require(!rows || !coeff_count || coefficients, "null coefficients");
    const uint64_t degree = coeff_count ? coeff_count - 1 : 0;
    if (!rows) return;
    Fr off = fp_one<FrParams>();
    if (offset) {
        require(fr_canonical(offset), "coset offset is not a canonical Fr element");
        uint32_t w[8];
        std::memcpy(w, offset, 32);
        off = canon(to_mont<FrParams>(w));
        require(!is_zero(off), "coset offset must be nonzero");
    }
    if (degree == 0) {  // kzg.rs:230-232: every opening is the identity
        std::memset(openings, 0, rows * domain_size * 64);
        return;
    }
    for (uint64_t i = 0; i < rows * coeff_count; ++i)
        require(fr_canonical(coefficients + 32 * i), "coefficient is not a canonical Fr element");
    uint64_t M = 1;
    while (M < 2 * degree - 1) M <<= 1;
    require(M <= DVGPU_MAX_NTT_SIZE, "opening convolution exceeds 2^20 points");
    require(rows <= (uint64_t(1) << 24) / M && rows <= (uint64_t(1) << 24) / domain_size,
            "too many polynomials for one call (rows x 2^20 points at most 2^24)");
    const int LM = log2_pow2(M), Ln = log2_pow2(domain_size);
    Owner* o = z.owner;
    hipStream_t st = o->stream;
    ScopedEvents<4> ev(timings != nullptr);
    if (timings) HIP_CHECK(hipEventRecord(ev[0], st));
    auto* cw = z.coeffs.as<uint32_t>(rows * coeff_count * 8);
    HIP_CHECK(hipMemcpyAsync(cw, coefficients, rows * coeff_count * 32, hipMemcpyHostToDevice, st));
    auto* sw = z.scal.as<uint32_t>((rows << LM) * 8);
    k_kzg_reversed<<<grid_for(rows << LM, 256), 256, 0, st>>>(cw, coeff_count, rows, LM, sw);
    HIP_CHECK_LAUNCH();
    if (LM) ntt_rows_mont(z.domain, sw, sw, LM, rows, false, z.ntt, st);
    if (z.basis_degree != degree) {  // kzg.rs:277-300: one retained degree
        z.basis_degree = 0;
        auto* bw = z.basis.as<uint32_t>(M * 32);
        k_kzg_basis_in<<<grid_for(M, 256), 256, 0, st>>>(static_cast<const uint32_t*>(z.powers.ptr), degree, LM, bw);
        HIP_CHECK_LAUNCH();
        gfft_rows(z.domain, bw, LM, 1, false, st);
        z.basis_degree = degree;
        ++z.basis_builds;
    }
    if (timings) HIP_CHECK(hipEventRecord(ev[1], st));
    uint32_t mw[8] = {0};
    mw[LM / 32] = 1u << (LM % 32);
    const Fr minv = canon(inv(to_mont<FrParams>(mw)));
    auto* conv = z.conv.as<uint32_t>((rows << LM) * 32);
    k_kzg_pointwise<<<grid_for(rows << LM, GFFT_THREADS), GFFT_THREADS, 0, st>>>(
        static_cast<const uint32_t*>(z.basis.ptr), sw, rows, LM, minv, conv);
    HIP_CHECK_LAUNCH();
    gfft_rows(z.domain, conv, LM, rows, true, st);
    auto* hw = z.h.as<uint32_t>((rows << Ln) * 32);
    k_kzg_truncate<<<grid_for(rows << Ln, 256), 256, 0, st>>>(conv, rows, LM, degree, Ln, hw);
    HIP_CHECK_LAUNCH();
    if (offset) {
        k_kzg_coset<<<grid_for(rows << Ln, GFFT_THREADS), GFFT_THREADS, 0, st>>>(hw, rows, Ln, off);
        HIP_CHECK_LAUNCH();
    }
    if (timings) HIP_CHECK(hipEventRecord(ev[2], st));
    gfft_rows(z.domain, hw, Ln, rows, false, st);
    auto* ow = z.out.as<uint32_t>((rows << Ln) * 16);
    k_kzg_affine<<<grid_for(rows << Ln, 256), 256, 0, st>>>(hw, rows, Ln, ow);
    HIP_CHECK_LAUNCH();
    if (timings) HIP_CHECK(hipEventRecord(ev[3], st));
    HIP_CHECK(hipMemcpyAsync(openings, ow, (rows << Ln) * 64, hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    if (timings) {
        timings->scalar_and_basis_ms = ev.ms(0, 1);
        timings->convolution_ms = ev.ms(1, 2);
        timings->domain_transform_ms = ev.ms(2, 3);
    }
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" {

int dvgpu_kzg_srs_create(void* backend, const uint8_t* g1_powers, size_t count, const uint8_t* root_2_20, void** srs,
                         char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(srs && root_2_20 && (g1_powers || !count), "invalid KZG SRS arguments");
        require(count <= (uint64_t(1) << 28), "KZG SRS too large");
        uint32_t root[8];
        std::memcpy(root, root_2_20, 32);
        require(ntt_root_has_order_2_20(root), "root does not have order 2^20");
        for (size_t i = 0; i < count; ++i) {
            uint32_t x[8], y[8];
            std::memcpy(x, g1_powers + 64 * i, 32);
            std::memcpy(y, g1_powers + 64 * i + 32, 32);
            bool any = false;
            for (int q = 0; q < 8; ++q) any |= (x[q] | y[q]) != 0;
            if (!any) continue;
            require(is_canonical<FqParams>(x) && is_canonical<FqParams>(y), "SRS point coordinate not canonical");
            const Affine<FqOps> a = {to_mont<FqParams>(x), to_mont<FqParams>(y), false};
            require(affine_on_curve<FqOps>(a), "SRS point is not on G1");
        }
        auto z = std::make_unique<KzgSrs>();
        z->owner = o;
        z->attach(&o->quota);
        z->count = count;
        if (count) {
            void* p = z->powers.ensure(count * 64);
            HIP_CHECK(hipMemcpyAsync(p, g1_powers, count * 64, hipMemcpyHostToDevice, o->stream));
        }
        ntt_domain_prepare(z->domain, root, o->stream);
        HIP_CHECK(hipStreamSynchronize(o->stream));
        o->gate.locked([&] { o->kzg.push_back(z.get()); });
        *srs = z.release();
        return 0;
    });
}

int dvgpu_kzg_srs_destroy(void* backend, void* srs) {
    Owner* o = as_owner(backend);
    if (!o) return -1;
    KzgSrs* z = nullptr;
    o->gate.locked([&] {
        auto it = std::find(o->kzg.begin(), o->kzg.end(), static_cast<KzgSrs*>(srs));
        if (it == o->kzg.end()) return;
        z = *it;
        o->kzg.erase(it);
    });
    if (!z) return -1;
    o->gate.wait([&] { return z->users == 0; });  // calls already on it finish first
    if (o->stream) (void)hipStreamSynchronize(o->stream);
    kzg_destroy(z);
    return 0;
}

int dvgpu_kzg_srs_basis(void* backend, void* srs, uint64_t* degree, uint64_t* builds) {
    Owner* o = as_owner(backend);
    if (!o || o->retired) return -1;
    try {
        KzgCall call(o, srs);
        if (degree) *degree = call.srs().basis_degree;
        if (builds) *builds = call.srs().basis_builds;
        return 0;
    } catch (const std::exception&) {
        return -1;
    }
}

int dvgpu_kzg_precompute_single_openings(void* backend, void* srs, const uint8_t* coefficients, size_t coeff_count,
                                         size_t poly_count, size_t domain_size, const uint8_t* coset_offset,
                                         uint8_t* openings, DvgpuKzgTimings* timings, char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        KzgCall call(o, srs);
        precompute(call.srs(), coefficients, coeff_count, poly_count, domain_size, coset_offset, openings, timings);
        return 0;
    });
}

int dvgpu_ext_glv_split(const uint8_t* scalars, size_t count, uint8_t* out) {
    if (count && (!scalars || !out)) return -1;
    for (size_t i = 0; i < count; ++i) {
        uint32_t k[8];
        std::memcpy(k, scalars + 32 * i, 32);
        if (!is_canonical<FrParams>(k)) return -1;
        const GlvSplit s = bn254::glv_split(k);
        std::memcpy(out + 40 * i, s.k1, 16);
        std::memcpy(out + 40 * i + 16, s.k2, 16);
        std::memcpy(out + 40 * i + 32, &s.neg1, 4);
        std::memcpy(out + 40 * i + 36, &s.neg2, 4);
    }
    return 0;
}

int dvgpu_ext_glv_jsf(const uint8_t* scalars, size_t count, uint8_t* out) {
    if (count && (!scalars || !out)) return -1;
    for (size_t i = 0; i < count; ++i) {
        uint32_t k[8];
        std::memcpy(k, scalars + 32 * i, 32);
        if (!is_canonical<FrParams>(k)) return -1;
        GlvSplit s = bn254::glv_split(k);
        const bn254::Jsf j = bn254::jsf_recode(s.k1, s.k2);
        std::memcpy(out + 64 * i, j.nz1, 16);
        std::memcpy(out + 64 * i + 16, j.sg1, 16);
        std::memcpy(out + 64 * i + 32, j.nz2, 16);
        std::memcpy(out + 64 * i + 48, j.sg2, 16);
    }
    return 0;
}

}  // extern "C"
