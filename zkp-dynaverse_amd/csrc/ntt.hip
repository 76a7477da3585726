// Natural-order radix-2 NTT over BN254 Fr (decimation in frequency).
//
// A transform of length 2^L runs as P = ceil(L / 11) LDS passes plus one
// permutation pass.  A pass owns a bit range [a, a + S) of the element index
// and runs those S butterfly stages on tiles of 2^S x C elements held in LDS
// (C consecutive low indices keep the global accesses contiguous).  The
// butterfly is (x, y) -> (x + y, (x - y) * w): x - y is formed limb-wise
// without a carry pass and fed straight into the Montgomery product.  The last
// pass is followed by a bit-reversal gather that also leaves Montgomery form
// (and applies 1/n on the inverse) with one product per element.
#include "ntt.hpp"

#include <cstring>

namespace dvgpu {

using namespace bn254;

namespace {

constexpr int TILE = 2048;  // elements per LDS tile (64 KiB)
constexpr int THREADS = 256;
constexpr int MAX_STAGES = 11;

__device__ __forceinline__ Fr ld8(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FrParams>(w);
}

__device__ __forceinline__ void st8(uint32_t* p, const Fr& v) {
    uint32_t w[8];
    pack(v, w);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// One DIF pass over index bits [a, a + S) of `rows` transforms of length 2^L.
// CANON_IN: the input holds canonical integers (converted on load).
template <bool CANON_IN>
__global__ void __launch_bounds__(THREADS) k_dif_pass(const uint32_t* __restrict__ in,
                                                      uint32_t* __restrict__ out, int L, int a, int S,
                                                      int C, const uint32_t* __restrict__ tw) {
    __shared__ uint4 lds[TILE * 2];
    const size_t n = size_t(1) << L;
    const uint32_t tile_elems = (1u << S) * C;
    const size_t tiles_per_row = n / tile_elems;
    const size_t g = blockIdx.x;
    const size_t row = g / tiles_per_row;
    const size_t t = g % tiles_per_row;
    const size_t lo_groups = (size_t(1) << a) / C;
    const size_t hi = t / lo_groups;
    const size_t lo0 = (t % lo_groups) * C;
    const size_t base = row * n + (hi << (a + S)) + lo0;

    for (uint32_t e = threadIdx.x; e < tile_elems; e += THREADS) {
        const uint32_t mid = e / C, c = e % C;
        const uint32_t* src = in + (base + (size_t(mid) << a) + c) * 8;
        uint4 v0 = reinterpret_cast<const uint4*>(src)[0];
        uint4 v1 = reinterpret_cast<const uint4*>(src)[1];
        if (CANON_IN) {
            const uint32_t w[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
            uint32_t o[8];
            pack(to_mont<FrParams>(w), o);
            v0 = make_uint4(o[0], o[1], o[2], o[3]);
            v1 = make_uint4(o[4], o[5], o[6], o[7]);
        }
        lds[2 * e] = v0;
        lds[2 * e + 1] = v1;
    }
    __syncthreads();

    const uint32_t half = tile_elems / 2;
    for (int sl = S - 1; sl >= 0; --sl) {
        const int s = a + sl;
        const uint32_t h = 1u << sl;
        for (uint32_t b = threadIdx.x; b < half; b += THREADS) {
            const uint32_t c = b % C, q = b / C;
            const uint32_t m1 = ((q >> sl) << (sl + 1)) | (q & (h - 1));
            const uint32_t m2 = m1 + h;
            const size_t j = (size_t(q & (h - 1)) << a) + lo0 + c;
            const Fr w = ld8(tw + (j << (NTT_MAX_LOG - 1 - s)) * 8);
            uint32_t* p1 = reinterpret_cast<uint32_t*>(&lds[2 * (m1 * C + c)]);
            uint32_t* p2 = reinterpret_cast<uint32_t*>(&lds[2 * (m2 * C + c)]);
            const Fr x = ld8(p1), y = ld8(p2);
            st8(p1, reduce(norm(add_w(x, y))));
            st8(p2, mul(sub_w(x, y), w));
        }
        __syncthreads();
    }

    for (uint32_t e = threadIdx.x; e < tile_elems; e += THREADS) {
        const uint32_t mid = e / C, c = e % C;
        uint32_t* dst = out + (base + (size_t(mid) << a) + c) * 8;
        reinterpret_cast<uint4*>(dst)[0] = lds[2 * e];
        reinterpret_cast<uint4*>(dst)[1] = lds[2 * e + 1];
    }
}

// out[row][k] = x[row][rev_L(k)] * scale, scale given as raw limbs: a plain
// integer leaves Montgomery form, a Montgomery value keeps it.
template <bool CANON_OUT>
__global__ void __launch_bounds__(THREADS) k_bitrev_scale(const uint32_t* __restrict__ in,
                                                          uint32_t* __restrict__ out, int L,
                                                          size_t total, Fr scale) {
    const size_t i = size_t(blockIdx.x) * THREADS + threadIdx.x;
    if (i >= total) return;
    const size_t n = size_t(1) << L;
    const size_t row = i >> L, k = i & (n - 1);
    const uint32_t src = L ? (__brev(uint32_t(k)) >> (32 - L)) : 0;
    Fr x = mul(ld8(in + (row * n + src) * 8), scale);
    if (CANON_OUT) x = canon(x);
    st8(out + i * 8, x);
}

__global__ void k_twiddles(const uint32_t* __restrict__ lo, const uint32_t* __restrict__ hi,
                           uint32_t* __restrict__ tw, uint32_t count) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    st8(tw + size_t(k) * 8, canon(mul(ld8(lo + (k & 1023) * 8), ld8(hi + (k >> 10) * 8))));
}

Fr fr_from_words(const uint32_t (&w)[8]) { return to_mont<FrParams>(w); }

bool fr_is_one(const Fr& x) {
    const Fr c = canon(x), one = canon(fp_one<FrParams>());
    for (int i = 0; i < LIMBS; ++i)
        if (c.l[i] != one.l[i]) return false;
    return true;
}

}  // namespace

bool ntt_root_has_order_2_20(const uint32_t (&root)[8]) {
    if (!is_canonical<FrParams>(root)) return false;
    Fr p = fr_from_words(root);
    for (int i = 0; i < 19; ++i) p = sqr(p);
    if (fr_is_one(p)) return false;
    return fr_is_one(sqr(p));
}

void ntt_domain_prepare(NttDomain& d, const uint32_t (&root)[8], hipStream_t stream) {
    if (d.ready && std::memcmp(d.root_words, root, 32) == 0) return;
    require(ntt_root_has_order_2_20(root), "NTT root must have order 2^20");
    d.ready = false;
    const uint32_t count = 1u << (NTT_MAX_LOG - 1);
    const Fr w = fr_from_words(root);
    const Fr wi = inv(w);
    DeviceBuffer stage;
    std::vector<uint32_t> host(size_t(1024 + 512) * 8 * 2);
    for (int dir = 0; dir < 2; ++dir) {
        const Fr base = dir ? wi : w;
        uint32_t* lo = host.data() + size_t(dir) * (1536 * 8);
        uint32_t* hi = lo + 1024 * 8;
        Fr acc = fp_one<FrParams>();
        for (int j = 0; j < 1024; ++j) {
            uint32_t o[8];
            pack(canon(acc), o);
            std::memcpy(lo + j * 8, o, 32);
            acc = mul(acc, base);
        }
        const Fr step = acc;  // base^1024
        acc = fp_one<FrParams>();
        for (int j = 0; j < 512; ++j) {
            uint32_t o[8];
            pack(canon(acc), o);
            std::memcpy(hi + j * 8, o, 32);
            acc = mul(acc, step);
        }
    }
    uint32_t* dev = stage.as<uint32_t>(host.size());
    HIP_CHECK(hipMemcpyAsync(dev, host.data(), host.size() * 4, hipMemcpyHostToDevice, stream));
    uint32_t* f = d.fwd.as<uint32_t>(size_t(count) * 8);
    uint32_t* b = d.inv.as<uint32_t>(size_t(count) * 8);
    k_twiddles<<<grid_for(count, 256), 256, 0, stream>>>(dev, dev + 1024 * 8, f, count);
    HIP_CHECK_LAUNCH();
    k_twiddles<<<grid_for(count, 256), 256, 0, stream>>>(dev + 1536 * 8, dev + 1536 * 8 + 1024 * 8, b, count);
    HIP_CHECK_LAUNCH();
    HIP_CHECK(hipStreamSynchronize(stream));  // `stage` and `host` die here
    std::memcpy(d.root_words, root, 32);
    d.ready = true;
}

static void run_rows(const NttDomain& d, const uint32_t* in, uint32_t* out, int L, size_t rows,
                     bool inverse, bool canonical, NttWorkspace& ws, hipStream_t stream) {
    require(d.ready, "NTT domain not prepared");
    require(L >= 0 && L <= NTT_MAX_LOG, "NTT length must be a power of two <= 2^20");
    const size_t n = size_t(1) << L;
    const size_t total = n * rows;
    if (!total) return;
    uint32_t* work = ws.a.as<uint32_t>(total * 8);
    const uint32_t* tw = static_cast<const uint32_t*>(inverse ? d.inv.ptr : d.fwd.ptr);
    if (L == 0) {
        HIP_CHECK(hipMemcpyAsync(out, in, total * 32, hipMemcpyDeviceToDevice, stream));
        return;
    }
    const int passes = (L + MAX_STAGES - 1) / MAX_STAGES;
    int top = L;
    bool first = true;
    for (int p = 0; p < passes; ++p) {
        const int remaining_passes = passes - p;
        const int S = (top + remaining_passes - 1) / remaining_passes;
        const int a = top - S;
        const int C = int(std::min<size_t>(size_t(1) << a, size_t(TILE >> S)));
        const size_t blocks = rows * (n / (size_t(1u << S) * C));
        if (first && canonical)
            k_dif_pass<true><<<unsigned(blocks), THREADS, 0, stream>>>(in, work, L, a, S, C, tw);
        else
            k_dif_pass<false><<<unsigned(blocks), THREADS, 0, stream>>>(first ? in : work, work, L, a, S, C, tw);
        HIP_CHECK_LAUNCH();
        first = false;
        top = a;
    }
    // Scale: canonical output wants a plain integer (1 or n^-1); Montgomery
    // output wants the Montgomery form of the same value.
    Fr scale = fp_zero<FrParams>();
    if (canonical) {
        if (!inverse) {
            scale.l[0] = 1;
        } else {
            uint32_t nw[8] = {0};
            nw[L / 32] = 1u << (L % 32);
            const Fr nm = to_mont<FrParams>(nw);
            const Fr ninv_m = inv(nm);
            uint32_t plain[8];
            from_mont(ninv_m, plain);
            scale = unpack<FrParams>(plain);
        }
    } else {
        if (!inverse) {
            scale = fp_one<FrParams>();
        } else {
            uint32_t nw[8] = {0};
            nw[L / 32] = 1u << (L % 32);
            scale = canon(inv(to_mont<FrParams>(nw)));
        }
    }
    if (canonical)
        k_bitrev_scale<true><<<grid_for(total, THREADS), THREADS, 0, stream>>>(work, out, L, total, scale);
    else
        k_bitrev_scale<false><<<grid_for(total, THREADS), THREADS, 0, stream>>>(work, out, L, total, scale);
    HIP_CHECK_LAUNCH();
}

void ntt_rows(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
              bool inverse, NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, true, ws, stream);
}

void ntt_rows_mont(const NttDomain& d, const uint32_t* in, uint32_t* out, int log_n, size_t rows,
                   bool inverse, NttWorkspace& ws, hipStream_t stream) {
    run_rows(d, in, out, log_n, rows, inverse, false, ws, stream);
}

}  // namespace dvgpu
