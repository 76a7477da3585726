// The NTT pass body (one LDS tile of sub-transforms), its twiddle tables and
// the pass geometry, written once for the device and for a host build.
//
// On the device every loop is strided by the block's threads and every round
// ends with a barrier.  The host build runs the same code with one emulated
// thread per block: within a round every group reads and writes its own LDS
// slots, so running the groups one after the other gives the same result.
#pragma once
#include <stdint.h>
#include <stddef.h>
#include <algorithm>

#include "field.hpp"

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

#if defined(__HIP_DEVICE_COMPILE__)
#define NTT_TID threadIdx.x
#define NTT_BDIM blockDim.x
#define NTT_SYNC() __syncthreads()
#define NTT_BREV(x) __brev(x)
#define NTT_CLZ(x) __clz(x)
#else
#define NTT_TID 0u
#define NTT_BDIM 1u
#define NTT_SYNC() ((void)0)
#define NTT_BREV(x) ::dvgpu::ntt_host_brev(x)
#define NTT_CLZ(x) __builtin_clz(x)
#endif

namespace dvgpu {

inline uint32_t ntt_host_brev(uint32_t x) {
    uint32_t r = 0;
    for (int i = 0; i < 32; ++i) r |= ((x >> i) & 1u) << (31 - i);
    return r;
}

namespace ntt {

using namespace bn254;

constexpr int MAX_LOG = 20;          // twiddle powers omega^k, k < 2^19, of the 2^20-th root
constexpr int SINGLE_MAX_LOG = 11;   // lengths up to 2^11 run as one LDS pass
constexpr int STAGE_TABLE_LOG = 11;  // shared stage table: local stages s < 11
constexpr int MAX_PASSES = 2;
constexpr uint32_t TILE = 2048;      // elements per LDS tile (72 KiB as nine limb planes)
constexpr int TILE_LOG = 11;
constexpr uint32_t TARGET_BLOCKS = 512;       // a launch widens its tiles only beyond this many
constexpr uint64_t R4_MIN_GROUPS = 1u << 17;  // radix-4 rounds from this many four-point groups

enum PassMode : uint32_t { PASS_ROWS = 0, PASS_COLS = 1, PASS_XPOSE = 2 };

// One pass of one entry, as the kernels read it.
struct NttPass {
    const uint32_t* in;
    uint32_t* out;
    const uint32_t* tw;    // stage table (rows, transpose) or the length's pass-1 table (columns)
    uint32_t mode, L, S;   // pass kind, log2 transform length, stages of this pass
    uint32_t logU;         // sub-transforms per tile, log2
    uint32_t count;        // rows: rows in total; columns / transpose: transforms
    uint32_t tiles;        // rows: tiles in total; columns / transpose: tiles per transform
    uint32_t logCT;        // columns: log2 of the pass-1 table's columns per group
    uint32_t scaled;       // outputs times `scale` (the inverse's 1/n, last pass)
    uint32_t scale[LIMBS]; // Montgomery limbs
};

HD Fr ld8(const uint32_t* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#else
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
#endif
    return unpack<FrParams>(w);
}

HD void st8(uint32_t* p, const Fr& v) {
    uint32_t w[8];
    pack(v, w);
#if defined(__HIP_DEVICE_COMPILE__)
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
#else
    for (int i = 0; i < 8; ++i) p[i] = w[i];
#endif
}

HD Fr lds_ld(const uint32_t* __restrict__ lds, uint32_t e) {
    Fr x;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) x.l[k] = lds[k * TILE + e];
    return x;
}

HD void lds_st(uint32_t* __restrict__ lds, uint32_t e, const Fr& x) {
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) lds[k * TILE + e] = x.l[k];
}

// Nine limb planes at p, plane stride `stride` words.
HD Fr planes_ld(const uint32_t* __restrict__ p, uint32_t stride) {
    Fr x;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) x.l[k] = p[k * stride];
    return x;
}

// Stage table: omega_(2^(s+1))^j, segment s at (2^s - 1) * 9 words.
HD Fr tw_stage(const uint32_t* __restrict__ tw, uint32_t s, uint32_t j) {
    return planes_ld(tw + ((1u << s) - 1) * LIMBS + j, 1u << s);
}

// Pass-1 table of a length: column x2, local stage s, index j.  Columns come
// in groups of CT = 2^logCT; a group holds its stages' segments one after the
// other, each as nine planes of CT 2^s words (entry j * CT + column).
HD Fr tw_cols(const uint32_t* __restrict__ tw, uint32_t S, uint32_t logCT, uint32_t s, uint32_t j, uint32_t x2) {
    const uint32_t ct = 1u << logCT;
    const uint32_t group = ct * ((1u << S) - 1) * LIMBS;
    return planes_ld(tw + (x2 >> logCT) * group + ct * ((1u << s) - 1) * LIMBS + j * ct + (x2 & (ct - 1)),
                     ct << s);
}

HD Fr twiddle(const NttPass& p, uint32_t s, uint32_t j, uint32_t x2) {
    return p.mode == PASS_COLS ? tw_cols(p.tw, p.S, p.logCT, s, j, x2) : tw_stage(p.tw, s, j);
}

// Blocks are dealt round-robin to the eight XCDs; this order gives each XCD a
// run of consecutive tiles, whose runs of U elements share cache lines.
HD uint32_t xcd_order(uint32_t b, uint32_t blocks) {
    return (blocks & 7) ? b : (b & 7) * (blocks >> 3) + (b >> 3);
}

HD bool unit2(const NttPass& p, int s) { return p.mode != PASS_COLS && s == 0; }

HD Fr tw2_of(const NttPass& p, uint32_t t, int s, uint32_t first) {
    const uint32_t u = t & ((1u << p.logU) - 1), q = t >> p.logU, j = q & ((1u << s) - 1);
    return unit2(p, s) ? fp_zero<FrParams>() : twiddle(p, s, j, first + u);
}

// Radix-2 butterfly t at local stage s with its twiddle w given (unused at
// unit stages).
HD void bfly2(uint32_t* __restrict__ lds, const NttPass& p, uint32_t t, int s, const Fr& w) {
    const uint32_t logU = p.logU, h = 1u << s;
    const uint32_t u = t & ((1u << logU) - 1), q = t >> logU;
    const uint32_t j = q & (h - 1);
    const uint32_t m1 = ((q >> s) << (s + 1)) | j;
    const uint32_t e1 = (m1 << logU) + u, e2 = e1 + (h << logU);
    const Fr x = lds_ld(lds, e1), y = lds_ld(lds, e2);
    lds_st(lds, e1, reduce(norm(add_w(x, y))));
    lds_st(lds, e2, unit2(p, s) ? reduce(norm(sub_w(x, y))) : mul(sub_w(x, y), w));
}

// One radix-2 round at local stage s (each butterfly loads its twiddle after
// its LDS reads).
HD void round2(uint32_t* __restrict__ lds, const NttPass& p, uint32_t tile, uint32_t first, int s) {
    const uint32_t logU = p.logU, h = 1u << s;
    const bool unit = unit2(p, s);
    for (uint32_t t = NTT_TID; t < (tile >> 1); t += NTT_BDIM) {
        const uint32_t u = t & ((1u << logU) - 1), q = t >> logU;
        const uint32_t j = q & (h - 1);
        const uint32_t m1 = ((q >> s) << (s + 1)) | j;
        const uint32_t e1 = (m1 << logU) + u, e2 = e1 + (h << logU);
        const Fr x = lds_ld(lds, e1), y = lds_ld(lds, e2);
        lds_st(lds, e1, reduce(norm(add_w(x, y))));
        lds_st(lds, e2, unit ? reduce(norm(sub_w(x, y))) : mul(sub_w(x, y), twiddle(p, s, j, first + u)));
    }
    NTT_SYNC();
}

// One radix-4 round over local stages s and s - 1 (s >= 1): the group at
// positions m0 + {0, h2, 2 h2, 3 h2}, h2 = 2^(s-1).  Inputs < 2m; the two
// sums of stage s stay unreduced (< 4m) into stage s - 1.  (The twiddles are
// loaded after the group's LDS reads: loading them first keeps them live
// longer, and the extra registers cost the pass occupancy.)
HD void round4(uint32_t* __restrict__ lds, const NttPass& p, uint32_t tile, uint32_t first, int s) {
    const uint32_t logU = p.logU, h2 = 1u << (s - 1);
    const bool unit = p.mode != PASS_COLS && s == 1;  // omega_4^0 and omega_2^0
    for (uint32_t t = NTT_TID; t < (tile >> 2); t += NTT_BDIM) {
        const uint32_t u = t & ((1u << logU) - 1), q = t >> logU;
        const uint32_t j0 = q & (h2 - 1);
        const uint32_t m0 = ((q >> (s - 1)) << (s + 1)) | j0;
        const uint32_t e0 = (m0 << logU) + u, d = h2 << logU;
        const Fr x0 = lds_ld(lds, e0), x1 = lds_ld(lds, e0 + d), x2 = lds_ld(lds, e0 + 2 * d),
                 x3 = lds_ld(lds, e0 + 3 * d);
        // stage s: (x0, x2) with omega^j0, (x1, x3) with omega^(j0 + h2)
        const Fr a0 = norm(add_w(x0, x2));  // < 4m
        const Fr a1 = norm(add_w(x1, x3));  // < 4m
        const Fr a3 = mul(sub_w(x1, x3), twiddle(p, s, j0 + h2, first + u));
        Fr y0, y1, y2, y3;
        if (unit) {
            const Fr a2 = norm(sub_w(x0, x2));  // < 6m
            y0 = reduce(norm(add_w(a0, a1)));
            y1 = reduce(norm(sub_w(a0, a1)));
            y2 = reduce(norm(add_w(a2, a3)));
            y3 = reduce(norm(sub_w(a2, a3)));
        } else {
            const Fr a2 = mul(sub_w(x0, x2), twiddle(p, s, j0, first + u));
            const Fr wc = twiddle(p, s - 1, j0, first + u);  // stage s - 1: both pairs
            y0 = reduce(norm(add_w(a0, a1)));
            y1 = mul(sub_w(a0, a1), wc);
            y2 = reduce(norm(add_w(a2, a3)));
            y3 = mul(sub_w(a2, a3), wc);
        }
        lds_st(lds, e0, y0);
        lds_st(lds, e0 + d, y1);
        lds_st(lds, e0 + 2 * d, y2);
        lds_st(lds, e0 + 3 * d, y3);
    }
    NTT_SYNC();
}

// One tile (logical block lb) of pass p: load, DIF rounds, un-bit-reversing
// store (with the optional scale), fully reduced.  On the device, when every
// thread owns one radix-2 butterfly per round, each round's twiddle is loaded
// during the round before (the first round's during the tile load), so its
// latency overlaps work.
template <bool R4>
HD void ntt_tile(uint32_t* __restrict__ lds, const NttPass& p, uint32_t lb) {
    const uint32_t S = p.S, logU = p.logU, U = 1u << logU;
    const uint32_t N = 1u << S, mask = N - 1, tile = N << logU;
    const uint32_t stride = p.L - S;  // log2 of the row pitch of columns / transposed outputs
    size_t base = 0;
    uint32_t g = lb;
    if (p.mode != PASS_ROWS) {
        const uint32_t tr = lb / p.tiles;
        g = lb - tr * p.tiles;
        base = size_t(tr) << p.L;
    }
    const uint32_t first = g << logU;  // first row, column or k1 of the tile
    int s = int(S) - 1;
    // Device, radix-2, one butterfly per thread and round: prefetched
    // twiddles.  (Radix-4 rounds keep loading their own: the three extra
    // twiddles in flight cost the radix-4 pass more occupancy than the
    // prefetch saves, measured on MI355X.)
#if defined(__HIP_DEVICE_COMPILE__)
    const bool one = !R4 && NTT_BDIM == (tile >> 1) && s >= 0;
#else
    const bool one = false;
#endif
    Fr w2;
    if (one) w2 = tw2_of(p, NTT_TID, s, first);

    for (uint32_t i = NTT_TID; i < tile; i += NTT_BDIM) {
        uint32_t u, pos;
        size_t gi;
        bool live = true;
        if (p.mode == PASS_COLS) {
            u = i & (U - 1);
            pos = i >> logU;
            gi = base + (size_t(pos) << stride) + first + u;
        } else {
            u = i >> S;
            pos = i & mask;
            gi = base + (size_t(first + u) << S) + pos;
            if (p.mode == PASS_ROWS) live = first + u < p.count;
        }
        Fr x = fp_zero<FrParams>();
        if (live) x = ld8(p.in + gi * 8);
        lds_st(lds, (pos << logU) + u, x);
    }
    NTT_SYNC();

    if (one) {
        for (; s >= 0; --s) {
            const Fr cur = w2;
            if (s >= 1) w2 = tw2_of(p, NTT_TID, s - 1, first);
            bfly2(lds, p, NTT_TID, s, cur);
            NTT_SYNC();
        }
    } else if (R4) {
        if (S & 1) round2(lds, p, tile, first, s--);
        for (; s >= 1; s -= 2) round4(lds, p, tile, first, s);
    } else {
        for (; s >= 0; --s) round2(lds, p, tile, first, s);
    }

    Fr scale;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) scale.l[k] = p.scale[k];
    for (uint32_t i = NTT_TID; i < tile; i += NTT_BDIM) {
        uint32_t u, k;
        size_t go;
        bool live = true;
        if (p.mode == PASS_ROWS) {
            u = i >> S;
            k = i & mask;
            go = (size_t(first + u) << S) + k;
            live = first + u < p.count;
        } else {
            u = i & (U - 1);
            k = i >> logU;
            go = base + (size_t(k) << stride) + first + u;
        }
        const uint32_t pos = S ? (NTT_BREV(k) >> (32 - S)) : 0u;
        Fr x = lds_ld(lds, (pos << logU) + u);
        if (p.scaled) x = mul(x, scale);
        if (live) st8(p.out + go * 8, canon(x));
    }
}

// ---- twiddle tables -----------------------------------------------------------

// Stage table entry e < 2^11 - 1: stage s = floor(log2(e + 1)), j = e + 1 - 2^s,
// omega_(2^(s+1))^j read from the power table pw (omega^k, k < 2^19).
HD void stage_table_entry(const uint32_t* __restrict__ pw, uint32_t* __restrict__ out, uint32_t e) {
    const uint32_t s = 31 - NTT_CLZ(e + 1), j = e + 1 - (1u << s);
    const Fr w = ld8(pw + (size_t(j) << (MAX_LOG - 1 - s)) * 8);
    uint32_t* o = out + ((1u << s) - 1) * LIMBS + j;
#pragma unroll
    for (int k = 0; k < LIMBS; ++k) o[k << s] = w.l[k];
}

// Pass-1 table entry e of length 2^L (S stages, a = L - S column bits,
// CT = 2^logCT columns per group): (group, s, j, column) holds
// omega^((j 2^a + x2) 2^(19 - a - s)).
HD void pass1_table_entry(const uint32_t* __restrict__ pw, uint32_t* __restrict__ out, uint32_t L, uint32_t S,
                          uint32_t logCT, uint32_t e) {
    const uint32_t a = L - S, ct = 1u << logCT;
    const uint32_t per_group = ct * ((1u << S) - 1);
    const uint32_t grp = e / per_group, r = e - grp * per_group;
    const uint32_t rr = r >> logCT, col = r & (ct - 1);  // rr = 2^s - 1 + j
    const uint32_t s = 31 - NTT_CLZ(rr + 1), j = rr + 1 - (1u << s);
    const uint32_t x2 = (grp << logCT) + col;
    const size_t k = ((size_t(j) << a) | x2) << (MAX_LOG - 1 - a - s);
    const Fr w = ld8(pw + k * 8);
    uint32_t* o = out + size_t(grp) * per_group * LIMBS + size_t(ct) * ((1u << s) - 1) * LIMBS + j * ct + col;
#pragma unroll
    for (int q = 0; q < LIMBS; ++q) o[size_t(q) * (ct << s)] = w.l[q];
}

inline uint32_t pass1_table_entries(int L) {
    const uint32_t S = uint32_t(L / 2), a = uint32_t(L) - S;
    return (1u << a) * ((1u << S) - 1);
}

// Columns per pass-1 table group for length 2^L.
inline uint32_t pass1_log_ct(int L) {
    const int S = L / 2, a = L - S;
    return uint32_t(std::min(a, TILE_LOG - S));
}

// ---- pass geometry ----------------------------------------------------------------

inline uint32_t pow2_floor(uint64_t x) {
    uint32_t p = 1;
    while (uint64_t(p) * 2 <= x && p < (1u << 30)) p *= 2;
    return p;
}

inline uint32_t log2_ceil(uint32_t x) {
    uint32_t l = 0;
    while ((1u << l) < x) ++l;
    return l;
}

// The passes of rows x 2^L (one or two) with their launch sizes; in -> out
// for one pass, in -> work -> out for two.  Twiddle pointers: the stage
// table and, for two passes, the length's pass-1 table; scale: the
// Montgomery limbs of 1/2^L (inverse) or null.  Tiles widen (more
// sub-transforms per tile, up to TILE elements) only beyond `target`
// tiles per launch.
struct LengthPlan {
    int passes = 0;
    NttPass pass[MAX_PASSES];
    uint32_t blocks[MAX_PASSES] = {0, 0};
    uint32_t tile[MAX_PASSES] = {0, 0};
};

inline LengthPlan plan_length(const uint32_t* in, uint32_t* out, uint32_t* work, int L, uint64_t rows,
                              const uint32_t* stage_tw, const uint32_t* pass1_tw, const Fr* scale,
                              uint32_t target = TARGET_BLOCKS) {
    LengthPlan lp;
    NttPass base{};
    base.L = uint32_t(L);
    base.count = uint32_t(rows);
    base.scaled = 0;
    for (int k = 0; k < LIMBS; ++k) base.scale[k] = 0;
    if (L <= SINGLE_MAX_LOG) {
        NttPass p = base;
        p.in = in;
        p.out = out;
        p.tw = stage_tw;
        p.mode = PASS_ROWS;
        p.S = uint32_t(L);
        const uint32_t u = std::min<uint32_t>(pow2_floor(std::max<uint64_t>(1, rows / target)), TILE >> L);
        p.logU = log2_ceil(u);
        p.tiles = uint32_t((rows + u - 1) / u);
        p.logCT = 0;
        if (scale && L) {
            p.scaled = 1;
            for (int k = 0; k < LIMBS; ++k) p.scale[k] = scale->l[k];
        }
        lp.passes = 1;
        lp.pass[0] = p;
        lp.blocks[0] = p.tiles;
        lp.tile[0] = u << L;
        return lp;
    }
    const uint32_t S1 = uint32_t(L / 2), a = uint32_t(L) - S1;
    NttPass p1 = base, p2 = base;
    p1.in = in;
    p1.out = work;
    p1.tw = pass1_tw;
    p1.mode = PASS_COLS;
    p1.S = S1;
    p1.logCT = pass1_log_ct(L);
    const uint32_t c = std::min<uint32_t>(pow2_floor(std::max<uint64_t>(1, (rows << a) / target)), 1u << p1.logCT);
    p1.logU = log2_ceil(c);
    p1.tiles = (1u << a) / c;
    p2.in = work;
    p2.out = out;
    p2.tw = stage_tw;
    p2.mode = PASS_XPOSE;
    p2.S = a;
    p2.logCT = 0;
    const uint32_t r = std::min<uint32_t>(pow2_floor(std::max<uint64_t>(1, (rows << S1) / target)),
                                          std::min<uint32_t>(1u << S1, TILE >> a));
    p2.logU = log2_ceil(r);
    p2.tiles = (1u << S1) / r;
    if (scale) {
        p2.scaled = 1;
        for (int k = 0; k < LIMBS; ++k) p2.scale[k] = scale->l[k];
    }
    lp.passes = 2;
    lp.pass[0] = p1;
    lp.pass[1] = p2;
    lp.blocks[0] = uint32_t(rows * p1.tiles);
    lp.blocks[1] = uint32_t(rows * p2.tiles);
    lp.tile[0] = c << S1;
    lp.tile[1] = r << a;
    return lp;
}

inline bool use_radix4(uint64_t blocks, uint32_t tile) { return blocks * (tile >> 2) >= R4_MIN_GROUPS; }

inline unsigned threads_for(uint32_t tile, bool r4) {
    return unsigned(std::max<uint32_t>(64, r4 ? tile >> 2 : tile >> 1));
}

}  // namespace ntt
}  // namespace dvgpu
