// Device memory forms of Fr and G1 shared by the resident engine and the
// opening-cache precompute: Fr and Fq as eight little-endian 32-bit words in
// Montgomery form, G1 as XYZZ (X, Y, ZZ, ZZZ: 32 words, 128 bytes).
#pragma once
#include "common.hpp"
#include "curve.hpp"
#include "glv.hpp"

namespace dvgpu {

using namespace bn254;
using G1 = Xyzz<FqOps>;

__device__ __forceinline__ Fr ldfr(const uint32_t* p) {
    const uint4 a = reinterpret_cast<const uint4*>(p)[0];
    const uint4 b = reinterpret_cast<const uint4*>(p)[1];
    const uint32_t w[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    return unpack<FrParams>(w);
}
__device__ __forceinline__ void stfr(uint32_t* p, const Fr& v) {
    uint32_t w[8];
    pack(v, w);
    reinterpret_cast<uint4*>(p)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(p)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}
__device__ __forceinline__ Fq ldfq(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return unpack<FqParams>(w);
}
__device__ __forceinline__ void stfq(uint32_t* p, const Fq& v) {
    uint32_t w[8];
    pack(v, w);
    for (int i = 0; i < 8; ++i) p[i] = w[i];
}
__device__ __forceinline__ G1 ldg1(const uint32_t* p) { return {ldfq(p), ldfq(p + 8), ldfq(p + 16), ldfq(p + 24)}; }
__device__ __forceinline__ void stg1(uint32_t* p, const G1& v) {
    stfq(p, v.X);
    stfq(p + 8, v.Y);
    stfq(p + 16, v.ZZ);
    stfq(p + 24, v.ZZZ);
}
// canonical scalar words of a Montgomery Fr value
__device__ __forceinline__ void fr_bits(const Fr& v, uint32_t (&w)[8]) { from_mont(v, w); }

// [s]P for a per-lane scalar (GLV split, joint double-and-add).
__device__ inline G1 g1_scale(const G1& p, const Fr& s) {
    uint32_t k[8];
    fr_bits(s, k);
    return g1_mul_scalar(p, k, false);
}

// Fr root^e for a 64-bit exponent (Montgomery in, Montgomery out).
__device__ inline Fr fr_pow64(const Fr& root, uint64_t e) {
    Fr r = fp_one<FrParams>(), b = root;
    while (e) {
        if (e & 1) r = mul(r, b);
        e >>= 1;
        if (e) b = sqr(b);
    }
    return r;
}

}  // namespace dvgpu
