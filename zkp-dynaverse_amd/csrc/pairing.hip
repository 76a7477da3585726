// Batched pairing-product checks (include/dynaverse_pairing.h): one thread per
// group of pairs; Miller loops multiplied into one Fq12, one final
// exponentiation, the product compared with one.
#include <cstring>
#include <string>

#include "../../include/dynaverse_pairing.h"
#include "owner.hpp"
#include "pairing.hpp"

namespace dvgpu {
namespace {

using namespace bn254;

__device__ __forceinline__ Fq ld_canon_fq(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return to_mont<FqParams>(w);
}

__device__ __forceinline__ void st_canon_fq(uint32_t* p, const Fq& v) {
    uint32_t w[8];
    from_mont(v, w);
    for (int i = 0; i < 8; ++i) p[i] = w[i];
}

__device__ __forceinline__ bool all_zero(const uint32_t* p, int n) {
    uint32_t any = 0;
    for (int i = 0; i < n; ++i) any |= p[i];
    return !any;
}

__device__ __forceinline__ bool canonical_fq(const uint32_t* p) {
    uint32_t w[8];
    for (int i = 0; i < 8; ++i) w[i] = p[i];
    return is_canonical<FqParams>(w);
}

__global__ void __launch_bounds__(64) k_pairing_check(const uint32_t* __restrict__ g1, const uint32_t* __restrict__ g2,
                                                      uint32_t terms, uint64_t count, uint8_t* __restrict__ ok,
                                                      uint32_t* __restrict__ gt) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    PairingInput in[PAIRING_MAX_TERMS];
    int k = 0;
    bool valid = true;
    for (uint32_t j = 0; j < terms; ++j) {
        const uint32_t* p = g1 + (g * terms + j) * 16;
        const uint32_t* q = g2 + (g * terms + j) * 32;
        const bool p0 = all_zero(p, 16), q0 = all_zero(q, 32);
        bool pv = p0, qv = q0;
        PairingInput e;
        if (!p0 && canonical_fq(p) && canonical_fq(p + 8)) {
            e.xp = ld_canon_fq(p);
            e.yp = ld_canon_fq(p + 8);
            pv = affine_on_curve<FqOps>({e.xp, e.yp, false});
        }
        if (!q0 && canonical_fq(q) && canonical_fq(q + 8) && canonical_fq(q + 16) && canonical_fq(q + 24)) {
            e.x2 = {ld_canon_fq(q), ld_canon_fq(q + 8)};
            e.y2 = {ld_canon_fq(q + 16), ld_canon_fq(q + 24)};
            qv = affine_on_curve<Fq2Ops>({e.x2, e.y2, false});
        }
        valid = valid && pv && qv;
        if (!p0 && !q0) in[k++] = e;  // e(O, Q) = e(P, O) = 1
    }
    if (!valid) {  // reported to the host, which fails the call
        ok[g] = 2;
        return;
    }
    const Fq12 r = final_exponentiation(k ? miller_product(in, k) : f12_one());
    ok[g] = f12_is_one(r) ? 1 : 0;
    if (gt) {
        uint32_t* o = gt + g * 96;
        const Fq6* c[2] = {&r.c0, &r.c1};
        for (int k = 0; k < 2; ++k) {
            const Fq2* a[3] = {&c[k]->a0, &c[k]->a1, &c[k]->a2};
            for (int i = 0; i < 3; ++i) {
                st_canon_fq(o + (k * 3 + i) * 16, a[i]->c0);
                st_canon_fq(o + (k * 3 + i) * 16 + 8, a[i]->c1);
            }
        }
    }
}

Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" int dvgpu_bn254_pairing_check(void* backend, const uint8_t* g1, const uint8_t* g2, size_t terms,
                                         size_t count, uint8_t* results, uint8_t* gt, float* kernel_ms,
                                         char* error, size_t capacity) {
    Owner* o = as_owner(backend);
    return guarded(o, error, capacity, [&]() -> int {
        require(terms >= 1 && terms <= PAIRING_MAX_TERMS, "terms must be 1..16");
        require(count <= (size_t(1) << 24), "too many pairing checks");
        if (!count) return 0;
        require(g1 && g2 && results, "null pairing buffers");
        const size_t n = terms * count;
        hipStream_t st = o->stream;
        // one device allocation: g1 | g2 | results | gt
        const size_t b1 = n * 64, b2 = n * 128, br = (count + 255) & ~size_t(255), bg = gt ? count * 384 : 0;
        auto* base = static_cast<uint8_t*>(o->stage_a.ensure(b1 + b2 + br + bg));
        HIP_CHECK(hipMemcpyAsync(base, g1, b1, hipMemcpyHostToDevice, st));
        HIP_CHECK(hipMemcpyAsync(base + b1, g2, b2, hipMemcpyHostToDevice, st));
        hipEvent_t ev[2] = {nullptr, nullptr};
        if (kernel_ms) {
            for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
            HIP_CHECK(hipEventRecord(ev[0], st));
        }
        k_pairing_check<<<grid_for(count, 64), 64, 0, st>>>(
            reinterpret_cast<const uint32_t*>(base), reinterpret_cast<const uint32_t*>(base + b1), uint32_t(terms),
            count, base + b1 + b2, gt ? reinterpret_cast<uint32_t*>(base + b1 + b2 + br) : nullptr);
        HIP_CHECK_LAUNCH();
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev[1], st));
        HIP_CHECK(hipMemcpyAsync(results, base + b1 + b2, count, hipMemcpyDeviceToHost, st));
        if (gt) HIP_CHECK(hipMemcpyAsync(gt, base + b1 + b2 + br, bg, hipMemcpyDeviceToHost, st));
        HIP_CHECK(hipStreamSynchronize(st));
        if (kernel_ms) {
            HIP_CHECK(hipEventElapsedTime(kernel_ms, ev[0], ev[1]));
            for (auto& e : ev) (void)hipEventDestroy(e);
        }
        const void* bad = std::memchr(results, 2, count);
        if (bad) {
            const size_t i = static_cast<const uint8_t*>(bad) - results;
            std::memset(results, 0, count);
            if (gt) std::memset(gt, 0, count * 384);
            throw InvalidArgument(("pairing check " + std::to_string(i) +
                                   ": a point is not canonical or not on its curve").c_str());
        }
        return 0;
    });
}
