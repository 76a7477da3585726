// Batched pairing-product checks (include/dynaverse_pairing.h): one thread per
// group of pairs; Miller loops multiplied into one Fq12, one final
// exponentiation, the product compared with one.
#include <cstring>

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

__global__ void __launch_bounds__(64) k_pairing_check(const uint32_t* __restrict__ g1, const uint32_t* __restrict__ g2,
                                                      uint32_t terms, uint64_t count, uint8_t* __restrict__ ok,
                                                      uint32_t* __restrict__ gt) {
    const uint64_t g = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= count) return;
    PairingInput in[PAIRING_MAX_TERMS];
    int k = 0;
    for (uint32_t j = 0; j < terms; ++j) {
        const uint32_t* p = g1 + (g * terms + j) * 16;
        const uint32_t* q = g2 + (g * terms + j) * 32;
        if (all_zero(p, 16) || all_zero(q, 32)) continue;  // e(O, Q) = e(P, O) = 1
        in[k++] = {ld_canon_fq(p), ld_canon_fq(p + 8), {ld_canon_fq(q), ld_canon_fq(q + 8)},
                   {ld_canon_fq(q + 16), ld_canon_fq(q + 24)}};
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

bool host_all_zero(const uint8_t* p, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (p[i]) return false;
    return true;
}

void validate_points(const uint8_t* g1, const uint8_t* g2, size_t n) {
    for (size_t i = 0; i < n; ++i) {
        const uint8_t* p = g1 + 64 * i;
        if (!host_all_zero(p, 64)) {
            uint32_t x[8], y[8];
            std::memcpy(x, p, 32);
            std::memcpy(y, p + 32, 32);
            require(is_canonical<FqParams>(x) && is_canonical<FqParams>(y), "G1 coordinate not canonical");
            const Affine<FqOps> a = {to_mont<FqParams>(x), to_mont<FqParams>(y), false};
            require(affine_on_curve<FqOps>(a), "G1 point is not on the curve");
        }
        const uint8_t* q = g2 + 128 * i;
        if (!host_all_zero(q, 128)) {
            uint32_t w[4][8];
            for (int k = 0; k < 4; ++k) {
                std::memcpy(w[k], q + 32 * k, 32);
                require(is_canonical<FqParams>(w[k]), "G2 coordinate not canonical");
            }
            const Affine<Fq2Ops> a = {{to_mont<FqParams>(w[0]), to_mont<FqParams>(w[1])},
                                      {to_mont<FqParams>(w[2]), to_mont<FqParams>(w[3])}, false};
            require(affine_on_curve<Fq2Ops>(a), "G2 point is not on the twist");
        }
    }
}

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
        validate_points(g1, g2, n);
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
        return 0;
    });
}
