// Device-resident extension entry points (dynaverse_hip_ext.h).
#include <algorithm>
#include <cstring>

#include "../../include/dynaverse_hip_ext.h"
#include "owner.hpp"

using namespace dvgpu;

namespace {
Owner* as_owner(void* p) { return static_cast<Owner*>(p); }

__global__ void k_copy16(const uint4* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    for (size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += size_t(gridDim.x) * blockDim.x)
        dst[i] = src[i];
}

using Events = ScopedEvents<MSM_PHASE_EVENTS>;

template <class F>
int msm_device(Owner* o, const void* points, const void* scalars, size_t n, int window_bits,
               uint8_t* result, float* phase_ms) {
    constexpr int XW = PointLayout<F>::XYZZ_WORDS;
    MsmConfig cfg = o->msm_config;
    cfg.window_bits = window_bits ? window_bits : msm_window_for(n, &o->window_table);
    hipEvent_t* ev = phase_ms ? o->msm_phase_events() : nullptr;
    const MsmPlan p0 = msm_plan(n, cfg);
    uint32_t* window_words = o->msm.windows.as<uint32_t>(p0.outputs * XW);
    o->msm.phase_events = ev;
    MsmPlan plan;
    try {
        plan = msm_enqueue<F>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(scalars),
                              n, o->msm, o->stream, cfg, window_words, p0.outputs);
    } catch (...) {
        o->msm.phase_events = nullptr;
        throw;
    }
    o->msm.phase_events = nullptr;
    uint32_t* host = o->msm.host_windows.as<uint32_t>(plan.outputs * XW);
    HIP_CHECK(hipMemcpyAsync(host, window_words, plan.outputs * XW * 4, hipMemcpyDeviceToHost, o->stream));
    HIP_CHECK(hipStreamSynchronize(o->stream));
    if (phase_ms)
        for (int k = 0; k + 1 < MSM_PHASE_EVENTS; ++k)
            HIP_CHECK(hipEventElapsedTime(&phase_ms[k], ev[k], ev[k + 1]));
    affine_to_bytes<F>(msm_combine_windows<F>(host, plan), result);
    return 0;
}
}  // namespace

extern "C" {

int dvgpu_ext_alloc(void* owner, size_t bytes, void** ptr, char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(ptr && bytes, "invalid allocation request");
        o->quota.charge(bytes);
        void* p = nullptr;
        const hipError_t st = hipMalloc(&p, bytes);
        if (st != hipSuccess) {
            o->quota.release(bytes);
            throw DeviceError(std::string("hipMalloc: ") + hipGetErrorString(st));
        }
        o->ext_allocations[p] = bytes;
        *ptr = p;
        return 0;
    });
}

int dvgpu_ext_free(void* owner, void* ptr) {
    Owner* o = as_owner(owner);
    if (!o) return -1;
    if (o->retired) return -4;
    auto it = o->ext_allocations.find(ptr);
    if (it == o->ext_allocations.end()) return -1;
    (void)hipStreamSynchronize(o->stream);
    (void)hipFree(it->first);
    o->quota.release(it->second);
    o->ext_allocations.erase(it);
    return 0;
}

int dvgpu_ext_copy_to_device(void* owner, void* dst, const void* src, size_t bytes, char* error,
                             size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require((dst && src) || !bytes, "null copy buffers");
        if (bytes) {
            HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, o->stream));
            HIP_CHECK(hipStreamSynchronize(o->stream));
        }
        return 0;
    });
}

int dvgpu_ext_copy_to_host(void* owner, void* dst, const void* src, size_t bytes, char* error,
                           size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require((dst && src) || !bytes, "null copy buffers");
        if (bytes) {
            HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, o->stream));
            HIP_CHECK(hipStreamSynchronize(o->stream));
        }
        return 0;
    });
}

int dvgpu_ext_device_copy(void* owner, void* dst, const void* src, size_t bytes, char* error,
                          size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(dst && src && bytes % 16 == 0, "copy needs device buffers and a multiple of 16 bytes");
        const size_t n = bytes / 16;
        if (n) {
            k_copy16<<<unsigned(std::min<size_t>(grid_for(n, 256), 8192)), 256, 0, o->stream>>>(
                static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
            HIP_CHECK_LAUNCH();
            HIP_CHECK(hipStreamSynchronize(o->stream));
        }
        return 0;
    });
}

int dvgpu_ext_stream(void* owner, void** stream) {
    Owner* o = as_owner(owner);
    if (!o || o->retired) return -4;
    if (!stream) return -1;
    *stream = o->stream;
    return 0;
}

int dvgpu_ext_points_to_mont(void* owner, const void* canonical, void* mont, size_t n, int g2,
                             char* error, size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require((canonical && mont) || !n, "null device buffers");
        require(g2 == 0 || g2 == 1, "g2 must be 0 or 1");
        if (g2)
            points_to_mont<Fq2Ops>(static_cast<const uint8_t*>(canonical), static_cast<uint32_t*>(mont), n, o->stream);
        else
            points_to_mont<FqOps>(static_cast<const uint8_t*>(canonical), static_cast<uint32_t*>(mont), n, o->stream);
        HIP_CHECK(hipStreamSynchronize(o->stream));
        return 0;
    });
}

int dvgpu_ext_msm_device(void* owner, const void* points, const void* scalars, size_t n, int g2,
                         int window_bits, uint8_t* result, float* phase_ms, char* error,
                         size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(result && n <= (size_t(1) << 24) && (!n || (points && scalars)), "invalid MSM arguments");
        require(g2 == 0 || g2 == 1, "g2 must be 0 or 1");
        if (!n) {
            std::memset(result, 0, g2 ? 128 : 64);
            return 0;
        }
        return g2 ? msm_device<Fq2Ops>(o, points, scalars, n, window_bits, result, phase_ms)
                  : msm_device<FqOps>(o, points, scalars, n, window_bits, result, phase_ms);
    });
}

int dvgpu_ext_ntt_device(void* owner, const void* in, void* out, size_t length, size_t rows,
                         int inverse, const uint8_t* root, float* kernel_ms, char* error,
                         size_t capacity) {
    Owner* o = as_owner(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(in && out && root && length && length <= DVGPU_MAX_NTT_SIZE && !(length & (length - 1)) &&
                    rows && (inverse == 0 || inverse == 1),
                "invalid NTT arguments");
        uint32_t rw[8];
        std::memcpy(rw, root, 32);
        require(ntt_root_has_order_2_20(rw), "NTT root must have order 2^20");
        ntt_domain_prepare(o->domain, rw, o->stream);
        int log_n = 0;
        while ((size_t(1) << log_n) < length) ++log_n;
        Events ev(kernel_ms != nullptr);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev.e[0], o->stream));
        ntt_rows(o->domain, static_cast<const uint32_t*>(in), static_cast<uint32_t*>(out), log_n, rows,
                 inverse != 0, o->ntt, o->stream);
        if (kernel_ms) HIP_CHECK(hipEventRecord(ev.e[1], o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        if (kernel_ms) HIP_CHECK(hipEventElapsedTime(kernel_ms, ev.e[0], ev.e[1]));
        return 0;
    });
}

}  // extern "C"
