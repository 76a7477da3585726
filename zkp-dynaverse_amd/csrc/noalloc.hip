// Allocation-free batched MSM: the dvgpu_ext_msm_batch_finalize computation
// with every intermediate buffer (digits, sorted entries, tasks, partial
// sums, chunk sums, descriptors, window values) carved out of one
// file-scope device array instead of the owner's growing workspace.
//
// The array's address differs per device and is looked up once per device
// with hipGetSymbolAddress while that device is current; host code uses only
// that address (the array's own name is a host-side stand-in).  Every call on
// one device shares the array and finds the previous call's contents in it;
// the enqueue clears what must start at zero (counters, histograms, the
// error word) with fill kernels, as it does for the regular workspace.
// Calls that overlap on one device (two threads, two owners) race on the
// array: callers serialise them.
#include <cstring>
#include <mutex>
#include <vector>

#include "../../include/dynaverse_hip_ext.h"
#include "owner.hpp"

namespace dvgpu {
namespace {

constexpr size_t SCRATCH_BYTES = size_t(DVGPU_NOALLOC_SCRATCH_BYTES);
static_assert(SCRATCH_BYTES % 256 == 0, "scratch slices are 256-byte aligned");

static __device__ __align__(256) uint8_t scratch[SCRATCH_BYTES];

// Owners run on device 0 (dvgpu_backend_create selects it).
constexpr int OWNER_DEVICE = 0;
constexpr int MAX_DEVICES = 64;

std::mutex scratch_lock;
uint8_t* scratch_ptr[MAX_DEVICES] = {};

// Device address of `scratch` on `device`, looked up on first use with that
// device current (the caller's current device is restored afterwards).
uint8_t* scratch_on(int device) {
    require(device >= 0 && device < MAX_DEVICES, "device ordinal out of range");
    std::lock_guard<std::mutex> g(scratch_lock);
    if (scratch_ptr[device]) return scratch_ptr[device];
    int current = 0;
    HIP_CHECK(hipGetDevice(&current));
    if (current != device) HIP_CHECK(hipSetDevice(device));
    void* p = nullptr;
    const hipError_t st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (current != device) HIP_CHECK(hipSetDevice(current));
    if (st != hipSuccess) throw DeviceError(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(st));
    require(p != nullptr, "scratch array has no device address");
    scratch_ptr[device] = static_cast<uint8_t*>(p);
    return scratch_ptr[device];
}

template <class F>
void enqueue(const void* points, const void* point_index, const void* scalars, const MsmBatchPlan& plan,
             MsmWorkspace& ws, size_t count, uint8_t* out, hipStream_t st) {
    constexpr int PB = PointLayout<F>::CANON_BYTES;
    uint32_t* win = static_cast<uint32_t*>(ws.windows.ptr);
    msm_batch_enqueue<F>(static_cast<const uint32_t*>(points), static_cast<const uint32_t*>(point_index),
                         static_cast<const uint32_t*>(scalars), plan, ws, st, win);
    device_fill_words(reinterpret_cast<uint32_t*>(out), count * PB / 4, 0u, st);
    msm_batch_finalize_enqueue<F>(win, plan, ws, out, nullptr, st);
}

}  // namespace
}  // namespace dvgpu

using namespace dvgpu;

extern "C" int dvgpu_ext_msm_batch_finalize_noalloc(void* owner, const void* points, size_t point_count,
                                                    const void* point_index, size_t index_count,
                                                    const void* scalars, size_t scalar_count,
                                                    const DvgpuMsmItem* items, size_t count, int g2,
                                                    void* results_device, uint8_t* results_host, char* error,
                                                    size_t capacity) {
    Owner* o = static_cast<Owner*>(owner);
    return guarded(o, error, capacity, [&]() -> int {
        require(items && count && (g2 == 0 || g2 == 1) && points && point_count && scalars && results_device,
                "invalid batch arguments");
        require(point_count <= 0xffffffffull, "too many points");
        for (size_t i = 0; i < count; ++i) {
            const DvgpuMsmItem& it = items[i];
            require(uint64_t(it.scalar_base) + it.n <= scalar_count, "MSM scalars out of range");
            if (point_index) require(uint64_t(it.point_base) + it.n <= index_count, "MSM point indices out of range");
            else require(uint64_t(it.point_base) + it.n <= point_count, "MSM points out of range");
        }
        std::vector<size_t> map;
        MsmBatchPlan plan = plan_items(items, count, o, map, g2 != 0);
        plan.point_limit = uint32_t(point_count);
        const size_t pb = g2 ? 128 : 64;
        uint8_t* out = static_cast<uint8_t*>(results_device);
        uint32_t err = 0;
        if (plan.descs.empty()) {
            device_fill_words(reinterpret_cast<uint32_t*>(out), count * pb / 4, 0u, o->stream);
        } else {
            const int xw = g2 ? PointLayout<Fq2Ops>::XYZZ_WORDS : PointLayout<FqOps>::XYZZ_WORDS;
            MsmWorkspace ws;  // every buffer borrowed from the scratch array
            msm_workspace_borrow(plan, xw, scratch_on(OWNER_DEVICE), SCRATCH_BYTES, ws);
            HIP_CHECK(hipMemcpyAsync(ws.descs.ptr, plan.descs.data(), plan.descs.size() * sizeof(MsmDesc),
                                     hipMemcpyHostToDevice, o->stream));
            if (g2) enqueue<Fq2Ops>(points, point_index, scalars, plan, ws, count, out, o->stream);
            else enqueue<FqOps>(points, point_index, scalars, plan, ws, count, out, o->stream);
            HIP_CHECK(hipMemcpyAsync(&err, static_cast<uint32_t*>(ws.total.ptr) + 3, 4, hipMemcpyDeviceToHost,
                                     o->stream));
        }
        if (results_host) HIP_CHECK(hipMemcpyAsync(results_host, out, count * pb, hipMemcpyDeviceToHost, o->stream));
        HIP_CHECK(hipStreamSynchronize(o->stream));
        require(err == 0, "MSM point index out of range");
        return 0;
    });
}
