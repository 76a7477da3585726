// Shared host-side plumbing: HIP error checks, a grow-only device arena and
// the exception types that map onto the C ABI's status codes.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>
#include <stdexcept>
#include <string>
#include <vector>

#include "errors.hpp"

namespace dvgpu {

inline void hip_check(hipError_t status, const char* what) {
    if (status != hipSuccess)
        throw DeviceError(std::string(what) + ": " + hipGetErrorString(status));
}
#define HIP_CHECK(expr) ::dvgpu::hip_check((expr), #expr)
#define HIP_CHECK_LAUNCH() ::dvgpu::hip_check(hipGetLastError(), "kernel launch")


// Memory accounting shared by every device allocation an owner makes.  The
// limit models the reference's per-owner quota; 0 means unlimited.
struct Quota {
    uint64_t limit = 0;
    uint64_t used = 0;
    uint64_t peak = 0;
    void charge(uint64_t bytes) {
        if (limit && (used > limit || bytes > limit - used))
            throw InvalidArgument("device memory quota exceeded");
        used += bytes;
        if (used > peak) peak = used;
    }
    void release(uint64_t bytes) { used = bytes > used ? 0 : used - bytes; }
};

// A named, grow-only device buffer.  Contents are not preserved on growth.
// A borrowed buffer views memory it does not own (a slice of a fixed scratch
// area): it never allocates or frees, and refuses to grow.
struct DeviceBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    Quota* quota = nullptr;
    bool borrowed = false;
    // Changes whenever the buffer starts to view other memory (allocation,
    // release, borrow), so a cached record of its contents can be checked.
    uint64_t generation = 0;
    DeviceBuffer() = default;
    DeviceBuffer(const DeviceBuffer&) = delete;
    DeviceBuffer& operator=(const DeviceBuffer&) = delete;
    ~DeviceBuffer() { release(); }
    void release() noexcept {
        if (ptr && !borrowed) {
            (void)hipFree(ptr);
            if (quota) quota->release(bytes);
        }
        ptr = nullptr;
        bytes = 0;
        borrowed = false;
        ++generation;
    }
    void borrow(void* p, size_t n) {
        release();
        ptr = p;
        bytes = n;
        borrowed = true;
    }
    void* ensure(size_t need) {
        if (need <= bytes) return ptr;
        if (borrowed) throw InvalidArgument("borrowed device buffer too small");
        const size_t grow = grown(need);
        if (quota) quota->charge(grow);
        void* fresh = nullptr;
        const hipError_t st = hipMalloc(&fresh, grow);
        if (st != hipSuccess) {
            if (quota) quota->release(grow);
            throw DeviceError(std::string("hipMalloc: ") + hipGetErrorString(st));
        }
        release();
        ptr = fresh;
        bytes = grow;
        return ptr;
    }
    template <class T>
    T* as(size_t count) { return static_cast<T*>(ensure(count * sizeof(T))); }
    // Bytes an ensure(need) on an empty buffer allocates (25% headroom,
    // 256-byte granules).
    static size_t grown(size_t need) { return ((need + need / 4) + 255) & ~size_t(255); }
};

// Pinned host staging buffer (hipHostMalloc), grow-only.
struct HostBuffer {
    void* ptr = nullptr;
    size_t bytes = 0;
    HostBuffer() = default;
    HostBuffer(const HostBuffer&) = delete;
    HostBuffer& operator=(const HostBuffer&) = delete;
    ~HostBuffer() { release(); }
    void release() noexcept {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        bytes = 0;
    }
    void* ensure(size_t need) {
        if (need <= bytes) return ptr;
        size_t grow = (need + need / 4 + 4095) & ~size_t(4095);
        void* fresh = nullptr;
        HIP_CHECK(hipHostMalloc(&fresh, grow, hipHostMallocDefault));
        release();
        ptr = fresh;
        bytes = grow;
        return ptr;
    }
    template <class T>
    T* as(size_t count) { return static_cast<T*>(ensure(count * sizeof(T))); }
};

// Create an event / a non-blocking stream; `out` is written only on success,
// so an owner's destructor never releases a handle that was not made.
inline void create_event(hipEvent_t& out, unsigned flags = hipEventDefault) {
    hipEvent_t e = nullptr;
    hip_check(hipEventCreateWithFlags(&e, flags), "hipEventCreateWithFlags");
    out = e;
}
inline void create_stream(hipStream_t& out) {
    hipStream_t s = nullptr;
    hip_check(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "hipStreamCreateWithFlags");
    out = s;
}

// Events owned by one call: created together (none when `on` is false) and
// destroyed on every exit path, a failed creation included.
template <int N>
struct ScopedEvents {
    hipEvent_t e[N] = {};
    explicit ScopedEvents(bool on = true) {
        if (!on) return;
        for (auto& x : e) {
            hipEvent_t made = nullptr;
            const hipError_t st = hipEventCreate(&made);
            if (st != hipSuccess) {
                destroy();
                hip_check(st, "hipEventCreate");
            }
            x = made;
        }
    }
    ~ScopedEvents() { destroy(); }
    ScopedEvents(const ScopedEvents&) = delete;
    ScopedEvents& operator=(const ScopedEvents&) = delete;
    hipEvent_t operator[](int i) const { return e[i]; }
    float ms(int a, int b) const {
        float v = 0;
        hip_check(hipEventElapsedTime(&v, e[a], e[b]), "hipEventElapsedTime");
        return v;
    }
    uint64_t ns(int a, int b) const { return uint64_t(double(ms(a, b)) * 1e6); }

  private:
    void destroy() noexcept {
        for (auto& x : e)
            if (x) {
                (void)hipEventDestroy(x);
                x = nullptr;
            }
    }
};

inline unsigned grid_for(size_t n, unsigned block) { return unsigned((n + block - 1) / block); }

}  // namespace dvgpu
