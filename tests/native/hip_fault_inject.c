/* Test-only LD_PRELOAD library: makes the k-th call of one HIP allocation
 * family fail, on the host, without calling the runtime (so no device work
 * is involved), and keeps per-family balances of live handles.
 *
 * Families: 0 device memory (hipMalloc / hipFree), 1 events
 * (hipEventCreate, hipEventCreateWithFlags / hipEventDestroy), 2 streams
 * (hipStreamCreate, hipStreamCreateWithFlags / hipStreamDestroy), 3 pinned
 * host memory (hipHostMalloc / hipHostFree).  The test process arms a
 * family through dvgpu_inject_arm() (found with ctypes in the preloaded
 * library) and reads the balances through dvgpu_inject_live().
 *
 * Reference counterpart: native/tests/fail_cuda_alloc.cpp (injected CUDA
 * allocation, event and stream failures for crates/gpu/tests/icicle_backend.rs).
 */
#define _GNU_SOURCE
#include <dlfcn.h>
#include <link.h>
#include <pthread.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

typedef int hipError_t; /* enum hipError_t: 0 = hipSuccess */
typedef void* hipEvent_t;
typedef void* hipStream_t;
enum { HIP_SUCCESS = 0, HIP_ERROR_OUT_OF_MEMORY = 2, HIP_ERROR_NOT_INITIALIZED = 3 };
enum { FAMILY_MEMORY, FAMILY_EVENT, FAMILY_STREAM, FAMILY_PINNED, FAMILIES };

static pthread_mutex_t lock = PTHREAD_MUTEX_INITIALIZER;
static int armed_family = -1;
static unsigned armed_countdown = 0;
static unsigned fired = 0;
static long live[FAMILIES];

/* Device and pinned pointers handed out, so frees of foreign pointers are
 * not counted (open addressing; a full table stops tracking, never fails). */
#define SLOTS 65536
static uintptr_t tracked[SLOTS];

static size_t slot_hash(uintptr_t p) { return (size_t)((p >> 8) * 0x9E3779B97F4A7C15ull) & (SLOTS - 1); }

static void track(uintptr_t p) {
    for (size_t i = slot_hash(p), k = 0; k < SLOTS; ++k, i = (i + 1) & (SLOTS - 1))
        if (tracked[i] == 0 || tracked[i] == 1) {
            tracked[i] = p;
            return;
        }
}

static int untrack(uintptr_t p) {
    for (size_t i = slot_hash(p), k = 0; k < SLOTS && tracked[i]; ++k, i = (i + 1) & (SLOTS - 1))
        if (tracked[i] == p) {
            tracked[i] = 1; /* tombstone */
            return 1;
        }
    return 0;
}

/* 1 when this call of `family` is the armed one (the arming is then spent). */
static int should_fail(int family) {
    int fail = 0;
    pthread_mutex_lock(&lock);
    if (armed_family == family && armed_countdown && --armed_countdown == 0) {
        armed_family = -1;
        ++fired;
        fail = 1;
    }
    pthread_mutex_unlock(&lock);
    return fail;
}

static void count(int family, long delta, uintptr_t p, int track_pointer) {
    pthread_mutex_lock(&lock);
    if (track_pointer) {
        if (delta > 0) track(p);
        else if (!untrack(p)) delta = 0;
    }
    live[family] += delta;
    pthread_mutex_unlock(&lock);
}

/* The runtime's own definition.  The HIP runtime usually arrives as a
 * dependency of a library loaded with RTLD_LOCAL (Python's ctypes), outside
 * the global scope RTLD_NEXT searches: then look it up in the loaded
 * libamdhip64 itself. */
static int find_runtime(struct dl_phdr_info* info, size_t size, void* data) {
    (void)size;
    if (info->dlpi_name && strstr(info->dlpi_name, "libamdhip64")) {
        *(const char**)data = info->dlpi_name;
        return 1;
    }
    return 0;
}

static void* next(const char* name) {
    void* f = dlsym(RTLD_NEXT, name);
    if (f) return f;
    static void* runtime;
    if (!runtime) {
        const char* path = NULL;
        dl_iterate_phdr(find_runtime, &path);
        if (path) runtime = dlopen(path, RTLD_NOW | RTLD_NOLOAD);
    }
    return runtime ? dlsym(runtime, name) : NULL;
}

void dvgpu_inject_arm(int family, unsigned k) {
    pthread_mutex_lock(&lock);
    armed_family = k ? family : -1;
    armed_countdown = k;
    pthread_mutex_unlock(&lock);
}

unsigned dvgpu_inject_fired(void) { return fired; }
long dvgpu_inject_live(int family) { return family >= 0 && family < FAMILIES ? live[family] : 0; }

hipError_t hipMalloc(void** ptr, size_t size) {
    static hipError_t (*real)(void**, size_t);
    if (!real) real = (hipError_t(*)(void**, size_t))next("hipMalloc");
    if (should_fail(FAMILY_MEMORY)) {
        if (ptr) *ptr = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(ptr, size);
    if (st == HIP_SUCCESS && ptr && *ptr) count(FAMILY_MEMORY, 1, (uintptr_t)*ptr, 1);
    return st;
}

hipError_t hipFree(void* ptr) {
    static hipError_t (*real)(void*);
    if (!real) real = (hipError_t(*)(void*))next("hipFree");
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(ptr);
    if (st == HIP_SUCCESS && ptr) count(FAMILY_MEMORY, -1, (uintptr_t)ptr, 1);
    return st;
}

hipError_t hipHostMalloc(void** ptr, size_t size, unsigned flags) {
    static hipError_t (*real)(void**, size_t, unsigned);
    if (!real) real = (hipError_t(*)(void**, size_t, unsigned))next("hipHostMalloc");
    if (should_fail(FAMILY_PINNED)) {
        if (ptr) *ptr = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(ptr, size, flags);
    if (st == HIP_SUCCESS && ptr && *ptr) count(FAMILY_PINNED, 1, (uintptr_t)*ptr, 1);
    return st;
}

hipError_t hipHostFree(void* ptr) {
    static hipError_t (*real)(void*);
    if (!real) real = (hipError_t(*)(void*))next("hipHostFree");
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(ptr);
    if (st == HIP_SUCCESS && ptr) count(FAMILY_PINNED, -1, (uintptr_t)ptr, 1);
    return st;
}

hipError_t hipEventCreateWithFlags(hipEvent_t* event, unsigned flags) {
    static hipError_t (*real)(hipEvent_t*, unsigned);
    if (!real) real = (hipError_t(*)(hipEvent_t*, unsigned))next("hipEventCreateWithFlags");
    if (should_fail(FAMILY_EVENT)) {
        if (event) *event = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(event, flags);
    if (st == HIP_SUCCESS) count(FAMILY_EVENT, 1, 0, 0);
    return st;
}

hipError_t hipEventCreate(hipEvent_t* event) {
    static hipError_t (*real)(hipEvent_t*);
    if (!real) real = (hipError_t(*)(hipEvent_t*))next("hipEventCreate");
    if (should_fail(FAMILY_EVENT)) {
        if (event) *event = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(event);
    if (st == HIP_SUCCESS) count(FAMILY_EVENT, 1, 0, 0);
    return st;
}

hipError_t hipEventDestroy(hipEvent_t event) {
    static hipError_t (*real)(hipEvent_t);
    if (!real) real = (hipError_t(*)(hipEvent_t))next("hipEventDestroy");
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(event);
    if (st == HIP_SUCCESS && event) count(FAMILY_EVENT, -1, 0, 0);
    return st;
}

hipError_t hipStreamCreateWithFlags(hipStream_t* stream, unsigned flags) {
    static hipError_t (*real)(hipStream_t*, unsigned);
    if (!real) real = (hipError_t(*)(hipStream_t*, unsigned))next("hipStreamCreateWithFlags");
    if (should_fail(FAMILY_STREAM)) {
        if (stream) *stream = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(stream, flags);
    if (st == HIP_SUCCESS) count(FAMILY_STREAM, 1, 0, 0);
    return st;
}

hipError_t hipStreamCreate(hipStream_t* stream) {
    static hipError_t (*real)(hipStream_t*);
    if (!real) real = (hipError_t(*)(hipStream_t*))next("hipStreamCreate");
    if (should_fail(FAMILY_STREAM)) {
        if (stream) *stream = NULL;
        return HIP_ERROR_OUT_OF_MEMORY;
    }
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(stream);
    if (st == HIP_SUCCESS) count(FAMILY_STREAM, 1, 0, 0);
    return st;
}

hipError_t hipStreamDestroy(hipStream_t stream) {
    static hipError_t (*real)(hipStream_t);
    if (!real) real = (hipError_t(*)(hipStream_t))next("hipStreamDestroy");
    const hipError_t st = !real ? HIP_ERROR_NOT_INITIALIZED : real(stream);
    if (st == HIP_SUCCESS && stream) count(FAMILY_STREAM, -1, 0, 0);
    return st;
}
