// Fixed-base SRS setup ([s_i] G for G1 / G2) and G2 subgroup validation.
#pragma once
#include "../../include/dynaverse_gpu.h"
#include "msm.hpp"

namespace dvgpu {

constexpr int SETUP_WINDOWS = 64;  // 4-bit windows over 256 scalar bits
constexpr int SETUP_TABLE = SETUP_WINDOWS * 16;

struct SetupState {
    int mode = 2;  // 0 ordinary multiplication, 1 table, 2 table + batch normalize
    DvgpuSetupStats stats{};
    DeviceBuffer table_g1, table_g2;  // affine Montgomery words, 64 x 16 entries
    bool have_g1 = false, have_g2 = false;
    DeviceBuffer scalars, points, masks, report, planes;  // planes: the repacked layout
    void attach(Quota* q) {
        for (DeviceBuffer* b : {&table_g1, &table_g2, &scalars, &points, &masks, &report, &planes}) b->quota = q;
    }
};

// Enqueue [s_i] G (device canonical scalars, 8 words each) -> canonical affine
// bytes (64 / 128 per point) in `out`.  Builds and retains the window table
// on first use.  `device_ns` (optional) gets table/generate/normalize spans.
template <class F>
void fixed_base_enqueue(SetupState& st, const uint32_t* scalars, size_t n, uint8_t* out_canonical,
                        hipStream_t stream);

// Same, but Montgomery affine words out (device-resident SRS generation).
template <class F>
void fixed_base_mont_enqueue(SetupState& st, const uint32_t* scalars, size_t n, uint32_t* out_mont,
                             hipStream_t stream);

// How one validation runs: the owner's kernel configuration (threads per
// block; schedule 1: the compact order of the seed predicate; layout 1: the
// records repacked into word planes in `scratch` (n * 128 bytes) first, when
// given; fused 1: the encoding and curve checks inside the subgroup kernel)
// and optional events after the parse, the repack and the subgroup kernel.
struct G2CheckLaunch {
    DvgpuG2KernelConfig config{128, 0, 0, 0};
    uint8_t* scratch = nullptr;
    hipEvent_t parsed = nullptr, repacked = nullptr, subgroup_done = nullptr;
};

// G2 wire validation of n points (128-byte canonical wire) on the device:
// per-point masks (1 encoding, 2 curve, 4 subgroup) and the 32-byte report.
// Kernels only (capturable).
void g2_validate_enqueue(const uint8_t* wire, size_t n, uint32_t algorithm, uint32_t* masks,
                         DvgpuG2ValidationReport* report, const G2CheckLaunch& launch, hipStream_t stream);

// Flip bit 0 of byte 0 of wire point `index` (test-only corruption).
void g2_corrupt_point(uint8_t* wire, size_t index, hipStream_t stream);

// psi, psi^2, psi^3 of homogeneous (X:Y:Z) G2 points; masks 1 = invalid.
void g2_psi_probe_enqueue(const uint8_t* points, size_t n, uint8_t* out, uint32_t* masks,
                          hipStream_t stream);

// Kernel attributes for the resources query.
void g2_kernel_attributes(uint32_t algorithm, unsigned threads, DvgpuG2KernelResources& out);

}  // namespace dvgpu
