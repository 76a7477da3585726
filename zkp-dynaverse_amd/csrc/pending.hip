// Entry points whose device implementation lands in later files.
#include "owner.hpp"
namespace dvgpu {
struct ResidentState {};
struct SetupState {};
void resident_destroy(ResidentState* s) noexcept { delete s; }
DvgpuResidentStats resident_stats(const ResidentState*) { return {}; }
DvgpuResidentDiagnostics resident_diagnostics(const ResidentState*) { return {}; }
void setup_destroy(SetupState* s) noexcept { delete s; }
}
// Temporary: entry points whose implementation is being written.
using namespace dvgpu;
extern "C" {
#define PENDING(msg) return fail(-1, msg, error, capacity)
int dvgpu_bn254_fixed_base(void*, const uint8_t*, size_t, int, uint8_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_bn254_fixed_base_g2_checked(void*, const uint8_t*, size_t, uint32_t, uint8_t*, DvgpuG2ValidationReport*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_bn254_g2_validation_inspect(void*, const uint8_t*, size_t, uint32_t, uint32_t*, DvgpuG2ValidationReport*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_g2_psi_probe(void*, const uint8_t*, size_t, uint8_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_resident_configure(void*, uint64_t, uint64_t, const uint8_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_resident_call(void*, const DvgpuResidentOp*, uint64_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_resident_call_profiled(void*, const DvgpuResidentOp*, uint64_t*, DvgpuResidentProfile*, char* error, size_t capacity) { PENDING("pending"); }
void dvgpu_resident_seal(void*) {}
void dvgpu_resident_commit(void*) {}
int dvgpu_resident_stats(void*, DvgpuResidentStats*) { return -1; }
int dvgpu_resident_diagnostics(void*, DvgpuResidentDiagnostics*) { return -1; }
int dvgpu_setup_mode(void*, int) { return -1; }
int dvgpu_setup_stats(void*, DvgpuSetupStats*) { return -1; }
int dvgpu_g2_experiment_stats(void*, DvgpuG2ExperimentStats*) { return -1; }
int dvgpu_g2_kernel_config_get(void*, DvgpuG2KernelConfig*) { return -1; }
int dvgpu_g2_kernel_config_set(void*, const DvgpuG2KernelConfig*) { return -1; }
int dvgpu_g2_kernel_resources(void*, uint32_t, DvgpuG2KernelResources*) { return -1; }
int dvgpu_g2_test_corrupt_next(void*, uint64_t) { return -1; }
int dvgpu_g2_validation_stats(void*, DvgpuG2ValidationStats*) { return -1; }
int dvgpu_g2_validation_traffic(void*, DvgpuG2ValidationTraffic*) { return -1; }
}
