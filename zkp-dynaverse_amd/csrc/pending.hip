// Entry points whose device implementation lands in later files.
#include "owner.hpp"
namespace dvgpu {
struct ResidentState {};
void resident_destroy(ResidentState* s) noexcept { delete s; }
DvgpuResidentStats resident_stats(const ResidentState*) { return {}; }
DvgpuResidentDiagnostics resident_diagnostics(const ResidentState*) { return {}; }
}
// Temporary: entry points whose implementation is being written.
using namespace dvgpu;
extern "C" {
#define PENDING(msg) return fail(-1, msg, error, capacity)
int dvgpu_resident_configure(void*, uint64_t, uint64_t, const uint8_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_resident_call(void*, const DvgpuResidentOp*, uint64_t*, char* error, size_t capacity) { PENDING("pending"); }
int dvgpu_resident_call_profiled(void*, const DvgpuResidentOp*, uint64_t*, DvgpuResidentProfile*, char* error, size_t capacity) { PENDING("pending"); }
void dvgpu_resident_seal(void*) {}
void dvgpu_resident_commit(void*) {}
int dvgpu_resident_stats(void*, DvgpuResidentStats*) { return -1; }
int dvgpu_resident_diagnostics(void*, DvgpuResidentDiagnostics*) { return -1; }
}
