// Exception types that map onto the C ABI's status codes (host-only, no HIP
// dependency, so host-side logic can be built and tested without a GPU).
#pragma once
#include <stdexcept>
#include <string>

namespace dvgpu {

// Invalid caller input: maps to status -1 and leaves the owner healthy.
struct InvalidArgument : std::invalid_argument {
    using std::invalid_argument::invalid_argument;
};

// Device/runtime failure: maps to status -2 and retires the owner.
struct DeviceError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// The owner was retired (by another thread's failure) while this call waited:
// maps to status -4 without retiring again.
struct OwnerRetired : std::exception {
    const char* what() const noexcept override { return "backend retired after a device error"; }
};

inline void require(bool ok, const char* message) {
    if (!ok) throw InvalidArgument(message);
}
inline void require(bool ok, const std::string& message) {
    if (!ok) throw InvalidArgument(message);
}

}  // namespace dvgpu
