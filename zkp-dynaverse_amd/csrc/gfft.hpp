// G1 group transforms over rows of XYZZ points (kzg.hip), shared with the
// resident engine.
#pragma once
#include <cstdint>

#include "ntt.hpp"

namespace dvgpu {

// rows x 2^L XYZZ points in place.  Forward (DIF): natural order in,
// bit-reversed order out.  Inverse (DIT, inverse root, unscaled):
// bit-reversed order in, natural order out.
void gfft_rows(const NttDomain& d, uint32_t* data, int L, uint64_t rows, bool inverse, hipStream_t st);

}  // namespace dvgpu
