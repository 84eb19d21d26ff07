// One variant of the persistent enumeration kernel: 4 waves per block,
// 8192 B of LDS list arena per wave, bitsets up to 16 words, 2 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_wide() { return {mbe::mbe_persistent<4, 8192, mbe::kMaxW, 2>, 4, 8192}; }
}  // namespace gmbe
