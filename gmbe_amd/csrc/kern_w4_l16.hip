// One variant of the persistent enumeration kernel: 4 waves per block,
// 16384 B of LDS list arena per wave, bitsets up to 4 words, 2 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_w4_l16() { return {mbe::mbe_persistent<4, 16384, mbe::kMainW, 2>, 4, 16384}; }
}  // namespace gmbe
