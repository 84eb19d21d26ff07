// One variant of the persistent enumeration kernel: 4 waves per block,
// 7168 B of LDS list arena per wave, bitsets up to 4 words, 4 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_w4_l7() { return {mbe::mbe_persistent<4, 7168, mbe::kMainW, 4>, 4, 7168}; }
}  // namespace gmbe
