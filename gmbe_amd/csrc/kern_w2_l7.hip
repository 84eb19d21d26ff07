// One variant of the persistent enumeration kernel: 2 waves per block,
// 7168 B of LDS list arena per wave, bitsets up to 4 words, 4 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_w2_l7() { return {mbe::mbe_persistent<2, 7168, mbe::kMainW, 4>, 2, 7168}; }
}  // namespace gmbe
