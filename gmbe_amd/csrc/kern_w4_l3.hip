// One variant of the persistent enumeration kernel: 4 waves per block,
// 3072 B of LDS list arena per wave, bitsets up to 4 words, 4 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_w4_l3() { return {mbe::mbe_persistent<4, 3072, mbe::kMainW, 4>, 4, 3072}; }
}  // namespace gmbe
