// One variant of the persistent enumeration kernel: 4 waves per block,
// 0 B of LDS list arena per wave, bitsets up to 4 words, 4 waves per SIMD.
#include "persistent.hpp"

namespace gmbe {
KernelChoice kernel_w4_l0() { return {mbe::mbe_persistent<4, 0, mbe::kMainW, 4>, 4, 0}; }
}  // namespace gmbe
