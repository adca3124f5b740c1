// step_coherent.hip -- the stepping kernel instantiated for the coherent exchange mode.
#include "step_kernel.h"

namespace femhip {
const void* step_kernel_coherent(int k, int fp32, int kr) { return pick_for_mode<kCoherent>(k, fp32, kr); }
}  // namespace femhip
