// step_fenced.hip -- the stepping kernel instantiated for the fenced exchange mode.
#include "step_kernel.h"

namespace femhip {
const void* step_kernel_fenced(int k, int fp32, int kr) { return pick_for_mode<kFenced>(k, fp32, kr); }
}  // namespace femhip
