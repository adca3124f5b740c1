// step_plain.hip -- the stepping kernel instantiated for the plain exchange mode.
#include "step_kernel.h"

namespace femhip {
const void* step_kernel_plain(int k, int fp32, int kr) { return pick_for_mode<kPlain>(k, fp32, kr); }
}  // namespace femhip
