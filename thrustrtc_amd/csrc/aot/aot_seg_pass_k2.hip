// Batched-segment digit passes for key kind 2 (float keys).
#include "aot_seg_pass.h"

namespace {
RegSegPass<2> reg_seg_pass_k2;
}  // namespace
