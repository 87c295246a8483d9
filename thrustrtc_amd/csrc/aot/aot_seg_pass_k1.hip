// Batched-segment digit passes for key kind 1 (signed keys).
#include "aot_seg_pass.h"

namespace {
RegSegPass<1> reg_seg_pass_k1;
}  // namespace
