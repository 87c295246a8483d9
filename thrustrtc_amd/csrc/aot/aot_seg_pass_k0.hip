// Batched-segment digit passes for key kind 0 (unsigned keys).
#include "aot_seg_pass.h"

namespace {
RegSegPass<0> reg_seg_pass_k0;
}  // namespace
