// Registration helpers for ahead-of-time kernel instantiations: each kernel is
// registered under the same name expression the host would otherwise hand to
// hipRTC, so both paths share one lookup.
#pragma once
#include <string>
#include "../device/trtc_kernels.hpp"
#include "../host/trtc_internal.h"

#define TRTC_REG(NAME, FN) trtc_host::register_aot((NAME), (const void*)(FN))

inline std::string vv(const char* t) { return std::string("VectorView<") + t + ">"; }
