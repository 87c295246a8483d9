"""GPU tests through the C++ API (include/*.h, libThrustRTC): a C++ caller
compiled here and run.  The C++ functions take size_t sizes where the C ABI
(python/api.h) takes int, so launches beyond 2^32 items are exercised here:
a TRTC_For over 2^32 + 1000 indices and the element-wise map (TRTC_Sequence)
over as many items both cover every index (their grids are capped and
grid-stride), and TRTC_Count reduces all of them."""
import os
import subprocess

import pytest

from helpers import ROOT, ensure_built

pytestmark = pytest.mark.gpu

PROGRAM = r'''
#include <cstdio>
#include <cstdint>
#include <vector>
#include "TRTCContext.h"
#include "DVVector.h"
#include "DeviceViewable.h"
#include "count.h"
#include "sequence.h"

int main()
{
    if (!TRTC_Try_Init()) return 2;
    const size_t n = ((size_t)1 << 32) + 1000;
    DVVector v("uint8_t", n);
    TRTC_For f({"v"}, "i", "    v[i] = (uint8_t)(i % 251);\n");
    const DeviceViewable* args[] = {&v};
    if (!f.launch_n(n, args)) return 3;
    DVUInt8 k(250);
    size_t c = 0;
    if (!TRTC_Count(v, k, c)) return 4;
    std::vector<uint8_t> tail(1000);
    v.to_host(tail.data(), n - 1000, n);
    size_t bad = 0;
    for (size_t i = 0; i < 1000; i++) bad += tail[i] != (uint8_t)((n - 1000 + i) % 251);
    printf("for_count %zu expected %zu tail_bad %zu\n", c, n / 251, bad);
    // element-wise map: v[i] = (uint8_t)i
    if (!TRTC_Sequence(v)) return 5;
    size_t c2 = 0;
    if (!TRTC_Count(v, k, c2)) return 6;
    printf("seq_count %zu expected %zu\n", c2, (n - 250 + 255) / 256);
    return (c == n / 251 && bad == 0 && c2 == (n - 250 + 255) / 256) ? 0 : 1;
}
'''


def test_cpp_api_beyond_2p32_items(tmp_path):
    ensure_built()
    src = tmp_path / "big.cpp"
    src.write_text(PROGRAM)
    exe = str(tmp_path / "big")
    libdir = os.path.join(ROOT, "thrustrtc_amd")
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-I", os.path.join(ROOT, "include"), str(src), "-L", libdir,
                        "-lThrustRTC", "-Wl,-rpath," + libdir, "-o", exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    print("running the C++ caller", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=240)
    print(r.stdout, r.stderr[-2000:], flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr[-2000:])
