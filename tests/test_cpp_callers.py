"""The C++ boundary: the reference's own C++ callers (cpp/test/*.cpp, which
include TRTCContext.h, DVVector.h, the algorithm headers and built_in.h and
link libThrustRTC, cpp/CMakeLists.txt:158) compile against include/ and link
against this library unchanged.  Compile and link only (no GPU, nothing is
run).  Skipped where the reference checkout is not present."""
import glob
import os
import subprocess

import pytest

from helpers import ROOT, ensure_built

REF_TESTS = "/root/reference/cpp/test"
SOURCES = sorted(glob.glob(os.path.join(REF_TESTS, "*.cpp")))


@pytest.fixture(scope="module")
def libdir():
    ensure_built()
    d = os.path.join(ROOT, "thrustrtc_amd")
    assert os.path.exists(os.path.join(d, "libThrustRTC.so"))
    return d


def test_reference_cpp_programs_are_present():
    if not os.path.isdir(REF_TESTS):
        pytest.skip("reference checkout not present")
    assert len(SOURCES) >= 39


@pytest.mark.skipif(not SOURCES, reason="reference checkout not present")
@pytest.mark.parametrize("src", SOURCES, ids=[os.path.basename(s) for s in SOURCES])
def test_reference_cpp_caller_compiles_and_links(src, libdir, tmp_path):
    exe = str(tmp_path / "a.out")
    cmd = ["g++", "-std=c++17", "-O0", "-I", os.path.join(ROOT, "include"), src, "-L", libdir, "-lThrustRTC",
           "-Wl,--no-undefined", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-3000:]
    assert os.path.getsize(exe) > 0
