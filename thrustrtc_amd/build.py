"""Build the native libraries in-tree.

  libPyThrustRTC.so     the runtime + C ABI (host C++ compiled by g++, the
                        ahead-of-time gfx950 kernels compiled by hipcc, linked
                        by hipcc against libamdhip64 and libhiprtc)
  libThrustRTC.so       symbolic link to it: the C++ API's library name
  libtrtc_workloads.so  deterministic synthetic-input generators (host only)

The device headers under csrc/device/ are also embedded into the runtime as
text (csrc/host/embedded_headers.inc) so that hipRTC can instantiate the same
kernel templates at run time for user functors.

Usage:  python thrustrtc_amd/build.py [-j N] [-v]   (run by path: importing the
        package initialises the GPU)
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = os.environ.get("TRTC_OFFLOAD_ARCH", "gfx950")

DEVICE_HEADERS = [
    "trtc_args.h",
    "trtc_builtin.hpp",
    "trtc_prims.hpp",
    "trtc_scan.hpp",
    "trtc_reduce.hpp",
    "trtc_copy_if.hpp",
    "trtc_scan_by_key.hpp",
    "trtc_radix.hpp",
    "trtc_msort.hpp",
    "trtc_select.hpp",
    "trtc_merge.hpp",
    "trtc_kernels.hpp",
]

HOST_SOURCES = ["context.cpp", "dvvector.cpp", "fake_vectors.cpp", "algorithms.cpp", "algorithms_ext.cpp", "batch.cpp", "segmented.cpp", "server.cpp", "capi.cpp"]

LIB = os.path.join(PKG, "libPyThrustRTC.so")
WORKLOADS_LIB = os.path.join(PKG, "libtrtc_workloads.so")

# hipcc flags for device code: gfx950 only; packed-FP32 instructions disabled.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fvisibility=hidden", "-D__HIP_PLATFORM_AMD__",
              f"-I{ROCM}/include", "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-result"]


def _embed_headers():
    """Write csrc/host/embedded_headers.inc (raw-string copies of the device headers)."""
    out = ["// Generated by thrustrtc_amd/build.py from csrc/device/*: do not edit.",
           f"static const int k_embedded_count = {len(DEVICE_HEADERS)};",
           "static const char* const k_embedded_names[] = {"]
    out += [f'    "{h}",' for h in DEVICE_HEADERS]
    out += ["};", "static const char* const k_embedded_sources[] = {"]
    for h in DEVICE_HEADERS:
        text = open(os.path.join(CSRC, "device", h)).read()
        assert ")TRTC_EMBED\"" not in text
        out.append('R"TRTC_EMBED(' + text + ')TRTC_EMBED",')
    out.append("};")
    path = os.path.join(CSRC, "host", "embedded_headers.inc")
    data = "\n".join(out) + "\n"
    if not os.path.exists(path) or open(path).read() != data:
        with open(path, "w") as f:
            f.write(data)


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l
                    and "is not a recognized feature for this target" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    if verbose and out.strip():
        print(out)
    return out


def _compile_host(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run(["g++", *HOST_FLAGS, f"-I{ROOT}/include", "-c", src, "-o", obj], verbose)
    return obj


def _compile_hip(src, verbose):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, f"-I{ROOT}/include", "-c", src, "-o", obj], verbose)
    return obj


def build(jobs=None, verbose=False):
    """Compile every native component (always recompiles)."""
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    _embed_headers()
    jobs = jobs or min(8, os.cpu_count() or 4)
    host = [os.path.join(CSRC, "host", s) for s in HOST_SOURCES]
    aot_dir = os.path.join(CSRC, "aot")
    hip = sorted(os.path.join(aot_dir, f) for f in os.listdir(aot_dir) if f.endswith(".hip"))
    with cf.ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_compile_hip, s, verbose) for s in hip]
        futs += [ex.submit(_compile_host, s, verbose) for s in host]
        futs.append(ex.submit(_run, ["g++", "-O3", "-std=c++17", "-shared", "-fPIC",
                                     os.path.join(CSRC, "workloads", "workloads.cpp"),
                                     "-o", WORKLOADS_LIB + ".tmp"], verbose))
        objs = []
        for f in futs:
            r = f.result()
            if isinstance(r, str) and r.endswith(".o"):
                objs.append(r)
    os.replace(WORKLOADS_LIB + ".tmp", WORKLOADS_LIB)
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
          f"-L{ROCM}/lib", "-lhiprtc", "-Wl,-rpath," + f"{ROCM}/lib", "-o", LIB + ".tmp"], verbose)
    os.replace(LIB + ".tmp", LIB)
    # the reference's C++ library name (cpp/CMakeLists.txt:158): the same
    # library, so C++ callers link -lThrustRTC unchanged
    alias = os.path.join(PKG, "libThrustRTC.so")
    if os.path.islink(alias) or os.path.exists(alias):
        os.remove(alias)
    os.symlink(os.path.basename(LIB), alias)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    print(build(a.j, a.v))


if __name__ == "__main__":
    main()
