"""Build libsuzerain_amd.so (every HIP kernel and the C ABI) for gfx950.

Run ``python -m suzerain_amd.build`` or ``__graft_entry__.build()``.  The
library is written next to this file so it travels with the source tree.
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libsuzerain_amd.so")
ARCH = os.environ.get("SZ_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 VALU forms.
# -ffp-contract=off: fused multiply-adds only where the source writes fma(),
# so every kernel has one defined rounding (shared with the oracle).
HIPFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]


def _compile(src):
    obj = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = ["hipcc", *HIPFLAGS, "-c", src, "-o", obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{res.stderr}")
    return obj


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) +
                  glob.glob(os.path.join(CSRC, "*.cpp")))
    jobs = min(8, len(srcs))
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs,
           "-o", tmp, "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx",
           "-Wl,-rpath,/opt/rocm/lib", "-Wl,-z,defs"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", file=sys.stderr)
    return LIB


DROPIN_SRC = os.path.join(os.path.dirname(HERE), "tests", "dropin",
                          "dropin_loop.cpp")
DROPIN_BIN = os.path.join(os.path.dirname(HERE), "tests", "dropin",
                          "dropin_loop")


def build_dropin(verbose=False):
    """Compile the reference-style C++ caller of the per-system solver
    (tests/dropin/dropin_loop.cpp) with the host C++ compiler against
    include/ and the library: the boundary as a C++ user links it."""
    root = os.path.dirname(HERE)
    cmd = ["g++", "-std=c++17", "-O2", "-Wall",
           "-I", os.path.join(root, "include"), "-I", "/opt/rocm/include",
           "-D__HIP_PLATFORM_AMD__", DROPIN_SRC, "-L", HERE, "-lsuzerain_amd",
           "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../suzerain_amd", "-Wl,-rpath,/opt/rocm/lib",
           "-o", DROPIN_BIN]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"g++ failed for {DROPIN_SRC}:\n{res.stderr}")
    if verbose:
        print(f"built {DROPIN_BIN}", file=sys.stderr)
    return DROPIN_BIN


if __name__ == "__main__":
    build(verbose=True)
    build_dropin(verbose=True)
