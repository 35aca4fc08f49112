"""Build the MI355X MPM workpiece solver (gfx950) in-tree.

Produces lib/libmpmw_hip.so (HIP kernels + C ABI + the C++
edgepredict::MPMWorkpieceSolver class).  Run as a script or call build().
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_DIR = os.path.join(HERE, "lib")
CSRC = os.path.join(HERE, "csrc")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # device float atomics compile to the hardware global/LDS add instructions
    "-munsafe-fp-atomics",
    # no packed-FP32 (v_pk_*_f32) code generation on this target
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if out.strip():
        print(out, flush=True)
    if r.returncode != 0:
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd)}")


def build(verbose=True):
    os.makedirs(LIB_DIR, exist_ok=True)
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    out = os.path.join(LIB_DIR, "libmpmw_hip.so")
    tmp = out + ".tmp"
    srcs = [os.path.join(CSRC, "mpmw_solver.hip"), os.path.join(CSRC, "mpmw_sort.hip"),
            os.path.join(CSRC, "mpms_solver.hip"), os.path.join(CSRC, "mpmc_solver.hip"),
            os.path.join(CSRC, "MPMWorkpieceSolver.cpp"), os.path.join(CSRC, "mpmw_toolmesh.cpp")]
    srcs = [s for s in srcs if os.path.exists(s)]
    _run([HIPCC, *HIP_FLAGS, "-shared", "-o", tmp, *srcs])
    os.replace(tmp, out)
    # the C++ drop-in class driven through the reference interfaces
    test_src = os.path.join(ROOT, "tests", "cpp", "solver_api_test.cpp")
    if os.path.exists(test_src):
        _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"), "-o",
              os.path.join(LIB_DIR, "solver_api_test"), test_src, "-L", LIB_DIR, "-lmpmw_hip",
              f"-Wl,-rpath,$ORIGIN", "-L/opt/rocm/lib", "-lamdhip64"])
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
