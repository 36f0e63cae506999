"""Build recipe for the HIP library (gfx950) and the C++ host driver.

    python -m amrsolver_amd.build              # library and the cns_driver executable -> amrsolver_amd/lib/
    python -m amrsolver_amd.build --no-driver  # the library only

Device code is compiled with -ffp-contract=off (every operation rounds as
the reference's CPU build does) and with packed-FP32 instructions disabled.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")
LIBNAME = os.path.join(LIB, "libamrsolver_hip.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("AMR_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-ffp-contract=off",
    "-I", os.path.join(ROOT, "include"), "-I", CSRC,
]
DEVICE = [
    f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]

SOURCES = ["capi.cpp", "unfused.hip", "fused.hip", "fill_reduce.hip", "adjoint.hip", "dir_flux.hip", "regrid.hip",
           "user_source.cpp"]


def _compile(src):
    out = os.path.join(OBJ, os.path.basename(src) + ".o")
    cmd = [HIPCC] + COMMON + DEVICE
    if src.endswith(".hip"):
        cmd += ["-x", "hip"]
    cmd += ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build_library(jobs=4, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIBNAME + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-z,defs", "-o", tmp] + objs + ["-lhiprtc"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIBNAME)
    if verbose:
        print(f"built {LIBNAME}")
    return LIBNAME


def build_driver(verbose=True):
    exe = os.path.join(LIB, "cns_driver")
    cmd = [HIPCC] + COMMON + [
        os.path.join(CSRC, "driver_main.cpp"),
        "-L", LIB, "-lamrsolver_hip", f"-Wl,-rpath,{LIB}", "-o", exe,
    ]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built {exe}")
    return exe


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--no-driver", action="store_true", help="build the library only")
    ap.add_argument("-j", type=int, default=4)
    a = ap.parse_args(argv)
    build_library(a.j)
    if not a.no_driver:
        build_driver()


if __name__ == "__main__":
    sys.exit(main())
