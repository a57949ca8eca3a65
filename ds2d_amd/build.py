"""Build libds2d.so (HIP, gfx950) in-tree with hipcc; also the CPU oracle.

Usage: python -m ds2d_amd.build [--jobs N] [--no-oracle]
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
LIB = os.path.join(HERE, "libds2d.so")
SOURCES = ["api.hip", "kernels_ref.hip", "kernels_setup.hip", "kernels_laplace.hip", "kernels_acoustic.hip", "kernels_fast.hip",
           "kernels_fused.hip", "kernels_fused2.hip", "kernels_acoustic_fused.hip", "kernels_eborn.hip", "kernels_eborn_fused.hip", "kernels_rot.hip", "kernels_rot_fused.hip", "unet_api.hip",
           "unet_kernels.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("DS2D_ARCH", "gfx950")
# Packed-FP32 instructions are disabled for all device code.
FLAGS = ["-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-I", os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed:\n" + " ".join(cmd) + "\n" + r.stdout + r.stderr)
    return r


def build_lib(jobs=8):
    os.makedirs(OBJ, exist_ok=True)
    objs = []

    def one(src):
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        _run([HIPCC, *FLAGS, "-c", os.path.join(CSRC, src), "-o", o])
        return o

    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(one, SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    return LIB


def build_driver():
    """ds2d_amd/bin/ds2d_lsrtm: the parameter-file driver (plain C++ over the C-ABI)."""
    out = os.path.join(HERE, "bin")
    os.makedirs(out, exist_ok=True)
    exe = os.path.join(out, "ds2d_lsrtm")
    _run(["g++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "driver", "ds2d_lsrtm.cpp"), "-o", exe, "-L", HERE, "-l:libds2d.so",
          "-Wl,-rpath,$ORIGIN/.."])
    return exe


def build_oracle():
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("--no-oracle", action="store_true")
    a = ap.parse_args(argv)
    print("built", build_lib(a.jobs))
    print("built", build_driver())
    if not a.no_oracle:
        build_oracle()
        print("built oracle")


if __name__ == "__main__":
    sys.exit(main())
