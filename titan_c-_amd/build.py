"""Builds libtitan_gpu.so (HIP kernels for gfx950 + the C++ host layer) in-tree.

Usage: python titan_c-_amd/build.py [--jobs N]

Flags per translation unit:
  * FP64 parity kernels (k_step64, k_cx, k_fit, k_mk): -ffp-contract=off, IEEE division / square root, so the
    arithmetic is evaluated in the reference's operation order;
  * mixed-precision flux (k_step32): FP32 with fast division / square root, multiply-adds contracted
    within an expression only (-ffp-contract=on), so a function inlined at two sites (a cell's
    Zone_1 frame in the cell phase and in a neighbouring tile's face phase) gives the same bits at
    both, and a face on a tile border gets the same flux from both tiles;
  * every HIP unit: packed-FP32 instructions disabled (-Xclang -target-feature -Xclang -packed-fp32-ops).
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lib")
OBJ_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "libtitan_gpu.so")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_COMMON = [
    "--offload-arch=" + ARCH, "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-result", "-Wno-unused-value",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]
UNITS = {
    "k_step64.hip": ["-ffp-contract=off"],
    "k_step64sc.hip": ["-ffp-contract=off"],
    "k_cx.hip": ["-ffp-contract=off"],
    "k_fit.hip": ["-ffp-contract=off"],
    "k_mk.hip": ["-ffp-contract=off"],
    "k_step32.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "k_step32sc.hip": ["-ffp-contract=on", "-fno-hip-fp32-correctly-rounded-divide-sqrt"],
    "k_mlp.hip": ["-ffp-contract=on"],
    "k_train.hip": ["-ffp-contract=fast"],
    "k_noalloc.hip": ["-ffp-contract=off"],
}
HOST_UNITS = ["titan_host.cpp", "titan_mesh.cpp", "titan_io.cpp", "titan_fit.cpp"]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-Wall", "-D__HIP_PLATFORM_AMD__", "-I" + os.path.join(ROCM, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s\n%s" % (" ".join(cmd), r.stdout, r.stderr))
    return r.stderr


def build(jobs=8, verbose=False):
    if not os.path.exists(HIPCC):
        raise RuntimeError("hipcc not found at %s" % HIPCC)
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    cmds = []
    objs = []
    for src, extra in UNITS.items():
        obj = os.path.join(OBJ_DIR, src + ".o")
        cmds.append([HIPCC] + HIP_COMMON + extra + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    for src in HOST_UNITS:
        obj = os.path.join(OBJ_DIR, src + ".o")
        cmds.append(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    with concurrent.futures.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for err in ex.map(_run, cmds):
            if verbose and err:
                sys.stderr.write(err)
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    shutil.rmtree(OBJ_DIR, ignore_errors=True)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.verbose))
