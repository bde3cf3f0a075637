"""Builds libvrgi.so (the HIP library behind include/vrgi.h) in-tree for gfx950.

Usage: python vulkan-voxel-restir-gi_amd/build.py [--jobs N] [--out PATH] [-D NAME=VALUE ...]
Every kernel is compiled with -ffp-contract=off (one rounding per written operation,
the evaluation order of the reference's Rust / GLSL) and without fast-math; packed
FP32 instructions are disabled. -D overrides a compile-time setting (VRGI_XCD_SWIZZLE,
VRGI_NEE_STREAMS, ...) for a variant library written to --out.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvrgi.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("VRGI_ARCH", "gfx950")

SOURCES = ["bvh_build.hip", "trace.hip", "restir.hip", "scene.hip", "sort.hip", "nrc.hip", "mesher.hip", "api.cpp", "world.cpp"]
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-fast-math", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]


def _compile(src, defines=(), obj_dir=OBJ):
    out = os.path.join(obj_dir, os.path.splitext(src)[0] + ".o")
    cmd = [HIPCC, *FLAGS, *[f"-D{d}" for d in defines], "-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return out, msgs


def build(jobs=6, verbose=False, out=LIB, defines=()):
    obj_dir = OBJ if not defines else os.path.join(OBJ, "variant")
    os.makedirs(obj_dir, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda src: _compile(src, defines, obj_dir), SOURCES))
    objs = []
    for obj, msgs in results:
        objs.append(obj)
        if verbose and msgs.strip():
            print(msgs)
    tmp = out + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=6)
    ap.add_argument("-v", action="store_true")
    ap.add_argument("--out", default=LIB)
    ap.add_argument("-D", dest="defines", action="append", default=[])
    a = ap.parse_args()
    print(build(a.jobs, a.v, os.path.abspath(a.out), tuple(a.defines)))
    sys.exit(0)
