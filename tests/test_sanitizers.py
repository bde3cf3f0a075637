"""CPU sanitizer runs (AddressSanitizer + UndefinedBehaviorSanitizer):

- the oracle, built with g++ -fsanitize=address,undefined, through scene builds, frames in
  both ReSTIR modes and every debug view, a learned-cache frame whose training records
  overflow their buffer, the NEE pdf, the standalone builders and the mesher;
- the library, built with hipcc with both sanitizers on the host pass only (-Xarch_host
  before each -fsanitize=; the device code is the ordinary one), through its host side
  without a GPU: world generation, the CPU mesher, 200k random-walk edits, configuration
  and every entry point's argument checks.

Each driver (tests/sanitize_driver.py) runs in a subprocess with the matching sanitizer
runtime preloaded and stops at the first report."""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "vulkan-voxel-restir-gi_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
ENV = {"ASAN_OPTIONS": "detect_leaks=0:halt_on_error=1:abort_on_error=0",
       "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1"}


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-2000:] + r.stderr[-4000:]


def _driver(what, env):
    e = dict(os.environ, **ENV, **env)
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "sanitize_driver.py"), what], env=e,
                       capture_output=True, text=True, timeout=900)
    report = [l for l in r.stderr.splitlines() if "AddressSanitizer" in l or "runtime error" in l]
    assert r.returncode == 0 and not report, r.stdout[-2000:] + r.stderr[-6000:]
    assert f"{what} ok" in r.stdout


def test_oracle_under_asan_ubsan(tmp_path):
    lib = str(tmp_path / "liboracle_asan.so")
    _run(["g++", "-O1", "-g", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-fopenmp",
          "-fsanitize=address,undefined", "-fno-omit-frame-pointer", "-shared", "-o", lib,
          os.path.join(ROOT, "oracle", "oracle.cpp"), os.path.join(ROOT, "oracle", "nrc.cpp")])
    rt = subprocess.run(["g++", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    _driver("oracle", {"ORACLE_LIB": lib, "LD_PRELOAD": rt})


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_library_host_side_under_asan_ubsan(tmp_path):
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")) + glob.glob(os.path.join(CSRC, "*.cpp")))
    flags = ["--offload-arch=gfx950", "-O1", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined"]

    def comp(src):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        _run([HIPCC, *flags, "-x", "hip", "-c", src, "-o", obj])
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(comp, srcs))
    lib = str(tmp_path / "libvrgi_asan.so")
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-Xarch_host", "-fsanitize=address",
          "-Xarch_host", "-fsanitize=undefined", "-shared-libsan", "-o", lib, *objs])
    rt = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    assert rt, "clang's ASan runtime not found"
    _driver("host", {"VRGI_LIB": lib, "LD_PRELOAD": rt[-1]})
