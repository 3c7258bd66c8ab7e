"""Build the gfx950 HIP library (libofrr.so) in-tree with hipcc.

Every translation unit under csrc/ is compiled for gfx950 only, in parallel,
then linked into mixedprecisionofrr_amd/libofrr.so.  Packed FP32 instructions
are disabled (``-target-feature -packed-fp32-ops``).
"""
import concurrent.futures as cf
import glob
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libofrr.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
CXXFLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", os.path.join(ROOT, "include"),
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    cmd = [HIPCC] + CXXFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose=True, jobs=None):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", tmp] + objs + ["-lamdhip64"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources")
    build_selftest(verbose)
    return LIB


SELFTEST = os.path.join(HERE, "msvd_selftest")


def build_selftest(verbose=True):
    """The C++ adapter's self-test (tests/cpp/msvd_selftest.cpp over
    include/ofrr/msvd.hpp), host code linked against libofrr.so.  Built with
    hipcc (clang), which gives the adapter the reference's __half element type."""
    src = os.path.join(ROOT, "tests", "cpp", "msvd_selftest.cpp")
    tmp = SELFTEST + ".tmp"
    cmd = [HIPCC, "-O2", "-std=c++17", "-Wall", f"--offload-arch={ARCH}", "-I", os.path.join(ROOT, "include"),
           src, "-L", HERE,
           "-lofrr", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/opt/rocm/lib", "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"self-test build failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, SELFTEST)
    if verbose:
        print(f"built {SELFTEST}")


if __name__ == "__main__":
    build()
    sys.exit(0)
