"""The C++ adapter (include/ofrr/msvd.hpp): the reference's msvd:: container
and solver interfaces over libofrr.so.

CPU: the header compiles as plain host C++ (g++ -fsyntax-only).
GPU: the self-test program (tests/cpp/msvd_selftest.cpp, built by build())
runs the reference's KATs and solver checks through the adapter."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "mixedprecisionofrr_amd", "msvd_selftest")


def test_header_compiles_as_host_cpp():
    src = os.path.join(ROOT, "tests", "cpp", "msvd_selftest.cpp")
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", "-I",
                        os.path.join(ROOT, "include"), src], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


def test_header_compiles_with_half_types():
    """Under clang (hipcc) the adapter has the reference's __half element type:
    the self-test's half-precision sections compile too."""
    src = os.path.join(ROOT, "tests", "cpp", "msvd_selftest.cpp")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror",
                        "-Wno-unused-command-line-argument", "--offload-arch=gfx950", "-I",
                        os.path.join(ROOT, "include"), src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]


@pytest.mark.gpu
def test_msvd_selftest():
    assert os.path.exists(BIN), "build() makes the self-test binary"
    env = dict(os.environ, OFRR_FIXTURES=os.path.join(ROOT, "tests", "golden", "suitesparse"))
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=240, env=env)
    print(r.stdout)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-2000:]
    assert "FAIL" not in r.stdout
