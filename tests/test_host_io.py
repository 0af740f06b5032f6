"""The C++ host layer's file readers and frame sources, on the CPU.

`host_io_checks.cpp` restates the reference's own tests of the spectral
library CSV (tests/test_spectra.cpp), ENVI cubes (tests/test_io.cpp:108-225)
and frame sources (tests/test_sources.cpp) against `hsi_host.hpp`; this test
compiles it with g++ against the built `libhsi_host.so` and runs it. Nothing
here touches a GPU.
"""
import os
import shutil
import subprocess

import pytest

from conftest import PKG, REPO

LIB = os.path.join(PKG, "lib")


def test_host_io_and_sources_match_the_reference_tests(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not installed")
    host_lib = os.path.join(LIB, "libhsi_host.so")
    assert os.path.exists(host_lib), "libhsi_host.so is not built: run __graft_entry__.build()"
    exe = str(tmp_path / "host_io_checks")
    cmd = ["g++", "-std=c++17", "-O1", "-Wall",
           "-I" + os.path.join(REPO, "include"), "-I" + os.path.join(PKG, "host"),
           "-I" + os.path.join(PKG, "csrc"), "-I/opt/rocm/include", "-D__HIP_PLATFORM_AMD__",
           os.path.join(REPO, "tests", "host_io_checks.cpp"), "-o", exe,
           "-L" + LIB, "-lhsi_host", "-lhsi_mi355x", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + LIB, "-Wl,-rpath,/opt/rocm/lib"]
    built = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert built.returncode == 0, built.stderr[-4000:]
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    run = subprocess.run([exe, str(scratch)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout[-4000:] + run.stderr[-2000:]
    assert run.stdout.strip().endswith("0 failed")
