"""A C11 program (tests/native/abi_c_test.c) that includes include/lib.h the
way cgo does (wrappers/go/lib/lib.go:7-12: flags as char) and links
-lzeknox unchanged: argument-error conventions and the host Merkle known
answers without a device, the Go cap known answers, a coset LDE and a
transpose on the GPU."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "abi_c_test.c")
LIBDIR = os.path.join(ROOT, "zeknox_amd")


@pytest.fixture(scope="module")
def harness(tmp_path_factory):
    assert os.path.exists(os.path.join(LIBDIR, "libzeknox.so")), "build libzeknox.so first"
    exe = str(tmp_path_factory.mktemp("cabi") / "abi_c_test")
    cmd = ["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include", SRC, "-o", exe,
           "-L" + LIBDIR, "-lzeknox", "-L/opt/rocm/lib", "-lamdhip64",
           "-Wl,-rpath," + LIBDIR, "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return exe


def run(exe, mode):
    r = subprocess.run([exe, mode], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failed" in r.stdout, r.stdout


def test_c_caller_without_device(harness):
    run(harness, "cpu")


@pytest.mark.gpu
def test_c_caller_on_gpu(harness):
    run(harness, "gpu")
