"""The relaxed Goldilocks operations of zeknox_amd/csrc/gl64.hpp (the forms the
device kernels run: 64x64 product and square, relaxed add/sub/reduce, powers of
two) against exact 128-bit integer arithmetic mod p on edge and random operands,
compiled for the host with clang (the carry builtins are clang's) and run on
the CPU."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLANG = "/opt/rocm/llvm/bin/clang++"


@pytest.mark.skipif(not os.path.exists(CLANG) and shutil.which("clang++") is None, reason="needs clang++")
def test_relaxed_field_ops_congruent_to_exact(tmp_path):
    cxx = CLANG if os.path.exists(CLANG) else shutil.which("clang++")
    exe = str(tmp_path / "field_relaxed_test")
    src = os.path.join(ROOT, "tests", "native", "field_relaxed_test.cpp")
    subprocess.run([cxx, "-O2", "-std=c++17", "-fsanitize=address,undefined", "-D__HIP_PLATFORM_AMD__",
                    "-I/opt/rocm/include", "-Wno-unused-result", "-o", exe, src], check=True,
                   capture_output=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
