"""Host logic of the commit server: its first-fit device-memory arena
(zeknox_amd/csrc/arena.hpp) under a randomised alloc/free workload, built with
g++ and run on the CPU."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_arena_randomised(tmp_path):
    exe = str(tmp_path / "arena_test")
    src = os.path.join(ROOT, "tests", "native", "arena_test.cpp")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fsanitize=address,undefined", "-o", exe, src], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
