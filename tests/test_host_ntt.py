"""The NTT pass kernels' per-thread code, run on the host thread by thread in
kernel order (tests/native/ntt_host_test.cpp, built for the host only), against
the oracle: the 10-bit pass in every role it has (first pass of an LDE with
the coset shift and the skipped replication stages, first and second pass of
forward and inverse 2^20 transforms through the blocked intermediate, the
inter-pass twiddles from a table and computed, the canonical recompute path
forced on every tile) and a 2^19 plan that mixes it with a 9-bit pass."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "ntt_host_test.cpp")
EXE = os.path.join(ROOT, "tests", "native", "_build", "ntt_host_test")


def build_harness():
    """hipcc, host code only (no device code is compiled or run)."""
    os.makedirs(os.path.dirname(EXE), exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17",
           "-I" + os.path.join(ROOT, "include"), SRC, "-o", EXE + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    os.replace(EXE + ".tmp", EXE)


@pytest.fixture(scope="module")
def harness():
    stale = not os.path.exists(EXE) or any(
        os.path.getmtime(p) > os.path.getmtime(EXE)
        for p in (SRC, os.path.join(ROOT, "zeknox_amd", "csrc", "ntt_pass.hpp"),
                  os.path.join(ROOT, "zeknox_amd", "csrc", "gl64.hpp")))
    if stale:
        build_harness()
    return EXE


CASES = [  # lg, lg_in, columns, inverse, coset, force_exact, twiddle table
    (20, 20, 1, 0, 0, 0, 1),
    (20, 17, 2, 0, 1, 0, 1),
    (20, 18, 1, 0, 1, 0, 0),
    (20, 20, 1, 1, 0, 0, 1),
    (20, 20, 1, 1, 1, 0, 0),
    (20, 20, 1, 0, 0, 1, 1),
    (20, 17, 1, 0, 1, 1, 1),
    (19, 19, 2, 0, 0, 0, 1),
    (19, 19, 1, 1, 0, 1, 0),
]


@pytest.mark.parametrize("lg,lg_in,C,inv,coset,force,table", CASES)
def test_host_run_of_pass_code_equals_oracle(harness, tmp_path, lg, lg_in, C, inv, coset, force, table):
    x = O.random(C << lg_in, 9000 + 16 * lg + lg_in + 3 * inv + 5 * coset + 7 * force)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    x.tofile(fin)
    r = subprocess.run([harness, str(lg), str(lg_in), str(C), str(inv), str(coset), str(force), str(table), fin,
                        fout], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(fout, dtype=np.uint64)
    if lg_in == lg:
        ref = O.ntt(x, lg, C, inverse=bool(inv), with_coset=bool(coset))
    else:
        ref = O.lde(x, lg_in, lg - lg_in, C, inverse=bool(inv), with_coset=bool(coset))
    assert (got == ref).all(), r.stdout
    if lg == 20:  # both passes ran the 10-bit kernel's code
        assert r.stdout.count("r10=1") == 2, r.stdout


TUNED = [  # lg, lg_in, columns, inverse, coset, plan (np b1 b2 b3 g1 g2 g3 r1 r2 r3; g < 0: auto)
    (20, 17, 1, 0, 1, (3, 7, 7, 6, -1, -1, -1, 1, 1, 1)),
    (20, 20, 1, 1, 0, (2, 10, 10, 0, 2, -1, -1, 0, 1, 1)),
    (20, 20, 1, 0, 0, (2, 10, 10, 0, -1, 1, -1, 1, 0, 1)),
    (17, 14, 3, 0, 1, (2, 10, 7, 0, -1, 2, -1, 1, 0, 1)),
    (17, 17, 2, 1, 1, (2, 7, 10, 0, 0, -1, -1, 1, 1, 1)),
    (14, 11, 5, 0, 1, (2, 8, 6, 0, 1, 0, -1, 1, 1, 1)),
    (12, 12, 7, 0, 0, (3, 4, 4, 4, -1, 0, 2, 1, 1, 1)),
]


@pytest.mark.parametrize("lg,lg_in,C,inv,coset,plan", TUNED)
def test_host_run_of_tuned_plans_equals_oracle(harness, tmp_path, lg, lg_in, C, inv, coset, plan):
    """Plans the autotuner may choose (other pass splits, fewer groups per
    workgroup, the 32-element kernel for a 10-bit pass) give the same
    transform."""
    x = O.random(C << lg_in, 7000 + 16 * lg + lg_in + 3 * inv + 5 * coset)
    fin, fout = str(tmp_path / "in.bin"), str(tmp_path / "out.bin")
    x.tofile(fin)
    r = subprocess.run([harness, str(lg), str(lg_in), str(C), str(inv), str(coset), "0", "1", fin, fout] +
                       [str(v) for v in plan], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("pass S=") == plan[0], r.stdout
    got = np.fromfile(fout, dtype=np.uint64)
    if lg_in == lg:
        ref = O.ntt(x, lg, C, inverse=bool(inv), with_coset=bool(coset))
    else:
        ref = O.lde(x, lg_in, lg - lg_in, C, inverse=bool(inv), with_coset=bool(coset))
    assert (got == ref).all(), r.stdout
