"""The lane-parallel Poseidon permutations of the narrow Merkle levels
(zeknox_amd/csrc/poseidon_hex.hpp: 16 lanes with one state element each;
poseidon_quad.hpp: 4 lanes with three) emulated on the host, the lanes of a
group in lock step with the gather done as the device exchange does it, against
the exact one-lane permutation: random states, edge values, and states whose
first S-box square takes the relaxed arithmetic's rare case (which the relaxed
path must report, and the exact path must get right).  Built for the host only
(hipcc --cuda-host-only) under AddressSanitizer and UBSan."""
import os
import subprocess

import pytest

from rare_inputs import first_sbox_rare_inputs

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_lane_parallel_permutations_equal_one_lane(tmp_path):
    exe = str(tmp_path / "poseidon_lanes_test")
    src = os.path.join(ROOT, "tests", "native", "poseidon_lanes_test.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Xarch_host",
                        "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-o", exe, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    args = []
    for element in (0, 1):
        for x in first_sbox_rare_inputs(element, 4):
            args += [str(element), "%x" % x]
    r = subprocess.run([exe] + args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
    relaxed_rare = int(r.stdout.split(",")[1].split()[0])
    assert relaxed_rare >= 16, r.stdout  # 8 rare states x (hex, quad)
