"""Launch plans of the batched Merkle forest (zeknox_amd/csrc/forest_map.hpp:
forest_tables, forest_find, forest_node) emulated on the host thread by thread
in launch order: every leaf hashed once into its slot, every tree node reads
children an earlier launch wrote, every digest and cap written once, nothing
outside its buffer, no launch above 2^32 threads
(tests/native/forest_plan_test.cpp).  Built for the host only under
AddressSanitizer and UBSan."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_forest_plans_cover_every_node_once_in_order(tmp_path):
    exe = str(tmp_path / "forest_plan_test")
    src = os.path.join(ROOT, "tests", "native", "forest_plan_test.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Xarch_host",
                        "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-o", exe, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr
