"""The fused commit's leaf-row writer (zeknox_amd/csrc/leaf_rows.hpp) driven by
the sponge's chunk hook on the host: every row lands where the row-major leaf
matrix puts it, for leaf sizes 1..20, 135 and 400 and every 8-byte misalignment
of the matrix, with nothing written outside it; sectors inside a row are stored
whole (tests/native/row_writer_test.cpp).  Built for the host only under
AddressSanitizer and UBSan."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="needs hipcc")
def test_row_writer_writes_whole_sectors_in_bounds(tmp_path):
    exe = str(tmp_path / "row_writer_test")
    src = os.path.join(ROOT, "tests", "native", "row_writer_test.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Xarch_host",
                        "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-o", exe, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
