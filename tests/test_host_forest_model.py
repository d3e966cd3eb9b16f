"""Host model of the work-queue protocol of the persistent multi-job Merkle
forest that an early batched commit ran in one launch (the kernel that ended in
a device memory fault and was removed): host threads stand in for workgroups,
claim items through one atomic queue head and wait on per-job, per-height done
counters.  Every digest, leaf and cap address is bounds-checked and every child
read is checked to have been written first (tests/native/forest_queue_model.cpp).

What it establishes, on the job list the kernel faulted on and on a mixed list:
the protocol's addresses stay in bounds in every mode, including a wait that
times out and proceeds (the design as it ran), which reads unwritten children
(wrong digests, not a fault); a wait that stops the workgroup instead (the rule
for any replacement) reads nothing unwritten and reports the error.  Built for
the host only under AddressSanitizer and UBSan."""
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def model(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("needs hipcc")
    exe = str(tmp_path_factory.mktemp("forest") / "forest_queue_model")
    src = os.path.join(ROOT, "tests", "native", "forest_queue_model.cpp")
    r = subprocess.run([HIPCC, "-x", "hip", "--cuda-host-only", "-O2", "-std=c++17", "-Xarch_host",
                        "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined", "-o", exe, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    return exe


def run(exe, workgroups, spin, proceed):
    r = subprocess.run([exe, str(workgroups), str(spin), "1" if proceed else "0"], capture_output=True,
                       text=True, timeout=600)
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    counts = [tuple(int(x) for x in m) for m in
              re.findall(r"(\d+) out-of-bounds, (\d+) unwritten reads, (\d+) jobs with wrong", r.stdout)]
    assert len(counts) == 2, r.stdout
    return r, counts


def test_protocol_is_correct_when_waits_complete(model):
    for w in (1, 8, 32):
        r, counts = run(model, w, 1 << 30, proceed=False)
        assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout
        assert all(c == (0, 0, 0) for c in counts), r.stdout


def test_timeouts_never_leave_the_buffers(model):
    # spin limit 0: a wait whose inputs are not all done times out at once
    for proceed in (True, False):
        r, counts = run(model, 8, 0, proceed)
        assert all(oob == 0 for oob, _, _ in counts), r.stdout
        if not proceed:
            # stopping instead of proceeding never reads an unwritten child
            assert all(u == 0 for _, u, _ in counts), r.stdout
