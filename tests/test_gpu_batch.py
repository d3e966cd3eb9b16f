"""Batched commits (zk_batch_commit) as first written against a persistent
Merkle-forest kernel: every job's LDE, digests and cap equal the oracle (small
jobs) and the single-job zk_commit (mixed sizes); hipGraph replays reproduce
the eager result.  Kept skipped: those first GPU runs ended in a device memory
fault whose cause was not found.  The current implementation (LDEs on four
streams, then fused forest launches in stream order, no device-side waits) is
covered by test_gpu_batch_streams.py, including these shapes against the
oracle, and by test_gpu_workloads.py."""
import numpy as np
import pytest

import oracle as O

pytestmark = [
    pytest.mark.gpu,
    pytest.mark.skip(reason="the persistent Merkle-forest implementation these cases first ran against ended "
                            "in a device memory fault whose cause was not found; the case stays off the GPU "
                            "(the current implementation is tested in test_gpu_batch_streams.py)"),
]
P = O.P


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def make_job(Z, lg_n, r, C, h, seed, leaves=False):
    from gpu_util import empty
    N = 1 << (lg_n + r)
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, seed)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    return {"coeffs": coeffs, "lde": empty(C * N), "leaves": empty(C * N) if leaves else None,
            "digests": empty(nd * 4), "cap": empty((1 << h) * 4), "lg_n": lg_n, "rate_bits": r, "ncols": C,
            "cap_height": h, "nd": nd, "seed": seed}


def oracle_commit(j):
    C, lg_n, r, h = j["ncols"], j["lg_n"], j["rate_bits"], j["cap_height"]
    coeffs = O.random(C << lg_n, j["seed"])
    lde = O.lde(coeffs, lg_n, r, C)
    dg, cp = O.merkle(O.transpose_rev(lde, lg_n + r, C), 1 << (lg_n + r), C, h)
    return lde, dg, cp


SMALL = [  # (lg_n, rate_bits, columns, cap_height)
    (3, 1, 5, 2), (4, 2, 7, 3), (6, 3, 16, 4), (5, 1, 1, 0), (7, 2, 9, 6), (2, 1, 3, 3),
    (8, 1, 12, 1), (9, 3, 4, 0), (6, 2, 135, 4),
]


def test_small_batch_equals_oracle(Z):
    from gpu_util import host
    jobs = [make_job(Z, *s, seed=50 + i, leaves=(i % 2 == 0)) for i, s in enumerate(SMALL)]
    Z.batch_commit(0, jobs, graph=False)
    for j in jobs:
        lde, dg, cp = oracle_commit(j)
        tag = (j["lg_n"], j["rate_bits"], j["ncols"], j["cap_height"])
        assert (host(j["lde"]) == lde).all(), tag
        assert (host(j["digests"], j["nd"] * 4) == dg).all(), tag
        assert (host(j["cap"]) == cp).all(), tag
        if j["leaves"] is not None:
            N = 1 << (j["lg_n"] + j["rate_bits"])
            assert (host(j["leaves"]) == O.transpose_rev(lde, j["lg_n"] + j["rate_bits"], j["ncols"])).all()


def mixed_jobs(Z, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        lg_n = int(rng.integers(8, 15))
        r = int(rng.integers(1, 4))
        C = int(rng.integers(4, 80))
        h = min(4, lg_n + r - 2)
        out.append(make_job(Z, lg_n, r, C, h, seed=1000 * seed + i))
    return out


def test_mixed_batch_equals_single_commits(Z):
    import torch
    jobs = mixed_jobs(Z, 12, 7)
    Z.batch_commit(0, jobs, graph=False)
    for j in jobs:
        from gpu_util import empty
        lde, dg, cp = empty(j["lde"].numel()), empty(j["digests"].numel()), empty(j["cap"].numel())
        Z.commit(0, j["coeffs"], lde, None, dg, cp, j["lg_n"], j["rate_bits"], j["ncols"], j["cap_height"])
        assert torch.equal(lde, j["lde"]) and torch.equal(dg, j["digests"]) and torch.equal(cp, j["cap"])


def test_graph_replays_match_eager(Z):
    """Eager first call, then captured graph replays; between replays the
    outputs are cleared and unrelated eager GPU work runs."""
    import torch
    jobs = mixed_jobs(Z, 8, 9)
    Z.batch_commit(0, jobs, graph=True)  # eager (builds the plan)
    ref = [(j["lde"].clone(), j["digests"].clone(), j["cap"].clone()) for j in jobs]
    other = make_job(Z, 10, 2, 20, 3, seed=77)
    for rep in range(4):
        for j in jobs:
            j["lde"].zero_()
            j["digests"].zero_()
            j["cap"].zero_()
        Z.commit(0, other["coeffs"], other["lde"], None, other["digests"], other["cap"], 10, 2, 20, 3)
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        Z.batch_commit(0, jobs, graph=True)  # capture on rep 0, replay afterwards
        assert float(x.sum()) == float(x.sum())
        for j, (l, d, c) in zip(jobs, ref):
            assert torch.equal(j["lde"], l) and torch.equal(j["digests"], d) and torch.equal(j["cap"], c), rep
    Z.batch_clear(0)
