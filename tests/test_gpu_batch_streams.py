"""Batched commits (zk_batch_commit) on parallel streams: every job's LDE,
leaves, digests and cap equal the same job committed alone by zk_commit (and
the oracle for the smaller ones), eagerly and from hipGraph replays that are
interleaved with unrelated eager GPU work."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def make_job(Z, lg_n, r, C, h, seed, leaves):
    from gpu_util import empty
    N = 1 << (lg_n + r)
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, seed)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    return {"coeffs": coeffs, "lde": empty(C * N), "leaves": empty(C * N) if leaves else None,
            "digests": empty(nd * 4), "cap": empty((1 << h) * 4), "lg_n": lg_n, "rate_bits": r, "ncols": C,
            "cap_height": h, "nd": nd, "seed": seed}


def alone(Z, j):
    """The same commit through zk_commit, into fresh buffers."""
    from gpu_util import empty
    out = {k: (empty(j[k].numel()) if j[k] is not None else None) for k in ("lde", "leaves", "digests", "cap")}
    Z.commit(0, j["coeffs"], out["lde"], out["leaves"], out["digests"], out["cap"], j["lg_n"], j["rate_bits"],
             j["ncols"], j["cap_height"])
    return out


def same(j, ref):
    import torch
    ok = torch.equal(j["lde"], ref["lde"]) and torch.equal(j["cap"], ref["cap"])
    ok = ok and torch.equal(j["digests"][: j["nd"] * 4], ref["digests"][: j["nd"] * 4])
    if j["leaves"] is not None:
        ok = ok and torch.equal(j["leaves"], ref["leaves"])
    return ok


def mixed_jobs(Z, n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        lg_n = int(rng.integers(3, 15))
        r = int(rng.integers(1, 4))
        C = int(rng.integers(1, 90))
        h = int(min(4, lg_n + r - 1))
        out.append(make_job(Z, lg_n, r, C, h, seed=1000 * seed + i, leaves=(i % 3 == 0)))
    return out


def test_single_job_batch_equals_commit(Z):
    j = make_job(Z, 10, 2, 20, 3, seed=5, leaves=True)
    Z.batch_commit(0, [j], graph=False)
    assert same(j, alone(Z, j))


def test_two_pass_2p20_batch_equals_commit_and_oracle(Z):
    """A 2^20-row job: the batched path runs its two 10-bit LDE passes in
    place (superblock tile order), zk_commit through the blocked
    intermediate in the library scratch; both equal the oracle's LDE."""
    from gpu_util import host
    j = make_job(Z, 17, 3, 3, 4, seed=77, leaves=False)
    Z.batch_commit(0, [j], graph=False)
    assert same(j, alone(Z, j))
    lde = O.lde(O.random(3 << 17, j["seed"]), 17, 3, 3)
    assert (host(j["lde"]) == lde).all()


def test_mixed_batch_equals_single_commits_and_oracle(Z):
    from gpu_util import host
    jobs = mixed_jobs(Z, 16, 3)
    Z.batch_commit(0, jobs, graph=False)
    for j in jobs:
        assert same(j, alone(Z, j)), (j["lg_n"], j["rate_bits"], j["ncols"], j["cap_height"])
    for j in jobs[:6]:  # the oracle for the smaller ones
        C, lg_n, r, h = j["ncols"], j["lg_n"], j["rate_bits"], j["cap_height"]
        if (C << (lg_n + r)) > (1 << 18):
            continue
        lde = O.lde(O.random(C << lg_n, j["seed"]), lg_n, r, C)
        dg, cp = O.merkle(O.transpose_rev(lde, lg_n + r, C), 1 << (lg_n + r), C, h)
        assert (host(j["lde"]) == lde).all() and (host(j["cap"]) == cp).all()
        assert (host(j["digests"], j["nd"] * 4) == dg).all()


EDGE = [  # (lg_n, rate_bits, columns, cap_height, leaf rows)
    (3, 1, 5, 2, True), (4, 2, 7, 3, False), (6, 3, 16, 4, True), (5, 1, 1, 0, False), (7, 2, 9, 6, True),
    (2, 1, 3, 3, False), (8, 1, 12, 1, True), (9, 3, 4, 0, False), (6, 2, 135, 4, True), (1, 1, 8, 1, True),
    (1, 1, 3, 1, False),
]


def test_edge_shapes_batch_equals_oracle(Z):
    """The batched forest's special cases in one call, against the oracle:
    caps == leaves (leaf hashes straight into the cap), two leaves per subtree
    (leaf hashes at d[0..2)), one subtree (cap_height 0), leaves of at most
    four elements (copied, not hashed), leaf rows on some jobs only.  The
    launch plans of these shapes are emulated thread by thread, with every
    access bounds-checked, by tests/native/forest_plan_test.cpp.  Eager, then
    graph capture and replay."""
    from gpu_util import host
    jobs = [make_job(Z, lg_n, r, C, h, seed=50 + i, leaves=rows) for i, (lg_n, r, C, h, rows) in enumerate(EDGE)]
    refs = []
    for j in jobs:
        C, lg_n, r, h = j["ncols"], j["lg_n"], j["rate_bits"], j["cap_height"]
        lde = O.lde(O.random(C << lg_n, j["seed"]), lg_n, r, C)
        dg, cp = O.merkle(O.transpose_rev(lde, lg_n + r, C), 1 << (lg_n + r), C, h)
        refs.append((lde, dg, cp))
    for call in range(3):  # eager, capture, replay
        for j in jobs:
            for k in ("lde", "leaves", "digests", "cap"):
                if j[k] is not None:
                    j[k].zero_()
        Z.batch_commit(0, jobs, graph=True)
        for j, (lde, dg, cp) in zip(jobs, refs):
            tag = (call, j["lg_n"], j["rate_bits"], j["ncols"], j["cap_height"])
            assert (host(j["lde"]) == lde).all(), tag
            assert (host(j["digests"], j["nd"] * 4) == dg).all(), tag
            assert (host(j["cap"]) == cp).all(), tag
            if j["leaves"] is not None:
                assert (host(j["leaves"]) == O.transpose_rev(lde, j["lg_n"] + j["rate_bits"], j["ncols"])).all(), tag
    Z.batch_clear(0)


def test_batch_graph_replays_match_eager(Z):
    """First call eager, second captures a hipGraph, later calls replay it;
    outputs are cleared and unrelated eager work runs between the calls."""
    import torch
    jobs = mixed_jobs(Z, 10, 9)
    refs = [alone(Z, j) for j in jobs]
    other = make_job(Z, 9, 3, 11, 2, seed=77, leaves=False)
    for rep in range(4):
        for j in jobs:
            for k in ("lde", "leaves", "digests", "cap"):
                if j[k] is not None:
                    j[k].zero_()
        Z.commit(0, other["coeffs"], other["lde"], None, other["digests"], other["cap"], 9, 3, 11, 2)
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        Z.batch_commit(0, jobs, graph=True)
        assert float(x.sum()) == float(x.sum())
        for j, ref in zip(jobs, refs):
            assert same(j, ref), rep
    Z.batch_clear(0)
