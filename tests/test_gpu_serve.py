"""Commit server (zk_server_*): requests submitted together and admitted by a
memory budget as others finish give exactly the cap (and, when kept, the LDE
and digests) of the same request committed alone by zk_commit; admission
respects the budget; kept requests hold their memory until released."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def alone(Z, lg_n, r, C, h, seed):
    from gpu_util import empty, host
    N = 1 << (lg_n + r)
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, seed)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    lde, dg, cp = empty(C * N), empty(nd * 4), empty(4 << h)
    Z.commit(0, coeffs, lde, None, dg, cp, lg_n, r, C, h)
    return host(lde), host(dg, nd * 4), host(cp)


def shapes(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        lg_rows = int(rng.integers(6, 17))
        out.append((lg_rows - 3, 3, int(rng.integers(1, 70)), int(min(4, lg_rows - 1)), 500 + i))
    return out


def test_served_results_equal_commit_alone(Z):
    from gpu_util import d2h
    reqs = shapes(24, 1)
    with Z.Server(0, budget_bytes=2 << 30, max_inflight=8) as s:
        tickets = [s.submit(lg_n, r, C, h, seed=seed, keep=(i % 4 == 0)) for i, (lg_n, r, C, h, seed) in
                   enumerate(reqs)]
        results = [s.wait(t) for t in tickets]
        for i, ((lg_n, r, C, h, seed), res) in enumerate(zip(reqs, results)):
            lde, dg, cp = alone(Z, lg_n, r, C, h, seed)
            assert (res["cap"] == cp).all(), (i, lg_n, C, h)
            assert res["submit_ns"] <= res["admit_ns"] <= res["done_ns"]
            if i % 4 == 0:
                assert (d2h(res["lde"], lde.size) == lde).all()
                assert (d2h(res["digests"], dg.size) == dg).all()
                s.release(tickets[i])
            else:
                assert res["lde"] is None and res["digests"] is None


def test_budget_limits_admission(Z):
    """Room for about two requests at a time: all complete, correctly, and
    never more bytes are admitted at once than the budget."""
    lg_n, r, C, h = 12, 3, 24, 4
    per = 8 * C * ((1 << lg_n) + (1 << (lg_n + r))) + 32 * 2 * (1 << (lg_n + r)) + 4096
    with Z.Server(0, budget_bytes=int(2.5 * per), max_inflight=8) as s:
        tickets = [s.submit(lg_n, r, C, h, seed=900 + i) for i in range(8)]
        res = [s.wait(t) for t in tickets]
    ref = alone(Z, lg_n, r, C, h, 903)[2]
    assert (res[3]["cap"] == ref).all()
    events = sorted([(x["admit_ns"], x["device_bytes"]) for x in res] + [(x["done_ns"], -x["device_bytes"]) for x in res])
    held = peak = 0
    for _, b in events:
        held += b
        peak = max(peak, held)
    assert peak <= int(2.5 * per)
    assert max(x["admit_ns"] for x in res) > min(x["done_ns"] for x in res)  # later ones waited for memory


def test_kept_requests_hold_memory_until_released(Z):
    lg_n, r, C, h = 11, 3, 16, 3
    per = 8 * C * ((1 << lg_n) + (1 << (lg_n + r))) + 32 * 2 * (1 << (lg_n + r)) + 4096
    with Z.Server(0, budget_bytes=int(2.5 * per), max_inflight=4) as s:
        a = s.submit(lg_n, r, C, h, seed=1, keep=True)
        b = s.submit(lg_n, r, C, h, seed=2, keep=True)
        ra, rb = s.wait(a), s.wait(b)
        c = s.submit(lg_n, r, C, h, seed=3)  # does not fit while a and b are kept
        s.release(a)
        rc = s.wait(c)
        assert rc["admit_ns"] >= ra["done_ns"]
        assert (rc["cap"] == alone(Z, lg_n, r, C, h, 3)[2]).all()
        s.release(b)


def request_bytes(lg_n, r, C, h):
    """The server's arena bytes for one request with synthetic coefficients."""
    a = lambda x: (x + 255) // 256 * 256  # noqa: E731
    n, N = 1 << lg_n, 1 << (lg_n + r)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    return a(8 * C * n) + a(8 * C * N) + a(32 * nd) + a(32 << h)


def test_special_shapes_admitted_as_one_group_equal_oracle(Z):
    """A budget that holds one large request keeps eight small ones waiting;
    when it finishes they are admitted together, as one group whose Merkle
    forests run as fused launches: caps == leaves (leaf hashes straight into
    the cap), two leaves per subtree, one subtree, leaves of at most four
    elements.  Every cap equals the oracle's and zk_commit's alone."""
    import oracle as O
    big = (14, 2, 120, 4, 77)
    small = [(3, 1, 5, 4, 1), (4, 2, 7, 5, 2), (5, 1, 1, 0, 3), (6, 2, 9, 7, 4), (2, 1, 3, 3, 5), (8, 1, 12, 1, 6),
             (6, 2, 135, 4, 7), (1, 1, 8, 1, 8)]
    budget = request_bytes(*big[:4])
    assert sum(request_bytes(*s[:4]) for s in small) <= budget
    with Z.Server(0, budget_bytes=budget, max_inflight=16) as s:
        tb = s.submit(*big[:4], seed=big[4])
        ts = [s.submit(lg_n, r, C, h, seed=seed) for lg_n, r, C, h, seed in small]
        rb = s.wait(tb)
        rs = [s.wait(t) for t in ts]
    assert (rb["cap"] == alone(Z, *big)[2]).all()
    for (lg_n, r, C, h, seed), res in zip(small, rs):
        assert res["admit_ns"] >= rb["done_ns"] - 1, "a small request was admitted beside the large one"
        lde = O.lde(O.random(C << lg_n, seed), lg_n, r, C)
        _, cp = O.merkle(O.transpose_rev(lde, lg_n + r, C), 1 << (lg_n + r), C, h)
        assert (res["cap"] == cp).all(), (lg_n, r, C, h)
        assert (res["cap"] == alone(Z, lg_n, r, C, h, seed)[2]).all(), (lg_n, r, C, h)
