"""The mixed batched workload and the served stream at the shapes the bench
runs them, checked job by job.

* Batched: 48 commits (log2 degree 10..18, 4..135 columns, blowup 2/4/8,
  cap_height min(4, log2 rows - 2), shapes and coefficients from seed 3)
  through zk_batch_commit, eagerly, captured into a hipGraph and replayed.
  Every job's LDE, leaves, digests and cap equal the reference's three-call
  path (compute_batched_lde -> compute_transpose_rev ->
  fill_digests_buf_linear_gpu_with_gpu_ptr) run alone; the smaller jobs equal
  the oracle outright, the larger ones its Horner evaluations, leaf hashes and
  node relations at sampled positions (the GPU-vs-CPU tree comparison of the
  reference's wrappers/rust/tests/merkle_tree.rs:161-223).
* Served: requests of the served stream (log2 rows 12..20, 8..200 columns,
  blowup 8, cap_height 4, seeds splitmix64(5 ^ i)) through the commit server
  while other requests run: every cap equals zk_commit of the same request
  alone, and the oracle's (outright for the smaller requests, by samples of
  the kept LDE and digests for the larger ones), also while another thread
  commits with a different coset shift.
"""
import threading

import numpy as np
import pytest

import bench
import oracle as O

pytestmark = pytest.mark.gpu

P = O.P
ORACLE_ELEMS = 1 << 21  # jobs with at most this many LDE elements are checked against the oracle outright


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def three_call(Z, coeffs, lg_n, r, C, h):
    from gpu_util import empty
    N = 1 << (lg_n + r)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    out = {"lde": empty(C * N), "leaves": empty(C * N), "digests": empty(nd * 4), "cap": empty(4 << h)}
    bench.abi_commit(Z, coeffs, out["lde"], out["leaves"], out["digests"], out["cap"], lg_n, r, C, h)
    return out


def rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2) if bits else 0


def check_against_oracle(lg_n, r, C, h, seed, lde, digests, cap, tag):
    """LDE / digests / cap of one commit (device tensors or raw device
    addresses) against the oracle: outright for small commits, else at
    sampled positions."""
    import torch
    from gpu_util import d2h, host
    n, N = 1 << lg_n, 1 << (lg_n + r)
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0

    def fetch(buf, start, count):
        if isinstance(buf, torch.Tensor):
            return host(buf[start:start + count])
        if isinstance(buf, np.ndarray):  # a result the server already copied to the host
            return buf.reshape(-1)[start:start + count]
        return d2h(buf + 8 * start, count)

    def lde_row(row):  # LDE row `row` across all columns
        if isinstance(lde, torch.Tensor):
            return host(lde[: C * N].view(C, N)[:, row].contiguous())
        return np.array([d2h(lde + 8 * (col * N + row), 1)[0] for col in range(C)], dtype=np.uint64)

    coeffs = O.random(C << lg_n, seed)
    if C * N <= ORACLE_ELEMS:
        o_lde = O.lde(coeffs, lg_n, r, C)
        o_dg, o_cp = O.merkle(O.transpose_rev(o_lde, lg_n + r, C), N, C, h)
        assert (fetch(lde, 0, C * N) == o_lde).all(), tag
        assert (fetch(digests, 0, nd * 4) == o_dg).all(), tag
        assert (fetch(cap, 0, 4 << h) == o_cp).all(), tag
        return
    rng = np.random.default_rng(seed & 0xFFFF)
    w = pow(7, (P - 1) >> (lg_n + r), P)
    rows = [0, N - 1] + rng.integers(0, N, 2).tolist()
    for col in (0, C - 1):
        c = coeffs[col * n:(col + 1) * n]
        for i in rows:
            got = int(fetch(lde, col * N + i, 1)[0])
            assert got == O.poly_eval(c, 7 * pow(w, i, P) % P), (tag, col, i)
    L, D, lg = N >> h, 2 * (N >> h) - 2, lg_n + r
    for leaf in [0, N - 1] + rng.integers(0, N, 2).tolist():
        s, i = divmod(leaf, L)
        assert (fetch(digests, 4 * (s * D + (D - L) + i), 4) == O.hash_one(lde_row(rev(leaf, lg)))).all(), \
            (tag, leaf)
    cp = fetch(cap, 0, 4 << h).reshape(-1, 4)
    for s in (0, (1 << h) - 1):
        sd = fetch(digests, 4 * s * D, 4 * D).reshape(-1, 4)
        for node in (0, 1, D - L - 1):
            assert (sd[node] == O.hash_two(sd[2 * node + 2], sd[2 * node + 3])).all(), (tag, s, node)
        assert (cp[s] == O.hash_two(sd[0], sd[1])).all(), (tag, s)


def test_batched_workload_equals_three_call_path_and_oracle(Z):
    import torch
    from gpu_util import empty
    shapes = bench.batched_shapes()
    assert len(shapes) == 48
    jobs, refs = [], []
    for k, r, C, h, seed in shapes:
        N = 1 << (k + r)
        nd = 2 * (N - (1 << h))
        coeffs = empty(C << k)
        Z.fill_random(0, coeffs, C << k, seed)
        jobs.append({"coeffs": coeffs, "lde": empty(C * N), "leaves": empty(C * N), "digests": empty(nd * 4),
                     "cap": empty(4 << h), "lg_n": k, "rate_bits": r, "ncols": C, "cap_height": h})
        refs.append(three_call(Z, coeffs, k, r, C, h))
    torch.cuda.synchronize()
    # eager (first call with this list), graph capture (second), replay (third)
    for call in range(3):
        for j in jobs:
            for key in ("lde", "leaves", "digests", "cap"):
                j[key].zero_()
        torch.cuda.synchronize()
        Z.batch_commit(0, jobs, graph=True)
        for i, (j, ref) in enumerate(zip(jobs, refs)):
            for key in ("lde", "leaves", "digests", "cap"):
                assert torch.equal(j[key], ref[key]), (call, i, key, shapes[i])
    Z.batch_clear(0)
    for i, ((k, r, C, h, seed), ref) in enumerate(zip(shapes, refs)):
        check_against_oracle(k, r, C, h, seed, ref["lde"], ref["digests"], ref["cap"], ("job", i))


def commit_alone(Z, lg_n, r, C, h, seed, shift=7):
    from gpu_util import empty, host
    N = 1 << (lg_n + r)
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, seed)
    nd = 2 * (N - (1 << h))
    lde, dg, cp = empty(C * N), empty(nd * 4), empty(4 << h)
    Z.commit(0, coeffs, lde, None, dg, cp, lg_n, r, C, h, shift=shift)
    return host(cp)


def test_served_stream_equals_commit_alone_and_oracle(Z):
    """32 requests of the served stream, submitted at once to a server with
    16 in flight; one in four is kept so its LDE and digests can be sampled.
    A second thread runs zk_commit with shift 11 throughout: the server's
    snapshot of the coset table keeps every served result at shift 7."""
    reqs = bench.served_requests(32)
    stop = threading.Event()
    errors = []

    def other_shift():
        try:
            while not stop.is_set():
                commit_alone(Z, 9, 3, 12, 2, 99, shift=11)
        except Exception as e:  # reported by the main thread
            errors.append(e)

    with Z.Server(0, budget_bytes=64 << 30, max_inflight=16) as s:
        th = threading.Thread(target=other_shift, daemon=True)
        th.start()
        try:
            keep = [i % 4 == 0 or (C << (k + r)) <= ORACLE_ELEMS for i, (k, r, C, h, _) in enumerate(reqs)]
            tickets = [s.submit(k, r, C, h, seed=seed, keep=kp) for (k, r, C, h, seed), kp in zip(reqs, keep)]
            # in order: each result is checked (and a kept one released) while
            # the later requests are still running
            for i, ((k, r, C, h, seed), t) in enumerate(zip(reqs, tickets)):
                res = s.wait(t)
                assert res["submit_ns"] <= res["admit_ns"] <= res["done_ns"]
                assert (res["cap"] == commit_alone(Z, k, r, C, h, seed)).all(), (i, k, r, C)
                if keep[i]:
                    check_against_oracle(k, r, C, h, seed, res["lde"], res["digests"], res["cap"], ("request", i))
                    s.release(t)
                else:
                    assert res["lde"] is None
        finally:
            stop.set()
            th.join()
        assert not errors, errors


def test_server_destroy_with_kept_request_in_flight(Z):
    """Room for one request: a kept request is running and a second one waits
    for its memory when the server is destroyed.  Destroy drops the kept
    result, lets the queued request run, and returns."""
    lg_n, r, C, h = 14, 3, 64, 4
    N = 1 << (lg_n + r)
    per = 8 * C * ((1 << lg_n) + N) + 32 * 2 * N + 4096
    s = Z.Server(0, budget_bytes=int(1.5 * per), max_inflight=4)
    a = s.submit(lg_n, r, C, h, seed=1, keep=True)
    b = s.submit(lg_n, r, C, h, seed=2)
    done = threading.Event()

    def close():
        s.close()
        done.set()

    th = threading.Thread(target=close, daemon=True)
    th.start()
    th.join(timeout=60)
    assert done.is_set(), "zk_server_destroy did not return"
    del a, b
