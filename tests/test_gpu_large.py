"""Large single commit on one GPU: 400 columns of 2^18 evaluations -> iNTT to
coefficients (compute_batched_ntt, inverse) -> coset LDE with blowup 8 to 2^21
rows (6.7 GB) -> Poseidon Merkle over 2^21 leaves of 400 elements (50
absorptions each), cap_height 4.  The oracle would need minutes here, so the
checks are size-independent: the fused zk_commit equals the three-call ABI
path everywhere; sampled coefficients re-evaluate to the input values; sampled
LDE entries equal direct evaluation at the coset points; sampled leaf rows,
leaf digests, internal nodes and every cap entry equal the oracle recomputed
from the device's own inputs and children."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu
P = O.P


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def rev(x, bits):
    return int(format(x, "0%db" % bits)[::-1], 2)


def test_large_commit_400x2p18(Z):
    import torch
    from gpu_util import empty, host
    C, lg_n, r, h = 400, 18, 3, 4
    n, N, lg = 1 << lg_n, 1 << (lg_n + r), lg_n + r
    vals = empty(C * n)
    Z.fill_random(0, vals, C * n, 4)
    vals_h = host(vals).reshape(C, n).copy()
    coeffs = vals  # in place: values -> coefficients
    Z.ntt_raw(0, coeffs, lg_n, 1, Z.NTTConfig(batches=C, are_inputs_on_device=True, are_outputs_on_device=True))
    coeffs_h = host(coeffs).reshape(C, n)
    w_n = pow(7, (P - 1) >> lg_n, P)
    rng = np.random.default_rng(11)
    cols = [0, 211, C - 1]
    for col in cols:
        for i in [0, n - 1] + rng.integers(0, n, 2).tolist():
            assert O.poly_eval(coeffs_h[col], pow(w_n, i, P)) == int(vals_h[col, i]), (col, i)

    nd = 2 * (N - (1 << h))
    lde, leaves, dg, cp = empty(C * N), empty(C * N), empty(nd * 4), empty(64)
    Z.commit(0, coeffs, lde, leaves, dg, cp, lg_n, r, C, h)

    # the three-call ABI path on the same coefficients
    a_lde = empty(C * N)
    Z.lde_batch(0, a_lde, coeffs, lg_n, Z.NTTConfig(batches=C, extension_rate_bits=r, with_coset=True,
                                                    are_inputs_on_device=True, are_outputs_on_device=True))
    assert torch.equal(a_lde, lde)
    a_leaves = empty(C * N)
    Z.transpose_rev_batch(0, a_leaves, lde, lg, Z.TransposeConfig(batches=C, are_inputs_on_device=True,
                                                                  are_outputs_on_device=True))
    assert torch.equal(a_leaves, leaves)
    del a_lde, a_leaves
    a_dg, a_cp = empty(nd * 4), empty(64)
    Z.fill_digests_buf_linear_gpu_with_gpu_ptr(a_dg, a_cp, leaves, nd, 1 << h, N, C, h)
    assert torch.equal(a_dg, dg) and torch.equal(a_cp, cp)

    lde2 = lde.view(C, N)
    w = pow(7, (P - 1) >> lg, P)
    rows = [0, 1, N - 1] + rng.integers(0, N, 3).tolist()
    sample = host(lde2[:, torch.tensor(rows, device="cuda")].contiguous()).reshape(C, len(rows))
    for col in cols:
        for k, i in enumerate(rows):
            assert int(sample[col, k]) == O.poly_eval(coeffs_h[col], 7 * pow(w, i, P) % P), (col, i)

    L, D = N >> h, 2 * (N >> h) - 2
    leaves2 = leaves.view(N, C)
    dg2 = dg.view(-1, 4)
    for k, row in enumerate(rows):
        leaf = rev(row, lg)
        lr = host(leaves2[leaf].contiguous())
        assert (lr == sample[:, k]).all(), row  # leaf rev(row) is LDE row `row`
        s, i = divmod(leaf, L)
        assert (host(dg2[s * D + (D - L) + i].contiguous()) == O.hash_one(lr)).all(), leaf
    cp_h = host(cp).reshape(-1, 4)
    for s in range(1 << h):
        sd = host(dg2[s * D:(s + 1) * D].contiguous()).reshape(-1, 4)
        # first and last node of the top levels and of the level above the leaves
        for node in (0, 1, 2, 5, L // 2 - 2, L - 4, L - 3):
            assert (sd[node] == O.hash_two(sd[2 * node + 2], sd[2 * node + 3])).all(), (s, node)
        assert (cp_h[s] == O.hash_two(sd[0], sd[1])).all(), s
