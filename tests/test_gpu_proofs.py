"""Merkle openings for FRI queries (zk_merkle_proofs): sibling paths equal the
oracle's extraction from the same digests, leaf rows equal the leaves, and
every opening recomputes its cap entry; also straight from a fused commit's
column-major LDE."""
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


@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [(2, 7, 0), (16, 3, 3), (64, 9, 0), (1024, 16, 4),
                                                           (4096, 135, 2), (256, 5, 8)])
def test_openings_equal_oracle_and_verify(Z, n_leaves, leaf_size, cap_height):
    from gpu_util import dev, empty, host
    leaves = O.random(n_leaves * leaf_size, 77 + n_leaves)
    ncaps = 1 << cap_height
    nd = 2 * (n_leaves - ncaps) if ncaps < n_leaves else 0
    dleaves, dg, cp = dev(leaves), empty(nd * 4), empty(ncaps * 4)
    Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dg, cp, dleaves, nd, ncaps, n_leaves, leaf_size, cap_height)
    dg_h, cp_h = host(dg, nd * 4), host(cp)
    rng = np.random.default_rng(n_leaves)
    idx = [0, n_leaves - 1] + rng.integers(0, n_leaves, 20).tolist()
    sib, rows = Z.merkle_proofs(0, dg, dleaves, n_leaves, leaf_size, cap_height, idx)
    lv = leaves.reshape(n_leaves, leaf_size)
    for k, i in enumerate(idx):
        assert (rows[k] == lv[i]).all()
        if nd:
            assert (sib[k] == O.merkle_proof(dg_h, n_leaves, cap_height, i)).all(), i
        assert O.merkle_verify(rows[k], sib[k], i, cap_height, n_leaves, cp_h), i


def test_openings_from_fused_commit_lde(Z):
    from gpu_util import empty, host
    lg_n, r, C, h = 11, 3, 40, 4
    N = 1 << (lg_n + r)
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, 12)
    lde, leaves, dg, cp = empty(C * N), empty(C * N), empty(2 * (N - 16) * 4), empty(64)
    Z.commit(0, coeffs, lde, leaves, dg, cp, lg_n, r, C, h)
    idx = np.random.default_rng(3).integers(0, N, 28).tolist()
    sib, rows = Z.merkle_proofs(0, dg, lde, N, C, h, idx, col_stride=N)
    sib2, rows2 = Z.merkle_proofs(0, dg, leaves, N, C, h, idx)
    assert (sib == sib2).all() and (rows == rows2).all()
    cp_h = host(cp)
    for k, i in enumerate(idx):
        assert O.merkle_verify(rows[k], sib[k], i, h, N, cp_h), i
