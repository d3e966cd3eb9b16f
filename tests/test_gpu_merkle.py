"""GPU Merkle (fill_digests_buf_linear_gpu_with_gpu_ptr) through the C ABI,
bit-exact against the oracle, the reference KATs and the golden fixtures
written by the reference's CPU Merkle."""
import hashlib
import os

import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    return zeknox_amd


def gpu_merkle(Z, leaves, n_leaves, leaf_size, cap_height, entry="gpu"):
    from gpu_util import dev, empty, host
    n_caps = 1 << cap_height
    nd = 0 if n_caps == n_leaves else 2 * (n_leaves - n_caps)
    dl = dev(leaves)
    dd = empty(nd * 4)
    dc = empty(n_caps * 4)
    if entry == "gpu":
        Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, dl, nd, n_caps, n_leaves, leaf_size, cap_height)
    else:
        Z.fill_digests_buf_linear_multigpu_with_gpu_ptr(dd, dc, dl, nd, n_caps, n_leaves, leaf_size, cap_height)
    return host(dd, nd * 4), host(dc)


def test_go_cap_kat(Z):
    from test_oracle_poseidon import GO_CASES, GO_LEAF
    for n_leaves, n_caps, expected in GO_CASES:
        leaves = np.tile(np.array(GO_LEAF, dtype=np.uint64), n_leaves)
        _, cp = gpu_merkle(Z, leaves, n_leaves, 7, int(np.log2(n_caps)))
        assert (cp.reshape(-1, 4) == np.array(expected, dtype=np.uint64)).all()


def test_hash_one_kat_through_caps(Z):
    # caps == leaves hashes each leaf into its cap slot: leaf k = KAT leaf prefix, zero padded
    from test_oracle_poseidon import KAT_HASH_ONE, KAT_LEAF
    for k in range(1, 10):
        leaves = np.array(KAT_LEAF[:k], dtype=np.uint64)
        _, cp = gpu_merkle(Z, leaves, 1, k, 0)
        assert cp.tolist() == KAT_HASH_ONE[k], k


@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [
    (8, 7, 3), (8, 7, 2), (16, 7, 1), (4, 6, 2), (32, 1, 2), (32, 4, 0), (64, 5, 3), (128, 8, 4),
    (256, 9, 2), (1024, 16, 4), (512, 135, 4), (64, 400, 2), (2, 12, 0), (1, 12, 0), (256, 24, 7),
    (4096, 17, 0), (2048, 3, 10), (131072, 5, 1), (65536, 9, 6),
])
def test_random_forest_equals_oracle(Z, n_leaves, leaf_size, cap_height):
    leaves = O.random(n_leaves * leaf_size, 300 + n_leaves + leaf_size + cap_height)
    dg, cp = gpu_merkle(Z, leaves, n_leaves, leaf_size, cap_height)
    odg, ocp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    assert (cp == ocp).all()
    assert (dg == odg).all()


@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [(512, 12, 2), (65536, 9, 3)])
def test_rare_case_leaves_equal_oracle(Z, n_leaves, leaf_size, cap_height):
    """Leaves whose first S-box hits the relaxed arithmetic's rare case, so the
    leaf hash takes its exact recompute path: 512 leaves use the quad-lane
    kernel, 65536 the one-lane kernel."""
    from rare_inputs import leaves_with_rare_cases
    leaves = leaves_with_rare_cases(O.random(n_leaves * leaf_size, 77 + n_leaves), n_leaves, leaf_size)
    print("rare leaves", n_leaves, leaf_size, cap_height, flush=True)
    dg, cp = gpu_merkle(Z, leaves, n_leaves, leaf_size, cap_height)
    odg, ocp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    assert (cp == ocp).all()
    assert (dg == odg).all()


def test_multigpu_entry_runs_single_gpu(Z):
    leaves = O.random(256 * 11, 333)
    dg, cp = gpu_merkle(Z, leaves, 256, 11, 3, entry="multigpu")
    odg, ocp = O.merkle(leaves, 256, 11, 3)
    assert (cp == ocp).all() and (dg == odg).all()


def test_golden_fixtures(Z):
    """Every tree of tests/golden/merkle_ref.npz (reference CPU Merkle output),
    including the 2^15 x 16, cap_height 4 tree of the 16-column commit."""
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "merkle_ref.npz"))
    for key in sorted(k for k in g.files if k.startswith("shape_")):
        tag = key[len("shape_"):]
        n, ls, h, seed = (int(v) for v in g[key])
        print("golden", tag, n, ls, h, flush=True)
        dg, cp = gpu_merkle(Z, O.random(n * ls, seed), n, ls, h)
        assert (cp == g["cap_" + tag]).all(), tag
        assert hashlib.sha256(dg.tobytes()).digest() == g["sha256_" + tag].tobytes(), tag


def test_unsupported_hash_type_leaves_buffers(Z, capfd):
    from gpu_util import dev, host
    dl = dev(O.random(64 * 8, 5))
    dc = dev(np.full(4 * 4, 7, dtype=np.uint64))
    dd = dev(np.full(2 * 60 * 4, 7, dtype=np.uint64))
    for ht in (2, 4, 9):  # Poseidon-BN128 and Monolith are not built; 9 is no hash type
        Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, dl, 120, 4, 64, 8, 2, hash_type=ht)
        assert "not implemented" in capfd.readouterr().err
        assert (host(dc) == 7).all()


@pytest.mark.parametrize("hash_type", [1, 3])
def test_other_hash_types_equal_reference_fixtures(Z, hash_type):
    """Keccak (1) and Poseidon2 (3) trees through the same entry point against
    the reference CPU Merkle's outputs (tests/golden/merkle_ref_hashtypes.npz)
    and the oracle, for row-major leaves."""
    from gpu_util import dev, empty, host
    g = np.load(os.path.join(os.path.dirname(__file__), "golden", "merkle_ref_hashtypes.npz"))
    tags = [k[len("shape_"):] for k in g.files if k.startswith("shape_") and int(g[k][0]) == hash_type]
    assert tags
    for tag in tags:
        _, n, ls, h, seed = (int(v) for v in g["shape_" + tag])
        ncaps = 1 << h
        nd = 2 * (n - ncaps) if ncaps < n else 0
        dl, dd, dc = dev(O.random(n * ls, seed)), empty(nd * 4), empty(ncaps * 4)
        Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dd, dc, dl, nd, ncaps, n, ls, h, hash_type=hash_type)
        assert (host(dc) == g["cap_" + tag]).all(), tag
        assert (host(dd, nd * 4) == g["digests_" + tag]).all(), tag
