"""The oracle's Poseidon sponge and Merkle layout, pinned by the reference's own
known-answer tests and by the reference's CPU Merkle compiled from source."""
import hashlib
import os

import numpy as np
import pytest

import oracle as O

# native/tests/tests.cu:279-292 (poseidon_test1): hash_one of leaf[0..k)
KAT_LEAF = [8395359103262935841, 1377884553022145855, 2370707998790318766, 3651132590097252162,
            1141848076261006345, 12736915248278257710, 9898074228282442027, 16154511938222758243,
            3651132590097252162]
KAT_HASH_ONE = {
    1: [8395359103262935841, 0, 0, 0],
    2: [8395359103262935841, 1377884553022145855, 0, 0],
    3: [8395359103262935841, 1377884553022145855, 2370707998790318766, 0],
    4: [8395359103262935841, 1377884553022145855, 2370707998790318766, 3651132590097252162],
    5: [3618821072812614426, 8353148445756493727, 4040525329700581442, 15983474240847269257],
    6: [16643938361881363776, 6653675298471110559, 12562058402463703932, 16154511938222758243],
    7: [7544909477878586743, 7431000548126831493, 17815668806142634286, 13168106265494210017],
    8: [6835933650993053111, 15978194778874965616, 2024081381896137659, 16520693669262110264],
    9: [9429914239539731992, 14881719063945231827, 15528667124986963891, 16465743531992249573],
}

# wrappers/go/lib/lib_test.go:194-214: a 7-element leaf replicated, expected cap entries
GO_LEAF = KAT_LEAF[:7]
GO_CASES = [  # (n_leaves, n_caps, expected cap digest, identical for every cap entry)
    (8, 8, [7544909477878586743, 7431000548126831493, 17815668806142634286, 13168106265494210017]),
    (8, 4, [7320113784629306061, 6779895905614599055, 11363085033710007978, 10376477785629036773]),
    (16, 2, [15643713602922591410, 15764162833178839023, 7963799186968795376, 10933634831850922421]),
]


@pytest.mark.parametrize("k", range(1, 10))
def test_hash_one_kat(k):
    assert O.hash_one(KAT_LEAF[:k]).tolist() == KAT_HASH_ONE[k]


@pytest.mark.parametrize("n_leaves,n_caps,expected", GO_CASES)
def test_go_cap_kat(n_leaves, n_caps, expected):
    leaves = np.tile(np.array(GO_LEAF, dtype=np.uint64), n_leaves)
    _, caps = O.merkle(leaves, n_leaves, 7, int(np.log2(n_caps)))
    assert (caps.reshape(-1, 4) == np.array(expected, dtype=np.uint64)).all()


def test_hash_two_is_permutation_of_concatenation():
    a, b = O.random(4, 11), O.random(4, 12)
    st = np.concatenate([a, b, np.zeros(4, dtype=np.uint64)])
    assert (O.hash_two(a, b) == O.permute(st)[:4]).all()


def test_sponge_overwrite_mode_matches_manual_absorb():
    # 9 elements: chunk 1 overwrites state[0..8), chunk 2 overwrites state[0] only
    x = O.random(9, 13)
    st = np.zeros(12, dtype=np.uint64)
    st[:8] = x[:8]
    st = O.permute(st)
    st[0] = x[8]
    st = O.permute(st)
    assert (O.hash_one(x) == st[:4]).all()


def test_naive_tree_definition_small():
    # 4 leaves, cap_height 0: digests [d0 d1 | l0 l1 l2 l3], d_i = H(d[2i+2], d[2i+3]), cap = H(d0, d1)
    leaves = O.random(4 * 7, 14).reshape(4, 7)
    dg, cap = O.merkle(leaves.ravel(), 4, 7, 0)
    dg = dg.reshape(-1, 4)
    lh = [O.hash_one(leaves[i]) for i in range(4)]
    assert all((dg[2 + i] == lh[i]).all() for i in range(4))
    assert (dg[0] == O.hash_two(lh[0], lh[1])).all() and (dg[1] == O.hash_two(lh[2], lh[3])).all()
    assert (cap == O.hash_two(dg[0], dg[1])).all()


needs_ref = pytest.mark.skipif(O.ref_lib() is None, reason="reference CPU Merkle (oracle/_ref) not built here")


@needs_ref
@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [
    (8, 7, 3), (8, 7, 2), (16, 7, 1), (4, 6, 2), (32, 1, 2), (32, 4, 0), (64, 5, 3), (128, 8, 4), (256, 9, 2),
    (64, 16, 4), (1024, 16, 4), (512, 135, 4), (64, 400, 2), (2, 12, 0), (1, 12, 0), (256, 24, 7),
])
def test_oracle_merkle_equals_reference_cpu(n_leaves, leaf_size, cap_height):
    leaves = O.random(n_leaves * leaf_size, 100 + n_leaves + leaf_size)
    dg, cp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    rdg, rcp = O.ref_merkle(leaves, n_leaves, leaf_size, cap_height)
    assert (cp == rcp).all()
    assert (dg == rdg).all()


@needs_ref
def test_oracle_hash_equals_reference_cpu_every_length():
    import ctypes
    R = O.ref_lib()
    x = O.random(40, 77)
    for n in range(0, 41):
        out = np.zeros(4, dtype=np.uint64)
        R._ZN14PoseidonHasher12cpu_hash_oneEPmmS0_(x.ctypes.data_as(ctypes.c_void_p), n,
                                                   out.ctypes.data_as(ctypes.c_void_p))
        assert (O.hash_one(x[:n]) == out).all(), n


def _golden(name):
    return os.path.join(os.path.dirname(__file__), "golden", name)


def test_oracle_matches_golden_fixtures():
    """Digests/caps written by the reference's compiled CPU Merkle (tests/golden/gen_golden.py)."""
    path = _golden("merkle_ref.npz")
    if not os.path.exists(path):
        pytest.skip("golden fixtures not generated")
    g = np.load(path)
    for key in sorted(k for k in g.files if k.startswith("shape_")):
        tag = key[len("shape_"):]
        n_leaves, leaf_size, cap_height, seed = (int(v) for v in g[key])
        leaves = O.random(n_leaves * leaf_size, seed)
        dg, cp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
        assert (cp == g["cap_" + tag]).all(), tag
        assert hashlib.sha256(dg.tobytes()).digest() == g["sha256_" + tag].tobytes(), tag
        if "digests_" + tag in g.files:
            assert (dg == g["digests_" + tag]).all(), tag


@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [(2, 7, 0), (8, 7, 2), (64, 9, 0), (256, 5, 3),
                                                           (1024, 16, 4), (16, 3, 3)])
def test_merkle_proofs_verify_to_the_cap(n_leaves, leaf_size, cap_height):
    """Proofs extracted from the linear layout recompute the cap (the tree
    itself is the reference-pinned oracle Merkle); a changed sibling or a
    wrong index does not."""
    leaves = O.random(n_leaves * leaf_size, 4242 + n_leaves)
    dg, cp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    rows = leaves.reshape(n_leaves, leaf_size)
    rng = np.random.default_rng(n_leaves)
    for idx in {0, n_leaves - 1, *rng.integers(0, n_leaves, 6).tolist()}:
        sib = O.merkle_proof(dg, n_leaves, cap_height, idx)
        assert sib.shape[0] == n_leaves.bit_length() - 1 - cap_height
        assert O.merkle_verify(rows[idx], sib, idx, cap_height, n_leaves, cp)
        if sib.shape[0]:
            bad = sib.copy()
            bad[-1, 0] ^= 1
            assert not O.merkle_verify(rows[idx], bad, idx, cap_height, n_leaves, cp)
            assert not O.merkle_verify(rows[idx], sib, idx ^ 1, cap_height, n_leaves, cp)
