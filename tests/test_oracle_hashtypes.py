"""The oracle's Keccak and Poseidon2 hashers (Merkle HashType 1 and 3), pinned
by published Keccak-256 vectors, the reference's known-answer tests
(native/tests/tests.cu:162-199 keccak_test, :405-439 poseidon2_test1), and the reference's
compiled CPU Merkle (oracle/_ref) or the fixtures it wrote
(tests/golden/merkle_ref_hashtypes.npz)."""
import os

import numpy as np
import pytest

import oracle as O

P = O.P


def test_keccak256_published_vectors():
    # Keccak-256 (original padding) as used by Ethereum
    assert O.keccak256(b"").hex() == "c5d2460186f7233c927e7db2dcc703c0e500b653ca82273b7bfad8045d85a470"
    assert O.keccak256(b"abc").hex() == "4e03657aea45a94fc7d47ba826c8d667c0d1e6e33a64a036ec44f58fa12d6c45"


def test_keccak_f_rate_boundaries_are_consistent():
    """Messages around the 136-byte rate: absorbing in one call equals the
    definition (padding 0x01 ... 0x80 in the last block)."""
    for n in (135, 136, 137, 271, 272, 300):
        msg = bytes((7 * i + 3) & 0xFF for i in range(n))
        blocks = msg + b"\x01" + b"\x00" * ((-n - 1) % 136)
        blocks = bytearray(blocks)
        blocks[-1] |= 0x80
        A = np.zeros(25, dtype=np.uint64)
        for off in range(0, len(blocks), 136):
            words = np.frombuffer(bytes(blocks[off:off + 136]), dtype="<u8")
            A[:17] ^= words
            O.lib().or_keccak_f(A.ctypes.data)
        assert A[:4].tobytes() == O.keccak256(msg), n


KECCAK_DATA = [13421290117754017454, 7401888676587830362, 15316685236050041751, 13588825262671526271,
               13421290117754017454, 7401888676587830362]
KECCAK_EXPECTED = {  # native/tests/tests.cu:166-174
    1: [13421290117754017454, 0, 0, 0],
    2: [13421290117754017454, 7401888676587830362, 0, 0],
    3: [13421290117754017454, 7401888676587830362, 15316685236050041751, 0],
    4: [9981707860959651334, 16351366398560378420, 4283762868800363615, 101],
    5: [708367124667950404, 17208681281141108820, 8334320481120086961, 134],
    6: [16109761546392287110, 4918745475135463511, 17110319063854316944, 103],
}


@pytest.mark.parametrize("n", sorted(KECCAK_EXPECTED))
def test_keccak_hasher_reference_kat(n):
    assert O.keccak_hash_one(KECCAK_DATA[:n]).tolist() == KECCAK_EXPECTED[n]


def test_poseidon2_reference_kat():
    inp = [0, 1]
    for _ in range(10):
        inp.append(inp[-2] + inp[-1])
    # native/tests/tests.cu:423-426
    assert O.p2_hash_one(inp).tolist() == [0x133a03eca11d93fb, 0x5365414fb618f58d, 0xfa49f50f3a2ba2e5,
                                           0xd16e53672c9832a4]


def test_hashtype_trees_equal_reference_fixtures():
    path = os.path.join(os.path.dirname(__file__), "golden", "merkle_ref_hashtypes.npz")
    g = np.load(path)
    tags = [k[len("shape_"):] for k in g.files if k.startswith("shape_")]
    assert len(tags) >= 9
    for tag in tags:
        ht, n, ls, h, seed = (int(v) for v in g["shape_" + tag])
        dg, cp = O.merkle(O.random(n * ls, seed), n, ls, h, ht)
        assert (cp == g["cap_" + tag]).all(), tag
        assert (dg == g["digests_" + tag]).all(), tag


@pytest.mark.parametrize("ht", [O.HASH_KECCAK, O.HASH_POSEIDON2])
def test_hashtype_trees_equal_compiled_reference(ht):
    if O.ref_lib() is None:
        pytest.skip("oracle/_ref (the reference's compiled CPU Merkle) is not built here")
    for n, ls, h in [(8, 7, 2), (64, 9, 0), (256, 5, 3), (16, 3, 1), (32, 40, 2), (2, 6, 0), (128, 4, 7)]:
        lv = O.random(n * ls, 100 + n + ls + ht)
        a, b = O.merkle(lv, n, ls, h, ht), O.ref_merkle(lv, n, ls, h, ht)
        assert (a[0] == b[0]).all() and (a[1] == b[1]).all(), (n, ls, h)
