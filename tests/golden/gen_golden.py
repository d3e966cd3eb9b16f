"""Writes tests/golden/merkle_ref.npz (Poseidon) and
tests/golden/merkle_ref_hashtypes.npz (Keccak and Poseidon2 through the same
Merkle entry point): digests and caps computed by the reference's own CPU
Merkle (okx/zeknox native/merkle/merkle.cpp with native/poseidon/poseidon.cpp,
native/keccak/keccak.cu and native/poseidon2/poseidon2.cu, compiled from
source into oracle/_ref by oracle/ref.mk) over seeded leaves from the
oracle's input generator.

Run in the build container (where /root/reference exists):
    python tests/golden/gen_golden.py
The fixture is data only: caps, SHA-256 of the digest buffer, and the full
digest buffer for small trees (inputs are regenerated from the stored seed)."""
import hashlib
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "oracle"))
import oracle as O  # noqa: E402

SHAPES = {  # tag: (n_leaves, leaf_size, cap_height, seed)
    "c1": (1 << 15, 16, 4, 1),           # the 16 x 2^15 commit's tree (cap_height 4)
    "caps_eq_leaves": (64, 7, 6, 21),
    "subtree2": (64, 7, 5, 22),
    "leaf135": (2048, 135, 4, 23),
    "leaf400": (256, 400, 2, 24),
    "short_leaf": (512, 3, 1, 25),
    "rate_exact": (1024, 8, 3, 26),
    "single_subtree": (4096, 12, 0, 27),
}


HASHTYPE_SHAPES = {  # tag: (hash_type, n_leaves, leaf_size, cap_height, seed)
    "keccak_a": (1, 1024, 16, 4, 31),
    "keccak_short": (1, 64, 3, 2, 32),     # leaves of fewer than 4 elements are copied
    "keccak_four": (1, 256, 4, 0, 33),     # 4 elements are hashed (Poseidon copies 4)
    "keccak_multiblock": (1, 128, 40, 3, 34),  # 320 bytes: three Keccak blocks
    "keccak_subtree2": (1, 32, 9, 4, 35),
    "poseidon2_a": (3, 1024, 16, 4, 41),
    "poseidon2_leaf135": (3, 256, 135, 2, 42),
    "poseidon2_caps_eq_leaves": (3, 16, 7, 4, 43),
    "poseidon2_subtree2": (3, 64, 5, 5, 44),
}


def hashtypes():
    out = {}
    for tag, (ht, n, ls, h, seed) in HASHTYPE_SHAPES.items():
        leaves = O.random(n * ls, seed)
        dg, cp = O.ref_merkle(leaves, n, ls, h, ht)
        out["shape_" + tag] = np.array([ht, n, ls, h, seed], dtype=np.uint64)
        out["cap_" + tag] = cp
        out["digests_" + tag] = dg
    np.savez_compressed(os.path.join(HERE, "merkle_ref_hashtypes.npz"), **out)
    print("wrote", os.path.join(HERE, "merkle_ref_hashtypes.npz"))


def main():
    O.build()
    if O.ref_lib() is None:
        raise SystemExit("oracle/_ref/libzkref.so is not built (needs /root/reference)")
    out = {}
    for tag, (n, ls, h, seed) in SHAPES.items():
        leaves = O.random(n * ls, seed)
        dg, cp = O.ref_merkle(leaves, n, ls, h)
        out["shape_" + tag] = np.array([n, ls, h, seed], dtype=np.uint64)
        out["cap_" + tag] = cp
        out["sha256_" + tag] = np.frombuffer(hashlib.sha256(dg.tobytes()).digest(), dtype=np.uint8)
        if dg.size <= 4 * 4096:  # full digests only for small trees
            out["digests_" + tag] = dg
    if "--hashtypes-only" not in sys.argv:
        np.savez_compressed(os.path.join(HERE, "merkle_ref.npz"), **out)
        print("wrote", os.path.join(HERE, "merkle_ref.npz"))
    hashtypes()


if __name__ == "__main__":
    main()
