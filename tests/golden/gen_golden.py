"""Writes tests/golden/merkle_ref.npz: digests and caps computed by the
reference's own CPU Merkle (okx/zeknox native/merkle/merkle.cpp +
native/poseidon/poseidon.cpp, compiled from source into oracle/_ref by
oracle/ref.mk) over seeded leaves from the oracle's input generator.

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
    np.savez_compressed(os.path.join(HERE, "merkle_ref.npz"), **out)
    print("wrote", os.path.join(HERE, "merkle_ref.npz"))


if __name__ == "__main__":
    main()
