"""ctypes binding of the CPU oracle (oracle/liboracle.so) and of the
reference's own CPU Merkle built from source (oracle/_ref/libzkref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and bench.py's
CPU-baseline leg, as the checker.  Nothing in zeknox_amd/ imports this.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
P = 0xFFFFFFFF00000001
_lib = None
_ref = None


def build():
    """Compiles liboracle.so (and the reference CPU Merkle when /root/reference exists)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    if os.path.isdir("/root/reference/native"):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "ref.mk")], check=True)


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        u64, u32, v = ctypes.c_uint64, ctypes.c_uint32, ctypes.c_void_p
        for name, res, args in [
            ("or_gl_add", u64, [u64, u64]), ("or_gl_sub", u64, [u64, u64]), ("or_gl_mul", u64, [u64, u64]),
            ("or_gl_pow", u64, [u64, u64]), ("or_gl_inv", u64, [u64]), ("or_gl_reduce128", u64, [u64, u64]),
            ("or_omega", u64, [u32]), ("or_omega_inv", u64, [u32]),
            ("or_poseidon_permute", None, [v]), ("or_hash_one", None, [v, u64, v]), ("or_hash_two", None, [v, v, v]),
            ("or_fill_digests_linear", ctypes.c_int, [v, v, v, u64, u64, u64, u64, u64]),
            ("or_ntt", None, [v, u32, u64, ctypes.c_int, ctypes.c_int, u64]),
            ("or_lde", None, [v, v, u32, u32, u64, ctypes.c_int, ctypes.c_int, u64]),
            ("or_dft_naive", None, [v, v, u32, ctypes.c_int, u64]),
            ("or_poly_eval", u64, [v, u64, u64]),
            ("or_merkle_proof", ctypes.c_int, [v, u64, u64, u64, v]),
            ("or_poseidon2_permute", None, [v]), ("or_p2_hash_one", None, [v, u64, v]),
            ("or_p2_hash_two", None, [v, v, v]), ("or_keccak_f", None, [v]), ("or_keccak256", None, [v, u64, v]),
            ("or_keccak_hash_one", None, [v, u64, v]), ("or_keccak_hash_two", None, [v, v, v]),
            ("or_fill_digests_linear_ht", ctypes.c_int, [v, v, v, u64, u64, u64, u64, u64, ctypes.c_int]),
            ("or_transpose_rev", None, [v, v, u32, u64]),
            ("or_fill_random", None, [v, u64, u64]), ("or_splitmix64", u64, [u64]),
            ("or_set_threads", None, [ctypes.c_int]), ("or_get_threads", ctypes.c_int, []),
        ]:
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        _lib = L
    return _lib


def ref_lib():
    """The reference's CPU Merkle (fill_digests_buf_linear_cpu and PoseidonHasher::cpu_hash_*),
    compiled from /root/reference by ref.mk; None where it is not built."""
    global _ref
    if _ref is None:
        path = os.path.join(HERE, "_ref", "libzkref.so")
        if not os.path.exists(path):
            return None
        R = ctypes.CDLL(path)
        v, u64 = ctypes.c_void_p, ctypes.c_uint64
        R.fill_digests_buf_linear_cpu.restype = None
        R.fill_digests_buf_linear_cpu.argtypes = [v, v, v] + [u64] * 6
        R._ZN14PoseidonHasher12cpu_hash_oneEPmmS0_.argtypes = [v, u64, v]
        R._ZN14PoseidonHasher12cpu_hash_twoEPmS0_S0_.argtypes = [v, v, v]
        _ref = R
    return _ref


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def u64arr(x):
    return np.ascontiguousarray(np.asarray(x, dtype=np.uint64))


def random(n, seed):
    out = np.empty(n, dtype=np.uint64)
    lib().or_fill_random(_p(out), n, seed)
    return out


def hash_one(x):
    x = u64arr(x)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_hash_one(_p(x) if x.size else None, x.size, _p(out))
    return out


def hash_two(a, b):
    a, b = u64arr(a), u64arr(b)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_hash_two(_p(a), _p(b), _p(out))
    return out


def permute(state):
    s = u64arr(state).copy()
    lib().or_poseidon_permute(_p(s))
    return s


HASH_POSEIDON, HASH_KECCAK, HASH_POSEIDON2 = 0, 1, 3  # HashType (native/merkle/merkle.h:21-28)


def merkle(leaves, n_leaves, leaf_size, cap_height, hash_type=HASH_POSEIDON):
    """(digests, caps) of the linear-layout forest over row-major leaves."""
    leaves = u64arr(leaves)
    n_caps = 1 << cap_height
    nd = 0 if n_caps == n_leaves else 2 * (n_leaves - n_caps)
    dg = np.zeros(max(nd, 1) * 4, dtype=np.uint64)
    cp = np.zeros(n_caps * 4, dtype=np.uint64)
    rc = lib().or_fill_digests_linear_ht(_p(dg), _p(cp), _p(leaves), nd, n_caps, n_leaves, leaf_size, cap_height,
                                         hash_type)
    if rc != 0:
        raise ValueError("invalid tree shape or hash type")
    return dg[: nd * 4], cp


def ref_merkle(leaves, n_leaves, leaf_size, cap_height, hash_type=HASH_POSEIDON):
    """Same as merkle() through the reference's compiled CPU path (its output
    buffers start zeroed, as here)."""
    R = ref_lib()
    leaves = u64arr(leaves)
    n_caps = 1 << cap_height
    nd = 0 if n_caps == n_leaves else 2 * (n_leaves - n_caps)
    dg = np.zeros(max(nd, 1) * 4, dtype=np.uint64)
    cp = np.zeros(n_caps * 4, dtype=np.uint64)
    R.fill_digests_buf_linear_cpu(_p(dg), _p(cp), _p(leaves), nd, n_caps, n_leaves, leaf_size, cap_height,
                                  hash_type)
    return dg[: nd * 4], cp


def p2_permute(state):
    s = u64arr(state).copy()
    lib().or_poseidon2_permute(_p(s))
    return s


def p2_hash_one(x):
    x = u64arr(x)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_p2_hash_one(_p(x) if x.size else None, x.size, _p(out))
    return out


def p2_hash_two(a, b):
    out = np.zeros(4, dtype=np.uint64)
    lib().or_p2_hash_two(_p(u64arr(a)), _p(u64arr(b)), _p(out))
    return out


def keccak256(data: bytes):
    buf = np.frombuffer(bytes(data) + b"\0", dtype=np.uint8)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_keccak256(_p(buf), len(data), _p(out))
    return out.tobytes()


def keccak_hash_one(x):
    x = u64arr(x)
    out = np.zeros(4, dtype=np.uint64)
    lib().or_keccak_hash_one(_p(x) if x.size else None, x.size, _p(out))
    return out


def keccak_hash_two(a, b):
    out = np.zeros(4, dtype=np.uint64)
    lib().or_keccak_hash_two(_p(u64arr(a)), _p(u64arr(b)), _p(out))
    return out


def ntt(x, lg, batches, inverse=False, with_coset=False, coset_gen=7):
    a = u64arr(x).copy()
    lib().or_ntt(_p(a), lg, batches, int(inverse), int(with_coset), coset_gen)
    return a


def lde(x, lg_n, rate_bits, batches, inverse=False, with_coset=True, coset_gen=7):
    x = u64arr(x)
    out = np.empty(batches << (lg_n + rate_bits), dtype=np.uint64)
    lib().or_lde(_p(out), _p(x), lg_n, rate_bits, batches, int(inverse), int(with_coset), coset_gen)
    return out


def dft_naive(x, lg, inverse=False, shift=1):
    x = u64arr(x)
    out = np.empty(1 << lg, dtype=np.uint64)
    lib().or_dft_naive(_p(out), _p(x), lg, int(inverse), shift)
    return out


def merkle_proof(digests, n_leaves, cap_height, index):
    """Sibling digests (k x 4) of leaf `index`, leaf level first."""
    digests = u64arr(digests)
    layers = n_leaves.bit_length() - 1 - cap_height
    out = np.zeros(max(layers, 1) * 4, dtype=np.uint64)
    k = lib().or_merkle_proof(_p(digests), n_leaves, cap_height, index, _p(out))
    if k < 0:
        raise ValueError("bad Merkle shape or index")
    return out[: 4 * k].reshape(k, 4)


def merkle_verify(leaf, siblings, index, cap_height, n_leaves, cap):
    """Recompute the cap entry from a leaf row and its proof (Plonky2's
    verify_merkle_proof_to_cap, restated with the oracle's hash): at each
    level an even position is the left child."""
    layers = n_leaves.bit_length() - 1 - cap_height
    h = hash_one(leaf)
    pos = index & ((1 << layers) - 1)
    for k in range(layers):
        h = hash_two(h, siblings[k]) if pos & 1 == 0 else hash_two(siblings[k], h)
        pos >>= 1
    return (h == u64arr(cap).reshape(-1, 4)[index >> layers]).all()


def poly_eval(c, x):
    """sum_j c[j] x^j mod p (Horner)."""
    c = u64arr(c)
    return int(lib().or_poly_eval(_p(c), c.size, x % P))


def transpose_rev(x, lg_n, batches):
    x = u64arr(x)
    out = np.empty_like(x)
    lib().or_transpose_rev(_p(out), _p(x), lg_n, batches)
    return out
