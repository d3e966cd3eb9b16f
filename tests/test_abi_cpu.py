"""libzeknox.so without a GPU: it loads, exports every function declared in
include/*.h, keeps the reference's argument-error conventions, and its host
Merkle entry point (fill_digests_buf_linear_cpu) equals the oracle."""
import glob
import os
import re

import numpy as np
import pytest

import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "**", "*.h"), recursive=True):
        text = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        for m in re.finditer(r"(?:EXTERN|ZK_EXTERNC|RustError|void|bool)\s+\**\s*([a-zA-Z_]\w*)\s*\(", text):
            names.add(m.group(1))
    names.discard("RustError")
    return names


def test_library_exports_every_declared_symbol():
    import zeknox_amd as Z
    names = declared_functions()
    assert {"compute_batched_lde", "compute_batched_ntt", "compute_transpose_rev",
            "fill_digests_buf_linear_gpu_with_gpu_ptr", "init_twiddle_factors", "zk_commit"} <= names
    for n in sorted(names):
        assert hasattr(Z.lib, n), n


def test_struct_layouts():
    import zeknox_amd as Z
    import ctypes
    assert ctypes.sizeof(Z.NTTConfig) == 24
    assert Z.NTTConfig.extension_rate_bits.offset == 12 and Z.NTTConfig.salt_size.offset == 20
    assert ctypes.sizeof(Z.TransposeConfig) == 8


def test_argument_errors_need_no_device():
    import zeknox_amd as Z
    with pytest.raises(Z.ZeknoxError) as e:
        Z.lde_batch(0, 0, 0, 0, Z.NTTConfig(batches=1, extension_rate_bits=3))
    assert e.value.code == 1 and e.value.message == "invalid args"  # ntt.cu:478-481
    with pytest.raises(Z.ZeknoxError) as e:
        Z.lde_batch(0, 0, 0, 10, Z.NTTConfig(batches=1, extension_rate_bits=0))
    assert e.value.code == 1
    with pytest.raises(Z.ZeknoxError) as e:
        Z.ntt_batch(0, 0, 0, Z.NTTConfig())
    assert e.value.code == 1 and e.value.message is None  # ntt.cu:255-256
    for shift in (0, O.P):  # a zero coset shift is rejected before any device work
        with pytest.raises(Z.ZeknoxError) as e:
            Z.commit(0, 1 << 20, 2 << 20, None, 3 << 20, 4 << 20, 4, 1, 1, 0, shift=shift)
        assert e.value.code == 1 and e.value.message == "invalid args"
    # the allocation-free commit never creates a device context (that
    # allocates its tables): on a device nothing has initialised it refuses
    with pytest.raises(Z.ZeknoxError) as e:
        Z.commit_noalloc(0, 1 << 20, 2 << 20, None, 3 << 20, 4 << 20, 4, 1, 1, 0)
    assert e.value.code == 1 and "not initialised" in e.value.message


def test_device_queries_without_gpu():
    import zeknox_amd as Z
    n = Z.get_number_of_gpus_rs()
    assert isinstance(n, int) and n >= 0
    assert Z.cuda_available() == (n > 0)


@pytest.mark.parametrize("n_leaves,leaf_size,cap_height", [
    (8, 7, 3), (8, 7, 2), (16, 7, 1), (64, 1, 2), (128, 4, 3), (256, 9, 2), (1024, 16, 4), (512, 135, 4),
    (64, 400, 2), (2, 12, 0), (4096, 8, 0),
])
def test_host_merkle_entry_point_equals_oracle(n_leaves, leaf_size, cap_height):
    import zeknox_amd as Z
    leaves = O.random(n_leaves * leaf_size, 200 + n_leaves + leaf_size)
    n_caps = 1 << cap_height
    nd = 0 if n_caps == n_leaves else 2 * (n_leaves - n_caps)
    dg = np.zeros(max(nd, 1) * 4, dtype=np.uint64)
    cp = np.zeros(n_caps * 4, dtype=np.uint64)
    Z.fill_digests_buf_linear_cpu(dg, cp, leaves, nd, n_caps, n_leaves, leaf_size, cap_height)
    odg, ocp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    assert (cp == ocp).all()
    assert (dg[: nd * 4] == odg).all()


def test_host_merkle_rare_case_leaves_equal_oracle():
    """Leaves whose first S-box hits the relaxed arithmetic's rare case: the
    host Poseidon must take its exact recompute path and still match."""
    import zeknox_amd as Z
    from rare_inputs import leaves_with_rare_cases
    n_leaves, leaf_size, cap_height = 256, 12, 2
    leaves = leaves_with_rare_cases(O.random(n_leaves * leaf_size, 78), n_leaves, leaf_size)
    n_caps = 1 << cap_height
    nd = 2 * (n_leaves - n_caps)
    dg = np.zeros(nd * 4, dtype=np.uint64)
    cp = np.zeros(n_caps * 4, dtype=np.uint64)
    Z.fill_digests_buf_linear_cpu(dg, cp, leaves, nd, n_caps, n_leaves, leaf_size, cap_height)
    odg, ocp = O.merkle(leaves, n_leaves, leaf_size, cap_height)
    assert (cp == ocp).all()
    assert (dg == odg).all()


def test_host_merkle_go_kat():
    import zeknox_amd as Z
    from test_oracle_poseidon import GO_CASES, GO_LEAF
    for n_leaves, n_caps, expected in GO_CASES:
        leaves = np.tile(np.array(GO_LEAF, dtype=np.uint64), n_leaves)
        h = int(np.log2(n_caps))
        nd = 2 * (n_leaves - n_caps)
        dg = np.zeros(max(nd, 1) * 4, dtype=np.uint64)
        cp = np.zeros(n_caps * 4, dtype=np.uint64)
        Z.fill_digests_buf_linear_cpu(dg, cp, leaves, nd, n_caps, n_leaves, 7, h)
        assert (cp.reshape(-1, 4) == np.array(expected, dtype=np.uint64)).all()


def test_host_merkle_rejects_bad_shape(capfd):
    import zeknox_amd as Z
    cp = np.full(8, 9, dtype=np.uint64)
    Z.fill_digests_buf_linear_cpu(None, cp, np.zeros(24, dtype=np.uint64), 5, 2, 3, 8, 1)  # 3 leaves
    assert "power of two" in capfd.readouterr().err
    assert (cp == 9).all()


def test_batch_rejects_invalid_jobs_before_device_work():
    import zeknox_amd as Z
    good = {"coeffs": 1 << 20, "lde": 2 << 20, "leaves": None, "digests": 3 << 20, "cap": 4 << 20, "lg_n": 4,
            "rate_bits": 1, "ncols": 3, "cap_height": 2}
    for bad in ({"lg_n": 0}, {"ncols": 0}, {"cap": None}, {"cap_height": 6}, {"digests": None},
                {"lg_n": 30, "rate_bits": 2}):
        with pytest.raises(Z.ZeknoxError) as e:
            Z.batch_commit(0, [good, dict(good, **bad)])
        assert e.value.code == 1 and e.value.message == "invalid args", bad


def test_overlapping_commit_buffers_are_refused_before_device_work():
    """An LDE written over its own coefficients (or leaves / digests / cap
    over another output) is refused with code 1 by zk_commit and
    zk_batch_commit; the addresses are never dereferenced."""
    import zeknox_amd as Z
    lg_n, r, C, h = 4, 1, 3, 2
    N = 1 << (lg_n + r)
    base = {"coeffs": 1 << 20, "lde": 2 << 20, "leaves": None, "digests": 3 << 20, "cap": 4 << 20,
            "lg_n": lg_n, "rate_bits": r, "ncols": C, "cap_height": h}
    cases = [{"lde": 1 << 20},                            # in-place LDE
             {"lde": (1 << 20) + 8 * (C << lg_n) - 8},    # last coefficient under the LDE's first
             {"leaves": (2 << 20) + 8 * C * N - 8},       # leaves over the LDE's tail
             {"cap": (3 << 20) + 32 * 2 * (N - (1 << h)) - 32}]  # cap over the last digest
    for bad in cases:
        j = dict(base, **bad)
        with pytest.raises(Z.ZeknoxError) as e:
            Z.commit(0, j["coeffs"], j["lde"], j["leaves"], j["digests"], j["cap"], lg_n, r, C, h)
        assert e.value.code == 1 and "overlapping" in e.value.message, bad
        with pytest.raises(Z.ZeknoxError) as e:
            Z.batch_commit(0, [base, j])
        assert e.value.code == 1 and "overlapping" in e.value.message, bad
    # adjacent (touching, not overlapping) ranges pass validation and only then
    # fail on the missing device, with a runtime (negative) code
    j = dict(base, lde=(1 << 20) + 8 * (C << lg_n))
    if Z.get_number_of_gpus_rs() == 0:
        with pytest.raises(Z.ZeknoxError) as e:
            Z.commit(0, j["coeffs"], j["lde"], None, j["digests"], j["cap"], lg_n, r, C, h)
        assert e.value.code < 0


def test_server_entry_points_without_gpu():
    import ctypes
    import zeknox_amd as Z
    if Z.get_number_of_gpus_rs() > 0:
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    e = Z.lib.zk_server_create(0, 1 << 20, 4, ctypes.byref(h))
    assert e.code != 0 and not h.value
    Z._check(Z.lib.zk_server_destroy(None))  # NULL is a no-op


def test_merkle_proofs_argument_errors_need_no_device():
    import ctypes
    import zeknox_amd as Z
    idx = np.array([3, 16], dtype=np.uint64)  # 16 is out of range for 16 leaves
    e = Z.lib.zk_merkle_proofs(0, 8, 8, 0, 16, 4, 2, idx.ctypes.data, 2, 8, 8)
    msg = ctypes.string_at(e.message).decode()
    Z._libc.free(e.message)
    assert e.code == 1 and "out of range" in msg
    e = Z.lib.zk_merkle_proofs(0, 8, 8, 0, 12, 4, 2, idx.ctypes.data, 1, 8, 8)  # 12 leaves
    assert e.code == 1
    Z._libc.free(e.message)


@pytest.mark.parametrize("hash_type", [1, 3])
def test_host_merkle_other_hash_types_equal_reference_fixtures(hash_type):
    """fill_digests_buf_linear_cpu with Keccak (1) / Poseidon2 (3) against the
    reference CPU Merkle's outputs and the oracle."""
    import zeknox_amd as Z
    g = np.load(os.path.join(ROOT, "tests", "golden", "merkle_ref_hashtypes.npz"))
    for k in g.files:
        if not k.startswith("shape_") or int(g[k][0]) != hash_type:
            continue
        tag = k[len("shape_"):]
        _, n, ls, h, seed = (int(v) for v in g[k])
        ncaps = 1 << h
        nd = 2 * (n - ncaps) if ncaps < n else 0
        leaves = O.random(n * ls, seed)
        dg = np.zeros(max(nd, 1) * 4, dtype=np.uint64)
        cp = np.zeros(ncaps * 4, dtype=np.uint64)
        Z.fill_digests_buf_linear_cpu(dg, cp, leaves, nd, ncaps, n, ls, h, hash_type=hash_type)
        assert (cp == g["cap_" + tag]).all(), tag
        assert (dg[: nd * 4] == g["digests_" + tag]).all(), tag


def test_ntt_tune_table_set_get_and_validation():
    """The NTT plan table (zk_ntt_tune_set/get) needs no device: it accepts
    runnable plans only and restores the default plan on an empty set."""
    import zeknox_amd as Z
    loaded = Z.ntt_tune_get(19)  # from the tuning table beside the library, if it has one
    try:
        Z.ntt_tune_set(19, (7, 6, 6), (-1, 2, 3), (1, 1, 1))
        assert Z.ntt_tune_get(19) == ((7, 6, 6), (-1, 2, 3), (1, 1, 1))
        Z.ntt_tune_set(19)
        assert Z.ntt_tune_get(19) is None
        for bits, gbits in [((10, 10), (-1, -1)), ((11, 8), (-1, -1)), ((10, 9), (7, -1)), ((0, 19), (-1, -1))]:
            with pytest.raises(Z.ZeknoxError) as e:
                Z.ntt_tune_set(19, bits, gbits)
            assert e.value.code == 1
        with pytest.raises(Z.ZeknoxError):
            Z.ntt_tune_set(31, (10, 10, 11))
        assert Z.ntt_tune_get(19) is None
    finally:
        Z.ntt_tune_set(19, *loaded) if loaded else Z.ntt_tune_set(19)
