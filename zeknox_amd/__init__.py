"""Python binding of libzeknox.so, the MI355X build of the zeknox C ABI.

Mirrors the Rust wrapper of the reference (okx/zeknox wrappers/rust/src/lib.rs:
`lde_batch`, `ntt_batch`, `intt_batch`, `transpose_rev_batch`,
`init_twiddle_factors_rs`, `init_coset_rs`, `init_cuda_rs`, ... and the raw
`fill_digests_*` functions) with the same argument meaning; like the Rust
wrapper, a non-zero RustError raises (ZeknoxError instead of a panic).

Pointers are plain integers (device or host addresses); `ptr(x)` takes them
from torch tensors or numpy arrays.  There is no fallback: importing this
module fails if libzeknox.so has not been built (python zeknox_amd/build.py).
"""
# PyTorch bundles a HIP runtime with the same soname as ROCm's
# (libamdhip64.so.7).  Loading torch first makes libzeknox.so bind to that
# copy, so the process holds one runtime; loaded the other way round, the
# second runtime to initialise can find no device.
try:
    import torch  # noqa: F401
except ImportError:
    pass
import ctypes
import os

__all__ = [
    "ZeknoxError", "NTTConfig", "TransposeConfig", "FORWARD", "INVERSE", "HashPoseidon",
    "lib", "ptr", "get_number_of_gpus_rs", "list_devices_info_rs", "init_cuda_rs", "init_cuda_degree_rs",
    "init_twiddle_factors_rs", "init_coset_rs", "lde_batch", "lde_batch_multi_gpu", "ntt_batch", "intt_batch",
    "transpose_rev_batch", "fill_digests_buf_linear_gpu_with_gpu_ptr", "fill_digests_buf_linear_multigpu_with_gpu_ptr",
    "fill_digests_buf_linear_cpu", "commit", "commit_noalloc", "fill_random", "profile_enable", "profile_reset", "profile_read",
    "cuda_available", "LIB_PATH", "P", "CommitJob", "batch_commit", "batch_clear", "BATCH_GRAPH",
]

P = 0xFFFFFFFF00000001
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libzeknox.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("zeknox_amd: %s is missing; build it with `python zeknox_amd/build.py`" % LIB_PATH)
lib = ctypes.CDLL(LIB_PATH)

FORWARD, INVERSE = 0, 1          # NTT_Direction (native/ntt/ntt.h:27-31)
HashPoseidon = 0                 # HashType (native/merkle/merkle.h:21-28)


class RustError(ctypes.Structure):
    _fields_ = [("code", ctypes.c_int), ("message", ctypes.c_void_p)]


class NTTConfig(ctypes.Structure):
    """NTT_Config (native/ntt/ntt.h:43-55); defaults as wrappers/rust/src/types.rs:47-61."""
    _fields_ = [
        ("batches", ctypes.c_uint32),
        ("order", ctypes.c_int),
        ("ntt_type", ctypes.c_int),
        ("extension_rate_bits", ctypes.c_uint32),
        ("are_inputs_on_device", ctypes.c_bool),
        ("are_outputs_on_device", ctypes.c_bool),
        ("with_coset", ctypes.c_bool),
        ("is_multi_gpu", ctypes.c_bool),
        ("salt_size", ctypes.c_uint32),
    ]

    def __init__(self, **kw):
        super().__init__(batches=1, order=0, ntt_type=0, extension_rate_bits=0, are_inputs_on_device=False,
                         are_outputs_on_device=False, with_coset=False, is_multi_gpu=False, salt_size=0)
        for k, v in kw.items():
            setattr(self, k, v)


class TransposeConfig(ctypes.Structure):
    """NTT_TransposeConfig (native/ntt/ntt.h:57-62)."""
    _fields_ = [("batches", ctypes.c_uint32), ("are_inputs_on_device", ctypes.c_bool),
                ("are_outputs_on_device", ctypes.c_bool)]

    def __init__(self, **kw):
        super().__init__(batches=1, are_inputs_on_device=False, are_outputs_on_device=False)
        for k, v in kw.items():
            setattr(self, k, v)


assert ctypes.sizeof(NTTConfig) == 24 and ctypes.sizeof(TransposeConfig) == 8

_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]

_v, _sz, _u32, _u64 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint32, ctypes.c_uint64


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype, f.argtypes = res, args
    return f


_sig("get_number_of_gpus", RustError, [ctypes.POINTER(ctypes.c_size_t)])
_sig("list_devices_info", RustError, [])
_sig("init_cuda", None, [])
_sig("init_cuda_degree", None, [_u32])
_sig("init_cuda_degree_and_generator", None, [_u32, _u64])
_sig("init_twiddle_factors", RustError, [_sz, _sz])
_sig("init_coset", RustError, [_sz, _sz, _u64])
_sig("compute_batched_ntt", RustError, [_sz, _v, _u32, ctypes.c_int, NTTConfig])
_sig("compute_batched_lde", RustError, [_sz, _v, _v, _u32, ctypes.c_int, NTTConfig])
_sig("compute_batched_lde_multi_gpu", RustError, [_v, _v, _u32, ctypes.c_int, NTTConfig, _u32, _sz, _sz])
_sig("compute_transpose_rev", RustError, [_sz, _v, _v, _u32, TransposeConfig])
_sig("fill_digests_buf_linear_gpu_with_gpu_ptr", None, [_v, _v, _v] + [_u64] * 7)
_sig("fill_digests_buf_linear_multigpu_with_gpu_ptr", None, [_v, _v, _v] + [_u64] * 6)
_sig("fill_digests_buf_linear_cpu", None, [_v, _v, _v] + [_u64] * 6)
_sig("cuda_available", ctypes.c_bool, [])
_sig("zk_commit", RustError, [_sz, _v, _v, _v, _v, _v, _u32, _u32, _u32, _u32, _u64])
_sig("zk_commit_noalloc", RustError, [_sz, _v, _v, _v, _v, _v, _u32, _u32, _u32, _u32, _u64])
_sig("zk_fill_random", RustError, [_sz, _v, _u64, _u64])
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i32p = ctypes.POINTER(ctypes.c_int32)
_sig("zk_ntt_tune_set", RustError, [_u32, _u32, _u32p, _i32p, _u32p])
_sig("zk_ntt_tune_get", RustError, [_u32, _u32p, _u32p, _i32p, _u32p])


def ntt_tune_set(lg, bits=(), gbits=None, kernel16=None):
    """Plan of the 2^lg transforms (include/zeknox_ext.h zk_ntt_tune_set):
    pass bits, groups per workgroup (log2, -1 auto), 16-element kernel flags;
    bits=() restores the default plan."""
    n = len(bits)
    b = (ctypes.c_uint32 * 3)(*(list(bits) + [0] * (3 - n)))
    g = (ctypes.c_int32 * 3)(*(list(gbits if gbits is not None else [-1] * n) + [-1] * (3 - n)))
    k = (ctypes.c_uint32 * 3)(*(list(kernel16 if kernel16 is not None else [1] * n) + [0] * (3 - n)))
    _check(lib.zk_ntt_tune_set(lg, n, b, g, k))


def ntt_tune_get(lg):
    """-> (bits, gbits, kernel16) of the tuned plan of 2^lg, or None."""
    n = ctypes.c_uint32()
    b, g, k = (ctypes.c_uint32 * 3)(), (ctypes.c_int32 * 3)(), (ctypes.c_uint32 * 3)()
    _check(lib.zk_ntt_tune_get(lg, ctypes.byref(n), b, g, k))
    if not n.value:
        return None
    return tuple(b[:n.value]), tuple(g[:n.value]), tuple(k[:n.value])


__all__ += ["ntt_tune_set", "ntt_tune_get"]


class CommitJob(ctypes.Structure):
    """zk_commit_job (include/zeknox_ext.h)."""
    _fields_ = [("coeffs", ctypes.c_void_p), ("lde", ctypes.c_void_p), ("leaves", ctypes.c_void_p),
                ("digests", ctypes.c_void_p), ("cap", ctypes.c_void_p), ("lg_n", ctypes.c_uint32),
                ("rate_bits", ctypes.c_uint32), ("ncols", ctypes.c_uint32), ("cap_height", ctypes.c_uint32)]


BATCH_GRAPH = 1
_sig("zk_batch_commit", RustError, [_sz, ctypes.POINTER(CommitJob), _u32, _u32])
_sig("zk_batch_clear", RustError, [_sz])
_sig("zk_profile_enable", None, [ctypes.c_int])
_sig("zk_profile_reset", None, [])
_sig("zk_profile_get", RustError, [ctypes.c_int, ctypes.c_char_p, _sz, ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_uint64)])


class ZeknoxError(RuntimeError):
    def __init__(self, code, message):
        super().__init__("zeknox error %d: %s" % (code, message))
        self.code = code
        self.message = message


def _check(e):
    """Raise on a non-zero RustError; the message string is freed here, as the
    Rust wrapper's `From<Error> for String` does (wrappers/rust/src/error.rs:11-19)."""
    msg = None
    if e.message:
        msg = ctypes.string_at(e.message).decode(errors="replace")
        _libc.free(e.message)
    if e.code != 0:
        raise ZeknoxError(e.code, msg)


def ptr(x):
    """Address of a torch tensor / numpy array / int."""
    if x is None:
        return None
    if isinstance(x, int):
        return x
    if hasattr(x, "data_ptr"):
        return x.data_ptr()
    if hasattr(x, "ctypes"):
        return x.ctypes.data
    raise TypeError("cannot take the address of %r" % type(x))


def get_number_of_gpus_rs():
    n = ctypes.c_size_t(0)
    _check(lib.get_number_of_gpus(ctypes.byref(n)))
    return n.value


def list_devices_info_rs():
    _check(lib.list_devices_info())


def cuda_available():
    return bool(lib.cuda_available())


def init_cuda_rs():
    lib.init_cuda()


def init_cuda_degree_rs(max_degree):
    lib.init_cuda_degree(max_degree)


def init_twiddle_factors_rs(device_id, lg_n):
    _check(lib.init_twiddle_factors(device_id, lg_n))


def init_coset_rs(device_id, lg_n, coset_gen):
    _check(lib.init_coset(device_id, lg_n, coset_gen))


def lde_batch(device_id, output, input, log_n_size, cfg):
    """Forward coset LDE (wrappers/rust/src/lib.rs:114-136)."""
    _check(lib.compute_batched_lde(device_id, ptr(output), ptr(input), log_n_size, FORWARD, cfg))


def lde_batch_multi_gpu(output, input, num_gpu, cfg, log_n_size, total_num_input_elements,
                        total_num_output_elements):
    _check(lib.compute_batched_lde_multi_gpu(ptr(output), ptr(input), num_gpu, FORWARD, cfg, log_n_size,
                                             total_num_input_elements, total_num_output_elements))


def ntt_batch(device_id, inout, log_n_size, cfg):
    _check(lib.compute_batched_ntt(device_id, ptr(inout), log_n_size, FORWARD, cfg))


def intt_batch(device_id, inout, log_n_size, cfg):
    _check(lib.compute_batched_ntt(device_id, ptr(inout), log_n_size, INVERSE, cfg))


def ntt_raw(device_id, inout, log_n_size, direction, cfg):
    _check(lib.compute_batched_ntt(device_id, ptr(inout), log_n_size, direction, cfg))


def lde_raw(device_id, output, input, log_n_size, direction, cfg):
    _check(lib.compute_batched_lde(device_id, ptr(output), ptr(input), log_n_size, direction, cfg))


def transpose_rev_batch(device_id, output, input, log_n_size, cfg):
    _check(lib.compute_transpose_rev(device_id, ptr(output), ptr(input), log_n_size, cfg))


def fill_digests_buf_linear_gpu_with_gpu_ptr(digests, caps, leaves, digests_buf_size, cap_buf_size,
                                             leaves_buf_size, leaf_size, cap_height, hash_type=HashPoseidon,
                                             gpu_id=0):
    lib.fill_digests_buf_linear_gpu_with_gpu_ptr(ptr(digests), ptr(caps), ptr(leaves), digests_buf_size,
                                                 cap_buf_size, leaves_buf_size, leaf_size, cap_height, hash_type,
                                                 gpu_id)


def fill_digests_buf_linear_multigpu_with_gpu_ptr(digests, caps, leaves, digests_buf_size, cap_buf_size,
                                                  leaves_buf_size, leaf_size, cap_height, hash_type=HashPoseidon):
    lib.fill_digests_buf_linear_multigpu_with_gpu_ptr(ptr(digests), ptr(caps), ptr(leaves), digests_buf_size,
                                                      cap_buf_size, leaves_buf_size, leaf_size, cap_height,
                                                      hash_type)


def fill_digests_buf_linear_cpu(digests, caps, leaves, digests_buf_size, cap_buf_size, leaves_buf_size, leaf_size,
                                cap_height, hash_type=HashPoseidon):
    lib.fill_digests_buf_linear_cpu(ptr(digests), ptr(caps), ptr(leaves), digests_buf_size, cap_buf_size,
                                    leaves_buf_size, leaf_size, cap_height, hash_type)


def commit(device_id, coeffs, lde, leaves, digests, cap, lg_n, rate_bits, ncols, cap_height, shift=7):
    """Fused LDE -> leaf hashes -> tree -> cap (include/zeknox_ext.h zk_commit)."""
    _check(lib.zk_commit(device_id, ptr(coeffs), ptr(lde), ptr(leaves), ptr(digests), ptr(cap), lg_n, rate_bits,
                         ncols, cap_height, shift))


def commit_noalloc(device_id, coeffs, lde, leaves, digests, cap, lg_n, rate_bits, ncols, cap_height, shift=7):
    """zk_commit without device allocation (include/zeknox_ext.h zk_commit_noalloc)."""
    _check(lib.zk_commit_noalloc(device_id, ptr(coeffs), ptr(lde), ptr(leaves), ptr(digests), ptr(cap), lg_n,
                                 rate_bits, ncols, cap_height, shift))


def batch_commit(device_id, jobs, graph=True):
    """Batched commits (include/zeknox_ext.h zk_batch_commit).  jobs: list of
    dicts with coeffs, lde, leaves (or None), digests, cap, lg_n, rate_bits,
    ncols, cap_height."""
    arr = (CommitJob * len(jobs))()
    for i, j in enumerate(jobs):
        arr[i] = CommitJob(ptr(j["coeffs"]), ptr(j["lde"]), ptr(j.get("leaves")), ptr(j["digests"]), ptr(j["cap"]),
                           j["lg_n"], j["rate_bits"], j["ncols"], j["cap_height"])
    _check(lib.zk_batch_commit(device_id, arr, len(jobs), BATCH_GRAPH if graph else 0))


def batch_clear(device_id=0):
    _check(lib.zk_batch_clear(device_id))


def fill_random(device_id, out, n, seed):
    _check(lib.zk_fill_random(device_id, ptr(out), n, seed))


def profile_enable(on=True):
    lib.zk_profile_enable(1 if on else 0)


def profile_reset():
    lib.zk_profile_reset()


def profile_read():
    """{kernel name: (total ms, launches)} accumulated since the last reset."""
    out = {}
    i = 0
    buf = ctypes.create_string_buffer(128)
    while True:
        ms, n = ctypes.c_double(0), ctypes.c_uint64(0)
        e = lib.zk_profile_get(i, buf, 128, ctypes.byref(ms), ctypes.byref(n))
        if e.code != 0:
            break
        out[buf.value.decode()] = (ms.value, n.value)
        i += 1
    return out


class ServeRequest(ctypes.Structure):
    """zk_serve_request (include/zeknox_ext.h)."""
    _fields_ = [("lg_n", ctypes.c_uint32), ("rate_bits", ctypes.c_uint32), ("ncols", ctypes.c_uint32),
                ("cap_height", ctypes.c_uint32), ("coeffs", ctypes.c_void_p), ("seed", ctypes.c_uint64),
                ("keep", ctypes.c_uint32), ("reserved", ctypes.c_uint32)]


class ServeResult(ctypes.Structure):
    """zk_serve_result (include/zeknox_ext.h)."""
    _fields_ = [("submit_ns", ctypes.c_uint64), ("admit_ns", ctypes.c_uint64), ("done_ns", ctypes.c_uint64),
                ("lde", ctypes.c_void_p), ("digests", ctypes.c_void_p), ("device_bytes", ctypes.c_uint64)]


_sig("zk_server_create", RustError, [_sz, _u64, _u32, ctypes.POINTER(ctypes.c_void_p)])
_sig("zk_server_submit", RustError, [_v, ctypes.POINTER(ServeRequest), _v, ctypes.POINTER(ctypes.c_uint64)])
_sig("zk_server_wait", RustError, [_v, _u64, ctypes.POINTER(ServeResult)])
_sig("zk_server_release", RustError, [_v, _u64])
_sig("zk_server_destroy", RustError, [_v])


class Server:
    """Commit server (include/zeknox_ext.h zk_server_*): submit() queues a
    request and returns at once; wait() blocks until its cap is ready."""

    def __init__(self, device_id=0, budget_bytes=8 << 30, max_inflight=16):
        import numpy as np
        self._np = np
        h = ctypes.c_void_p()
        _check(lib.zk_server_create(device_id, budget_bytes, max_inflight, ctypes.byref(h)))
        self._h = h
        self._caps = {}

    def submit(self, lg_n, rate_bits, ncols, cap_height, coeffs=None, seed=0, keep=False):
        req = ServeRequest(lg_n, rate_bits, ncols, cap_height, ptr(coeffs), seed, int(keep), 0)
        cap = self._np.zeros(4 << cap_height, dtype=self._np.uint64)
        t = ctypes.c_uint64()
        _check(lib.zk_server_submit(self._h, ctypes.byref(req), cap.ctypes.data, ctypes.byref(t)))
        self._caps[t.value] = cap
        return t.value

    def wait(self, ticket):
        """-> dict(cap, submit_ns, admit_ns, done_ns, lde, digests, device_bytes)."""
        r = ServeResult()
        try:
            _check(lib.zk_server_wait(self._h, ticket, ctypes.byref(r)))
        finally:
            cap = self._caps.pop(ticket, None)
        return {"cap": cap, "submit_ns": r.submit_ns, "admit_ns": r.admit_ns, "done_ns": r.done_ns,
                "lde": r.lde, "digests": r.digests, "device_bytes": r.device_bytes}

    def release(self, ticket):
        _check(lib.zk_server_release(self._h, ticket))

    def close(self):
        if self._h:
            _check(lib.zk_server_destroy(self._h))
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


__all__ += ["Server", "ServeRequest", "ServeResult"]


_sig("zk_merkle_proofs", RustError, [_sz, _v, _v, _u64, _u64, _u64, _u64, _v, _u64, _v, _v])


def merkle_proofs(device_id, digests, leaves, n_leaves, leaf_size, cap_height, indices, col_stride=0,
                  with_rows=True):
    """FRI openings (include/zeknox_ext.h zk_merkle_proofs): returns
    (siblings: n x layers x 4, rows: n x leaf_size or None) as numpy u64."""
    import numpy as np
    import torch
    idx = np.ascontiguousarray(np.asarray(indices, dtype=np.uint64))
    layers = n_leaves.bit_length() - 1 - cap_height
    n = idx.size
    sib = torch.empty(max(n * layers * 4, 1), dtype=torch.int64, device="cuda")
    rows = torch.empty(max(n * leaf_size, 1), dtype=torch.int64, device="cuda") if with_rows else None
    _check(lib.zk_merkle_proofs(device_id, ptr(digests), ptr(leaves), col_stride, n_leaves, leaf_size, cap_height,
                                idx.ctypes.data, n, ptr(sib), ptr(rows)))
    s = sib.cpu().numpy().view(np.uint64)[: n * layers * 4].reshape(n, layers, 4)
    r = rows.cpu().numpy().view(np.uint64)[: n * leaf_size].reshape(n, leaf_size) if with_rows else None
    return s, r


__all__ += ["merkle_proofs"]
