"""Helpers shared by the GPU tests: device buffers from numpy u64 arrays."""
import numpy as np
import torch


def dev(a):
    """numpy u64 -> device tensor (int64 view of the same bits)."""
    a = np.ascontiguousarray(np.asarray(a, dtype=np.uint64))
    return torch.from_numpy(a.view(np.int64).copy()).to("cuda")


def empty(n):
    return torch.empty(max(int(n), 1), dtype=torch.int64, device="cuda")


def host(t, n=None):
    a = t.detach().cpu().numpy().view(np.uint64)
    return a if n is None else a[:n]


def d2h(addr, n):
    """n u64 at a raw device address -> numpy (hipMemcpy through the HIP runtime)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    out = np.empty(int(n), dtype=np.uint64)
    rc = hip.hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(addr), ctypes.c_size_t(out.nbytes),
                       ctypes.c_int(2))  # hipMemcpyDeviceToHost
    assert rc == 0, rc
    return out
