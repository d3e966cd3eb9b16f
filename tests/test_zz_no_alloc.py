"""zk_commit_noalloc (the commit with no device allocation) against zk_commit,
through the C ABI: LDE, row-major leaves, every digest and the cap equal bit
for bit, on calls in a row where each later call is smaller than the first
(the fixed scratch array keeps the previous call's contents).

The first call runs before any zk_commit of its shape, so its second LDE pass
computes the inter-pass twiddles itself (no table exists yet and the
allocation-free path builds none); the later calls find the table zk_commit
built.  40 columns of 2^20 rows take two column chunks (32 + 8) through the
256 MiB array, 9 columns one chunk, and the 2^17-row commit uses no scratch.
"""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


def commit_outputs(Z, fn, lg_n, r, C, h, seed, **kw):
    from gpu_util import empty
    N = 1 << (lg_n + r)
    nd = 2 * (N - (1 << h))
    coeffs = empty(C << lg_n)
    Z.fill_random(0, coeffs, C << lg_n, seed)
    out = {"lde": empty(C * N), "leaves": empty(C * N), "digests": empty(nd * 4), "cap": empty(4 << h)}
    for t in out.values():
        t.fill_(-1)  # every element must be written
    fn(0, coeffs, out["lde"], out["leaves"], out["digests"], out["cap"], lg_n, r, C, h, **kw)
    return out


@pytest.mark.parametrize("calls", [[(17, 3, 40, 4, 11), (17, 3, 9, 4, 12), (14, 3, 33, 3, 13)]])
def test_noalloc_commit_equals_commit(Z, calls):
    import torch
    for i, (lg_n, r, C, h, seed) in enumerate(calls):
        print("zk_commit_noalloc call %d: %d columns x 2^%d -> 2^%d rows, cap_height %d"
              % (i, C, lg_n, lg_n + r, h), flush=True)
        a = commit_outputs(Z, Z.commit_noalloc, lg_n, r, C, h, seed)
        b = commit_outputs(Z, Z.commit, lg_n, r, C, h, seed)
        for key in ("lde", "leaves", "digests", "cap"):
            assert torch.equal(a[key], b[key]), (i, key)
        del a, b
        torch.cuda.empty_cache()


def test_noalloc_commit_refuses_another_shift(Z):
    with pytest.raises(Z.ZeknoxError) as e:
        commit_outputs(Z, Z.commit_noalloc, 10, 1, 4, 2, 1, shift=11)
    assert e.value.code == 1
