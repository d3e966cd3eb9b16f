"""NTT plans the autotuner may pick (zk_ntt_tune_set) compute the same
coset LDE as the default plan, bit for bit, through compute_batched_lde --
including a plan whose first pass has 2 groups of 8 points per 256-thread
workgroup, so that its grid exceeds 2^32 threads and is launched in chunks.
The default plans are pinned against the oracle by test_gpu_ntt.py."""
import pytest

pytestmark = pytest.mark.gpu

PLANS = [  # lg, columns, (bits, gbits, kernel16)
    (21, 135, ((3, 8, 10), (1, 1, 1), (1, 1, 1))),   # 17.7 M workgroups of 256 threads in pass 1
    (21, 135, ((8, 3, 10), (1, 1, 1), (1, 1, 1))),
    (21, 40, ((4, 7, 10), (2, 2, 2), (1, 1, 0))),
    (20, 135, ((10, 10), (-1, -1), (0, 1))),
    (20, 17, ((7, 7, 6), (3, 2, 1), (1, 1, 1))),
    (17, 64, ((10, 7), (-1, 1), (1, 1))),
    (14, 33, ((6, 8), (0, 2), (1, 1))),
    (12, 5, ((4, 4, 4), (-1, 0, 2), (1, 1, 1))),
]


@pytest.fixture(scope="module")
def Z():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    import zeknox_amd
    zeknox_amd.init_cuda_rs()
    return zeknox_amd


@pytest.mark.parametrize("lg,C,plan", PLANS)
def test_tuned_plan_equals_default_plan(Z, lg, C, plan):
    import torch
    from gpu_util import empty
    r = 3
    n, N = 1 << (lg - r), 1 << lg
    coeffs = empty(C * n)
    Z.fill_random(0, coeffs, C * n, 500 + lg + C)
    cfg = Z.NTTConfig(batches=C, extension_rate_bits=r, with_coset=True, are_inputs_on_device=True,
                      are_outputs_on_device=True)
    loaded = Z.ntt_tune_get(lg)
    try:
        Z.ntt_tune_set(lg)
        ref = empty(C * N)
        Z.lde_batch(0, ref, coeffs, lg - r, cfg)
        Z.ntt_tune_set(lg, *plan)
        out = empty(C * N)
        out.fill_(-1)
        Z.lde_batch(0, out, coeffs, lg - r, cfg)
        torch.cuda.synchronize()
        assert torch.equal(out, ref), (lg, C, plan)
    finally:
        Z.ntt_tune_set(lg, *loaded) if loaded else Z.ntt_tune_set(lg)
