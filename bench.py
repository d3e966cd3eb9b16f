"""Benchmark: Plonky2 commit (Goldilocks coset LDE -> Poseidon Merkle) rows/s on one MI355X.

Workload (one step): 135 polynomials of 2^17 coefficients (u64, uniform in
[0, p), seed 2, generated on the device), coset LDE with blowup 8 (shift 7) to
2^20 rows, then the Poseidon Merkle forest over the 2^20 leaves of 135
elements with cap_height 4.  Inputs are resident in HBM before timing.

The timed path is the fused entry point zk_commit writing the LDE, the
row-major leaves, every digest and the cap (the same outputs as the three-call
ABI path compute_batched_lde -> compute_transpose_rev ->
fill_digests_buf_linear_gpu_with_gpu_ptr, which is timed too and reported as
"plain_abi").  Kernel durations come from HIP events the library records on
its own stream around every launch during the timed steps.

The CPU baseline runs the oracle (a C port of the path, OpenMP) on a bounded
sample of the same workload on the host's cores.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "LDE+Merkle commit rows/s at 1 GPU; batched vs plain speedup; p99 latency"
HBM_PEAK = 8.0e12  # MI355X HBM3E, bytes/s
W = 8  # bytes per field element


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--columns", type=int, default=135)
    ap.add_argument("--lg-n", type=int, default=17)
    ap.add_argument("--rate-bits", type=int, default=3)
    ap.add_argument("--cap-height", type=int, default=4)
    ap.add_argument("--workload", choices=["plain", "large"], default="plain",
                    help="plain: 135 x 2^17 -> 2^20 commit; large: 400 x 2^18 values, iNTT, -> 2^21 commit")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-lg-n", type=int, default=13, help="degree of the CPU baseline sample")
    return ap.parse_args()


def cpu_baseline(C, lg_n, r, h):
    """Oracle (C + OpenMP) on the host: LDE + transpose_rev + Merkle, rows/s."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    L = O.lib()
    threads = L.or_get_threads()
    N = 1 << (lg_n + r)
    coeffs = O.random(C << lg_n, 2)
    times = []
    for _ in range(2):
        t0 = time.perf_counter()
        lde = O.lde(coeffs, lg_n, r, C)
        leaves = O.transpose_rev(lde, lg_n + r, C)
        O.merkle(leaves, N, C, h)
        times.append(time.perf_counter() - t0)
    t = min(times)
    return {"value": N / t, "unit": "rows/s", "cores": threads, "kind": "port",
            "sample": "%d columns x 2^%d coefficients -> 2^%d rows, cap_height %d (same shape as the GPU "
                      "workload at a 2^%d-times smaller degree), best of 2, %.2f s per commit"
                      % (C, lg_n, lg_n + r, h, 17 - lg_n, t)}


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: this build runs on exactly one GPU; --gpus %d is not supported "
              "(the commit is single-GPU; see DESIGN.md for how it would shard)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import zeknox_amd as Z

    intt = a.workload == "large"
    if intt:
        a.columns, a.lg_n = 400, 18
    C, lg_n, r, h = a.columns, a.lg_n, a.rate_bits, a.cap_height
    n, N = 1 << lg_n, 1 << (lg_n + r)
    nd = 2 * (N - (1 << h))
    dev = torch.device("cuda:0")
    coeffs = torch.empty(C * n, dtype=torch.int64, device=dev)
    Z.fill_random(0, coeffs, C * n, 2)
    lde = torch.empty(C * N, dtype=torch.int64, device=dev)
    leaves = torch.empty(C * N, dtype=torch.int64, device=dev)
    dg = torch.empty(nd * 4, dtype=torch.int64, device=dev)
    cp = torch.empty((1 << h) * 4, dtype=torch.int64, device=dev)

    icfg = Z.NTTConfig(batches=C, are_inputs_on_device=True, are_outputs_on_device=True)

    def fused():
        if intt:  # evaluations -> coefficients, in place (the data changes every step, the work does not)
            Z.intt_batch(0, coeffs, lg_n, icfg)
        Z.commit(0, coeffs, lde, leaves, dg, cp, lg_n, r, C, h)

    ncfg = Z.NTTConfig(batches=C, extension_rate_bits=r, with_coset=True, are_inputs_on_device=True,
                       are_outputs_on_device=True)
    tcfg = Z.TransposeConfig(batches=C, are_inputs_on_device=True, are_outputs_on_device=True)

    def plain():
        if intt:
            Z.intt_batch(0, coeffs, lg_n, icfg)
        Z.lde_batch(0, lde, coeffs, lg_n, ncfg)
        Z.transpose_rev_batch(0, leaves, lde, lg_n + r, tcfg)
        Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dg, cp, leaves, nd, 1 << h, N, C, h)

    def timed(fn, steps, warmup, profile):
        for _ in range(warmup):
            fn()
        Z.profile_reset()
        Z.profile_enable(profile)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        Z.profile_enable(False)
        prof = Z.profile_read()
        return (t1 - t0) / steps, prof

    t_fused, prof = timed(fused, a.steps, a.warmup, True)
    cap_fused = cp.clone()
    t_plain, prof_plain = timed(plain, max(2, a.steps // 2), 1, True)
    if not intt:
        assert torch.equal(cap_fused, cp), "fused and three-call commits disagree"

    def avg(p, k):
        ms, cnt = p.get(k, (0.0, 0))
        return (ms / cnt * 1e-3) if cnt else None

    steps = a.steps
    kern = {k: {"avg_ms": v[0] / v[1], "launches_per_step": v[1] / steps} for k, v in prof.items()}
    t_p1, t_p2 = avg(prof, "lde.pass1"), avg(prof, "lde.pass2")
    bytes_p1 = W * C * (n + N)          # read coefficients, write evaluations
    bytes_p2 = 2 * W * C * N            # read + write evaluations
    b_lde = W * C * (n + N)
    t_p3 = avg(prof, "lde.pass3")
    t_lde = (t_p1 or 0) + (t_p2 or 0) + (t_p3 or 0)
    perms_leaf = N * ((C + 7) // 8) if C > 4 else 0
    perms_tree = N - (1 << h)
    leaf_ms = prof.get("merkle.leaves_colmajor", (0, 1))
    t_leaf = leaf_ms[0] / leaf_ms[1] * 1e-3
    tree_ms = sum(prof.get(k, (0, 0))[0] for k in ("merkle.levels", "merkle.caps")) / steps * 1e-3
    result = {
        "metric": METRIC,
        "value": N / t_fused,
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": t_fused * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": ("large commit: %d columns x 2^%d evaluations, iNTT to coefficients, "
                                % (C, lg_n) if intt else "plain commit: %d polynomials x 2^%d coefficients, "
                                % (C, lg_n)) + "coset LDE blowup %d -> 2^%d rows, Poseidon Merkle cap_height %d"
                               % (1 << r, lg_n + r, h),
                   "columns": C, "lg_n": lg_n, "rate_bits": r, "rows": N, "cap_height": h, "shift": 7,
                   "path": ("compute_batched_ntt (inverse) + " if intt else "")
                   + "zk_commit (LDE + leaves + digests + cap)"},
        "roofline": {"bound": "hbm", "kernel": "lde.pass2", "achieved": bytes_p2 / t_p2 / 1e9 if t_p2 else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": bytes_p2 / t_p2 / HBM_PEAK if t_p2 else None, "traffic": None,
                     "algorithmic_bytes": bytes_p2},
        "lde": {"pass1_GBps": bytes_p1 / t_p1 / 1e9 if t_p1 else None,
                "pass2_GBps": bytes_p2 / t_p2 / 1e9 if t_p2 else None,
                "pass3_GBps": bytes_p2 / t_p3 / 1e9 if t_p3 else None,
                "lde_ms": t_lde * 1e3, "lde_GBps_algorithmic": b_lde / t_lde / 1e9 if t_lde else None,
                "lde_frac": b_lde / t_lde / HBM_PEAK if t_lde else None},
        "merkle": {"leaf_hash_ms": t_leaf * 1e3, "tree_ms": tree_ms * 1e3,
                   "leaf_permutations": perms_leaf, "tree_permutations": perms_tree,
                   "leaves_hashed_per_s": N / (t_leaf + tree_ms) if t_leaf else None,
                   "permutations_per_s": perms_leaf / t_leaf if t_leaf else None},
        "plain_abi": {"ms_per_step": t_plain * 1e3, "rows_per_s": N / t_plain,
                      "kernels": {k: v[0] / v[1] for k, v in prof_plain.items()}},
        "kernels": kern,
    }
    if intt:
        result["intt_ms"] = sum(avg(prof, "intt.pass%d" % k) or 0 for k in (1, 2, 3)) * 1e3
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(C, a.cpu_lg_n, r, h)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
