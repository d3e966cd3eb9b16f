"""Benchmark: Plonky2 commit (Goldilocks coset LDE -> Poseidon Merkle) on one MI355X.

The metric has three parts, all measured by the default run:

* rows/s of the plain commit (the JSON line's `value`): 135 polynomials of
  2^17 coefficients (u64, uniform in [0, p), seed 2, generated on the device),
  coset LDE with blowup 8 (shift 7) to 2^20 rows, Poseidon Merkle forest over
  the 2^20 leaves of 135 elements, cap_height 4; inputs resident in HBM.  The
  timed path is zk_commit writing the LDE, the row-major leaves, every digest
  and the cap (the outputs of compute_batched_lde -> compute_transpose_rev ->
  fill_digests_buf_linear_gpu_with_gpu_ptr, timed beside it as "plain_abi").
  "passes" gives each NTT pass's HBM throughput; "large" the same for the
  400 x 2^18 evaluations -> iNTT -> 2^21-row commit.
* batched vs plain speedup ("batched"): 48 commits with log2 degree uniform in
  10..18, 4..135 columns, blowup 2/4/8, cap_height min(4, log2 rows - 2)
  (shapes and coefficients from seed 3), through zk_batch_commit replayed
  from a hipGraph, against the same 48 jobs as sequential three-call ABI
  commits; with the kernel time of the jobs one at a time, the kernel time
  of one batched call (its one leaf-hash launch over all jobs, tree levels,
  LDEs), and the jobs' work at the plain commit's full-GPU kernel rates.
* p50/p99 latency of a served stream ("served"): requests with log2 rows
  uniform in 12..20, 8..200 columns, blowup 8, cap_height 4, synthetic
  coefficients per request seed, through the commit server (zk_server_*,
  200 GB budget) against the plain per-request path (the same work one
  request at a time: zk_commit without a leaf matrix, cap to the host):
  saturation throughput with every request queued at once, then latency under
  Poisson arrivals at 80% of the plain path's capacity.

Kernel durations come from HIP events the library records around every
launch during the timed steps.  `roofline` is the dominant LDE pass against
the HBM peak; `roofline_hash` the leaf-hash kernel against the VALU issue
capacity (its instruction mix's SIMD-cycles per permutation, stored in
profiles/commit_135x2p20_traffic.json, against 256 CUs x 4 SIMDs x 2.4 GHz).  The CPU baseline commits the plain workload's
full shape on the host cores (median of 5 runs after a warm-up): the
reference's compiled CPU Merkle and the oracle's C port of the LDE.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu] [--no-large]
                    [--workload plain|large] [--no-batched] [--no-served]
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
# MI355X VALU issue capacity: 256 CUs x 4 SIMDs at 2.4 GHz, in SIMD-cycles per
# second.  The leaf-hash roofline's peak is the issue-limited rate of the
# library's own permutation on this part: the leaf kernel's sponge with its
# inputs computed in registers, no memory traffic, as many waves per SIMD as
# the registers allow (tests/native/perm_rate.hip, result kept in the traffic
# profile); the SIMD-cycles per wave-permutation it measures are reported
# beside it.
SIMD_CYCLES_PER_S = 256 * 4 * 2.4e9
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "commit_135x2p20_traffic.json")
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
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-requests", type=int, default=96)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device-to-device twice (a kernel of known HBM traffic for counter runs)")
    ap.add_argument("--cpu-lg-n", type=int, default=17, help="log2 degree of the CPU baseline commit")
    ap.add_argument("--cpu-reps", type=int, default=5, help="timed CPU baseline runs (after one warm-up)")
    ap.add_argument("--no-large", action="store_true", help="skip the 400 x 2^18 -> 2^21 sub-record")
    ap.add_argument("--large-steps", type=int, default=3)
    return ap.parse_args()


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(C, lg_n, r, h, reps):
    """The commit on the host's cores, rows/s: one warm-up, then the median of
    `reps` runs.  The Merkle forest (most of the CPU time) runs the
    reference's own CPU code (native/merkle/merkle.cpp, compiled from its
    sources into oracle/_ref by oracle/ref.mk) when that library is present;
    the LDE and the leaf transpose, which the reference has no CPU version
    of, run the oracle's C port (OpenMP)."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    O.build()
    L = O.lib()
    threads = L.or_get_threads()
    ref = O.ref_lib() is not None
    merkle = O.ref_merkle if ref else O.merkle
    N = 1 << (lg_n + r)
    coeffs = O.random(C << lg_n, 2)
    times = []
    for i in range(reps + 1):
        t0 = time.perf_counter()
        lde = O.lde(coeffs, lg_n, r, C)
        leaves = O.transpose_rev(lde, lg_n + r, C)
        del lde
        merkle(leaves, N, C, h)
        del leaves
        if i:  # the first run is the warm-up
            times.append(time.perf_counter() - t0)
        print("cpu baseline: run %d of %d, %.2f s" % (i, reps, time.perf_counter() - t0), file=sys.stderr,
              flush=True)
    t = float(np.median(times))
    what = ("Merkle: the reference's compiled CPU code; LDE and transpose: the oracle's C port" if ref
            else "LDE, transpose and Merkle: the oracle's C port")
    return {"value": N / t, "unit": "rows/s", "cores": threads, "cpu_model": cpu_model(),
            "kind": "reference" if ref else "port",
            "sample": "%d columns x 2^%d coefficients -> 2^%d rows, cap_height %d, median of %d runs after one "
                      "warm-up (%.2f s per commit, runs %s s); %s"
                      % (C, lg_n, lg_n + r, h, reps, t, ", ".join("%.2f" % x for x in times), what)}


def empty(n):
    import torch
    return torch.empty(max(int(n), 1), dtype=torch.int64, device="cuda")


def abi_commit(Z, coeffs, lde, leaves, dg, cp, lg_n, r, C, h):
    """The reference's three-call commit (each call synchronous)."""
    N = 1 << (lg_n + r)
    Z.lde_batch(0, lde, coeffs, lg_n, Z.NTTConfig(batches=C, extension_rate_bits=r, with_coset=True,
                                                  are_inputs_on_device=True, are_outputs_on_device=True))
    Z.transpose_rev_batch(0, leaves, lde, lg_n + r, Z.TransposeConfig(batches=C, are_inputs_on_device=True,
                                                                      are_outputs_on_device=True))
    nd = 2 * (N - (1 << h)) if (1 << h) < N else 0
    Z.fill_digests_buf_linear_gpu_with_gpu_ptr(dg, cp, leaves, nd, 1 << h, N, C, h)


def batched_shapes(n=48, seed=3):
    """The mixed batched workload: (log2 degree, rate_bits, columns, cap_height,
    coefficient seed) per job; log2 degree uniform in 10..18, 4..135 columns,
    blowup 2/4/8, cap_height min(4, log2 rows - 2)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        k = int(rng.integers(10, 19))
        C = int(rng.integers(4, 136))
        r = int(rng.choice([1, 2, 3]))
        out.append((k, r, C, min(4, k + r - 2), 3000 + i))
    return out


def served_requests(n, seed=5):
    """The served stream: (log2 degree, rate_bits, columns, cap_height, seed)
    per request; log2 rows uniform in 12..20, 8..200 columns, blowup 8,
    cap_height 4, per-request seed splitmix64(5 ^ i)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        lg_rows = int(rng.integers(12, 21))
        out.append((lg_rows - 3, 3, int(rng.integers(8, 201)), 4, splitmix64(5 ^ i)))
    return out


def kernel_ms(Z, fn):
    """Runs fn once with the library's per-kernel timing on; returns
    (sum of kernel milliseconds, {kernel: ms})."""
    import torch
    torch.cuda.synchronize()
    Z.profile_reset()
    Z.profile_enable(True)
    fn()
    torch.cuda.synchronize()
    Z.profile_enable(False)
    prof = Z.profile_read()
    return sum(v[0] for v in prof.values()), {k: v[0] for k, v in prof.items()}


def work_ms_at_plain_rates(shapes, rates):
    """Device time of commits if every kernel ran at the rate it reaches on the
    plain 135 x 2^20 commit (a full GPU): leaf and tree permutations at the
    leaf kernel's permutation rate, LDE bytes at the plain LDE's byte rate.
    Small commits cannot fill the GPU alone; a batched or served form that
    overlaps them, or fuses their kernels, approaches this estimate (and can
    go below it where its kernels outrun the plain commit's, e.g. one-pass
    LDEs or leaves of at most 8 elements)."""
    perms = lde_bytes = 0
    for k, r, C, h, _ in shapes:
        N = 1 << (k + r)
        perms += N * ((C + 7) // 8 if C > 4 else 0) + (N - (1 << h))
        lde_bytes += W * C * ((1 << k) + N) + (2 * W * C * N if k + r > 10 else 0) + \
            (2 * W * C * N if k + r > 20 else 0)
    return 1e3 * (perms / rates["perm_per_s"] + lde_bytes / rates["lde_Bps"])


def run_batched(Z, rates, reps=5):
    """48 mixed commits: zk_batch_commit (graph replay) vs 48 sequential plain
    calls, with the device-busy time of the same jobs run one at a time
    (three-call path and zk_commit), the kernel time of one batched call and
    the jobs' work at the plain commit's kernel rates."""
    import torch
    jobs = []
    for k, r, C, h, seed in batched_shapes():
        N = 1 << (k + r)
        nd = 2 * (N - (1 << h))
        coeffs = empty(C << k)
        Z.fill_random(0, coeffs, C << k, seed)
        jobs.append({"coeffs": coeffs, "lde": empty(C * N), "leaves": empty(C * N), "digests": empty(nd * 4),
                     "cap": empty(4 << h), "lg_n": k, "rate_bits": r, "ncols": C, "cap_height": h})
    rows = sum(1 << (j["lg_n"] + j["rate_bits"]) for j in jobs)

    def plain():
        for j in jobs:
            abi_commit(Z, j["coeffs"], j["lde"], j["leaves"], j["digests"], j["cap"], j["lg_n"], j["rate_bits"],
                       j["ncols"], j["cap_height"])

    def fused_one_by_one():
        for j in jobs:
            Z.commit(0, j["coeffs"], j["lde"], j["leaves"], j["digests"], j["cap"], j["lg_n"], j["rate_bits"],
                     j["ncols"], j["cap_height"])

    plain()
    caps_plain = [j["cap"].clone() for j in jobs]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        plain()
    torch.cuda.synchronize()
    t_plain = (time.perf_counter() - t0) / reps
    busy_plain, k_plain = kernel_ms(Z, plain)
    fused_one_by_one()
    busy_fused, k_fused = kernel_ms(Z, fused_one_by_one)
    for j in jobs:
        j["cap"].zero_()
    Z.batch_commit(0, jobs, graph=True)  # eager
    Z.batch_commit(0, jobs, graph=True)  # captures the graph
    assert all(torch.equal(j["cap"], c) for j, c in zip(jobs, caps_plain)), "batched and plain commits disagree"
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        Z.batch_commit(0, jobs, graph=True)
    torch.cuda.synchronize()
    t_batch = (time.perf_counter() - t0) / reps
    # the kernels of one batched call (eager, so the library times each launch)
    _, k_batch = kernel_ms(Z, lambda: Z.batch_commit(0, jobs, graph=False))
    Z.batch_clear(0)
    leaf_fused = sum(v for k, v in k_fused.items() if k.startswith("merkle.leaves"))
    leaf_perms = sum((1 << (k + r)) * ((C + 7) // 8 if C > 4 else 0) for k, r, C, h, _ in batched_shapes())
    t_leaves = k_batch.get("batch.leaves", 0.0)
    return {"workload": "48 commits, log2 degree 10..18, 4..135 columns, blowup 2/4/8, cap_height "
                        "min(4, log2 rows - 2), seed 3", "jobs": 48, "rows": rows,
            "batched_ms": t_batch * 1e3, "plain_ms": t_plain * 1e3, "batched_rows_per_s": rows / t_batch,
            "plain_rows_per_s": rows / t_plain, "speedup": t_plain / t_batch,
            "device_busy_ms": {"plain_three_call": busy_plain, "fused_one_by_one": busy_fused,
                               "fused_leaf_hash": leaf_fused},
            "batched_kernels_ms": {"leaves_one_launch": t_leaves, "levels": k_batch.get("batch.levels", 0.0),
                                   "lde": sum(v for k, v in k_batch.items() if k.startswith("lde."))},
            "leaf_launch_perm_per_s": leaf_perms / (t_leaves / 1e3) if t_leaves else None,
            "two_x_needs_ms": t_plain * 1e3 / 2,
            "work_ms_at_plain_rates": work_ms_at_plain_rates(batched_shapes(), rates),
            "note": "2x the plain calls needs two_x_needs_ms for the whole batch; the one leaf-hash launch over "
                    "all 48 jobs (91 M permutations on the full GPU) takes leaves_one_launch alone, so 2x "
                    "needs a faster leaf kernel, not more overlap.  work_ms_at_plain_rates: the jobs' work "
                    "at the plain 135 x 2^20 commit's kernel rates (an estimate, not a bound)",
            "path": "zk_batch_commit (LDEs on 4 streams, then one leaf-hash launch and one launch per tree "
                    "height over all jobs; hipGraph replay) vs 48 x (compute_batched_lde + "
                    "compute_transpose_rev + fill_digests_buf_linear_gpu_with_gpu_ptr)"}


def splitmix64(x):
    x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    return x ^ (x >> 31)


def pct(v, q):
    import numpy as np
    return float(np.percentile(np.asarray(v), q))


def run_served(Z, nreq, rates):
    """Commit server vs the plain per-request path: saturation throughput and
    p50/p99 latency under Poisson arrivals at 80% of plain capacity.  The
    plain path does what the server does for a request, one request at a
    time: synthetic coefficients, zk_commit (LDE, digests, cap; no leaf
    matrix), the cap copied to the host.  The reference's three-call path
    (which also writes the leaf matrix) is timed beside it at saturation."""
    import threading
    import numpy as np
    import torch
    rng = np.random.default_rng(5)
    reqs = served_requests(nreq)
    for _ in range(nreq):  # the arrival draws follow the shape draws of the same generator
        rng.integers(12, 21), rng.integers(8, 201)
    rows = [1 << (k + r) for k, r, _, _, _ in reqs]
    n_coef = max(C << k for k, r, C, h, _ in reqs)
    n_lde = max(C << (k + r) for k, r, C, h, _ in reqs)
    n_dg = max(8 << (k + r) for k, r, C, h, _ in reqs)
    buf = {"coeffs": empty(n_coef), "lde": empty(n_lde), "leaves": empty(n_lde), "dg": empty(n_dg), "cp": empty(64)}

    def plain_one(q):
        k, r, C, h, seed = q
        Z.fill_random(0, buf["coeffs"], C << k, seed)
        Z.commit(0, buf["coeffs"], buf["lde"], None, buf["dg"], buf["cp"], k, r, C, h)
        return buf["cp"][: 4 << h].cpu()

    def plain_abi_one(q):
        k, r, C, h, seed = q
        Z.fill_random(0, buf["coeffs"], C << k, seed)
        abi_commit(Z, buf["coeffs"], buf["lde"], buf["leaves"], buf["dg"], buf["cp"], k, r, C, h)
        return buf["cp"][: 4 << h].cpu()

    for q in reqs[:4]:
        plain_one(q)
        plain_abi_one(q)
    torch.cuda.synchronize()
    # saturation: all requests at t = 0
    t0 = time.perf_counter()
    caps_plain = [plain_one(q) for q in reqs]
    t_plain_sat = time.perf_counter() - t0
    t0 = time.perf_counter()
    for q in reqs:
        plain_abi_one(q)
    t_abi_sat = time.perf_counter() - t0
    busy_plain, _ = kernel_ms(Z, lambda: [plain_one(q) for q in reqs])
    free, _ = torch.cuda.mem_get_info()
    budget = int(min(200e9, free - 16e9))
    srv = Z.Server(0, budget_bytes=budget, max_inflight=16)
    try:
        for q in reqs[:4]:
            srv.wait(srv.submit(q[0], q[1], q[2], q[3], seed=q[4]))
        t0 = time.perf_counter()
        tickets = [srv.submit(q[0], q[1], q[2], q[3], seed=q[4]) for q in reqs]
        res = [srv.wait(t) for t in tickets]
        t_srv_sat = time.perf_counter() - t0
        assert all((x["cap"] == c.numpy().view(np.uint64)).all() for x, c in zip(res, caps_plain)), \
            "served and plain commits disagree"
        # Poisson arrivals at 80% of the plain path's capacity
        lam = 0.8 * nreq / t_plain_sat
        arrivals = np.cumsum(rng.exponential(1.0 / lam, nreq))
        # plain: one request at a time, first come first served
        lat_plain = []
        start = time.perf_counter()
        for a, q in zip(arrivals, reqs):
            now = time.perf_counter() - start
            if now < a:
                time.sleep(a - now)
            plain_one(q)
            lat_plain.append(time.perf_counter() - start - a)
        # served: a submitter thread follows the same arrival times
        tick = [None] * nreq
        start = time.perf_counter()

        def submitter():
            for i, (a, q) in enumerate(zip(arrivals, reqs)):
                now = time.perf_counter() - start
                if now < a:
                    time.sleep(a - now)
                tick[i] = srv.submit(q[0], q[1], q[2], q[3], seed=q[4])

        th = threading.Thread(target=submitter)
        th.start()
        th.join()
        res2 = [srv.wait(t) for t in tick]
        lat_srv = [(x["done_ns"] - x["submit_ns"]) / 1e9 for x in res2]
    finally:
        srv.close()
    tot = sum(rows)
    return {"workload": "%d requests, log2 rows 12..20, 8..200 columns, blowup 8, cap_height 4, per-request "
                        "seeds splitmix64(5 ^ i)" % nreq, "requests": nreq, "budget_bytes": budget,
            "max_inflight": 16,
            "saturation": {"served_rows_per_s": tot / t_srv_sat, "plain_rows_per_s": tot / t_plain_sat,
                           "speedup": t_plain_sat / t_srv_sat, "served_ms": t_srv_sat * 1e3,
                           "plain_ms": t_plain_sat * 1e3, "plain_device_busy_ms": busy_plain,
                           "work_ms_at_plain_rates": work_ms_at_plain_rates(reqs, rates),
                           "plain_over_work_at_plain_rates": t_plain_sat * 1e3 / work_ms_at_plain_rates(reqs, rates),
                           "plain_abi_rows_per_s": tot / t_abi_sat, "speedup_vs_plain_abi": t_abi_sat / t_srv_sat},
            "poisson": {"rate_per_s": lam, "load_vs_plain_capacity": 0.8,
                        "served_p50_ms": pct(lat_srv, 50) * 1e3, "served_p99_ms": pct(lat_srv, 99) * 1e3,
                        "plain_p50_ms": pct(lat_plain, 50) * 1e3, "plain_p99_ms": pct(lat_plain, 99) * 1e3},
            "path": "zk_server_* (memory-budget admission; the requests admitted together run as one group: LDEs on 4 streams, fused forest launches) vs one request at a time: zk_fill_random + "
                    "zk_commit (no leaf matrix, like the server) + cap to host; plain_abi: zk_fill_random + "
                    "compute_batched_lde + compute_transpose_rev + fill_digests_buf_linear_gpu_with_gpu_ptr"}


def commit_run(Z, C, lg_n, r, h, steps, warmup, intt, with_plain):
    """Times the fused commit (zk_commit writing LDE, leaves, digests, cap;
    preceded by an in-place iNTT of the inputs for the large workload) and,
    optionally, the reference's three-call path on the same buffers.
    Returns a dict of the measurements, per-kernel times included."""
    import torch
    n, N = 1 << lg_n, 1 << (lg_n + r)
    nd = 2 * (N - (1 << h))
    dev = torch.device("cuda:0")
    coeffs = torch.empty(C * n, dtype=torch.int64, device=dev)
    Z.fill_random(0, coeffs, C * n, 4 if intt else 2)
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

    # Wall time without the library's per-kernel events (each timed kernel
    # gets an event record before and after it, which leaves a gap of about
    # 10 us between kernels on the GPU); the kernel times come from a second,
    # instrumented pass over the same steps.
    out = {}
    out["t_fused"], _ = timed(fused, steps, warmup, False)
    _, out["prof"] = timed(fused, steps, 0, True)
    if with_plain:
        cap_fused = cp.clone()
        out["t_plain"], _ = timed(plain, max(2, steps // 2), 1, False)
        _, out["prof_plain"] = timed(plain, max(2, steps // 2), 0, True)
        if not intt:
            assert torch.equal(cap_fused, cp), "fused and three-call commits disagree"
    del coeffs, lde, leaves, dg, cp
    torch.cuda.empty_cache()
    return out


def commit_summary(run, C, lg_n, r, h, steps, intt):
    """rows/s, per-pass HBM throughput of the LDE (and iNTT), leaf hashing
    and tree times of one commit_run."""
    n, N = 1 << lg_n, 1 << (lg_n + r)
    prof = run["prof"]

    def avg(k):
        ms, cnt = prof.get(k, (0.0, 0))
        return (ms / cnt * 1e-3) if cnt else None

    passes = {}
    bytes_first = W * C * (n + N)       # read coefficients, write evaluations
    bytes_later = 2 * W * C * N         # read + write evaluations
    for p in (1, 2, 3):
        t = avg("lde.pass%d" % p)
        if t:
            b = bytes_first if p == 1 else bytes_later
            passes["lde.pass%d" % p] = {"ms": t * 1e3, "GBps": b / t / 1e9, "frac": b / t / HBM_PEAK}
        t = avg("intt.pass%d" % p)
        if t:
            b = 2 * W * C * n
            passes["intt.pass%d" % p] = {"ms": t * 1e3, "GBps": b / t / 1e9, "frac": b / t / HBM_PEAK}
    t_lde = sum(v["ms"] for k, v in passes.items() if k.startswith("lde")) * 1e-3
    b_lde = W * C * (n + N)
    perms_leaf = N * ((C + 7) // 8) if C > 4 else 0
    leaf = prof.get("merkle.leaves_colmajor", (0, 1))
    t_leaf = leaf[0] / leaf[1] * 1e-3
    tree_ms = sum(prof.get(k, (0, 0))[0] for k in ("merkle.levels", "merkle.top")) / steps * 1e-3
    return {
        "rows_per_s": N / run["t_fused"], "ms_per_step": run["t_fused"] * 1e3,
        "passes": passes,
        "lde": {"lde_ms": t_lde * 1e3, "lde_GBps_algorithmic": b_lde / t_lde / 1e9 if t_lde else None,
                "lde_frac": b_lde / t_lde / HBM_PEAK if t_lde else None},
        "merkle": {"leaf_hash_ms": t_leaf * 1e3, "tree_ms": tree_ms * 1e3, "leaf_permutations": perms_leaf,
                   "tree_permutations": N - (1 << h),
                   "leaves_hashed_per_s": N / (t_leaf + tree_ms) if t_leaf else None,
                   "permutations_per_s": perms_leaf / t_leaf if t_leaf else None},
        "kernels": {k: {"avg_ms": v[0] / v[1], "launches_per_step": v[1] / steps} for k, v in prof.items()},
    }


def main():
    a = parse()
    if a.gpus != 1:
        print("bench.py: the commit runs on exactly one GPU; --gpus %d is not supported "
              "(a multi-GPU commit would split the tree into 2^cap_height subtrees, one "
              "coset LDE and subtree per GPU)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    import torch
    import zeknox_amd as Z

    intt = a.workload == "large"
    if intt:
        a.columns, a.lg_n = 400, 18
    C, lg_n, r, h = a.columns, a.lg_n, a.rate_bits, a.cap_height
    N = 1 << (lg_n + r)
    if a.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(2):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    steps = a.steps
    run = commit_run(Z, C, lg_n, r, h, steps, a.warmup, intt, with_plain=True)
    summ = commit_summary(run, C, lg_n, r, h, steps, intt)
    p2 = summ["passes"].get("lde.pass2")
    bytes_p2 = 2 * W * C * N
    result = {
        "metric": METRIC,
        "value": N / run["t_fused"],
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": a.warmup,
        "ms_per_step": run["t_fused"] * 1e3,
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
        "roofline": {"bound": "hbm", "kernel": "lde.pass2", "achieved": p2["GBps"] if p2 else None,
                     "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": p2["frac"] if p2 else None, "traffic": None,
                     "algorithmic_bytes": bytes_p2},
        "passes": summ["passes"],
        "lde": summ["lde"],
        "merkle": summ["merkle"],
        "plain_abi": {"ms_per_step": run["t_plain"] * 1e3, "rows_per_s": N / run["t_plain"],
                      "kernels": {k: v[0] / v[1] for k, v in run["prof_plain"].items()}},
        "kernels": summ["kernels"],
    }
    prof_c2 = (C, lg_n, r, h) == (135, 17, 3, 4) and not intt and os.path.exists(TRAFFIC_PROFILE)
    if prof_c2:
        # HBM bytes per launch from the calibrated FETCH_SIZE / WRITE_SIZE runs
        # (profiles/commit_135x2p20_traffic.json), the leaf-hash kernel's VALU
        # instructions per permutation from its SQ_INSTS_VALU run, and the
        # issue-limited permutation rate its roofline divides by
        tp = json.load(open(TRAFFIC_PROFILE))
        result["roofline"]["traffic"] = tp["kernels"]["lde.pass2"]["bytes"]
        perms_leaf = summ["merkle"]["leaf_permutations"]
        t_leaf = summ["merkle"]["leaf_hash_ms"] * 1e-3
        v = tp["valu_issue"]["leaf_hash_kernel<ColMajorRev, rows>"]
        pr = tp["permutation_rate"]
        rate = perms_leaf / t_leaf if t_leaf else None
        peak = pr["registers_only"]["permutations_per_s"]
        result["roofline_hash"] = {
            "bound": "valu-issue", "kernel": "merkle.leaves_colmajor", "unit": "permutations/s",
            "achieved": rate, "peak": peak, "frac": rate / peak if rate else None,
            "frac_same_occupancy": rate / pr["four_waves"]["permutations_per_s"] if rate else None,
            "peak_simd_cycles_per_wave_permutation": pr["registers_only"]["simd_cycles_per_wave_permutation"],
            "simd_cycles_per_s": SIMD_CYCLES_PER_S,
            "valu_instr_per_permutation": v["valu_instr_per_permutation"],
            "traffic": tp["kernels"]["merkle.leaves_colmajor"]["bytes"],
            "method": pr["method"]}
    if intt:
        result["intt_ms"] = sum(v["ms"] for k, v in summ["passes"].items() if k.startswith("intt"))
    else:
        torch.cuda.empty_cache()
        if not a.no_large:
            lrun = commit_run(Z, 400, 18, 3, 4, a.large_steps, 1, True, with_plain=False)
            large = commit_summary(lrun, 400, 18, 3, 4, a.large_steps, True)
            large["workload"] = ("large commit: 400 columns x 2^18 evaluations, iNTT to coefficients, coset LDE "
                                 "blowup 8 -> 2^21 rows, Poseidon Merkle cap_height 4; compute_batched_ntt "
                                 "(inverse) + zk_commit (LDE + leaves + digests + cap)")
            large["steps"] = a.large_steps
            del large["kernels"]
            result["large"] = large
        torch.cuda.empty_cache()
        # the plain commit's rates, for the batched and served work estimates
        lde_pass_bytes = sum(W * C * ((1 << lg_n) + N) if k == "lde.pass1" else 2 * W * C * N
                             for k in summ["passes"] if k.startswith("lde"))
        rates = {"perm_per_s": summ["merkle"]["permutations_per_s"],
                 "lde_Bps": lde_pass_bytes / (summ["lde"]["lde_ms"] * 1e-3)}
        if not a.no_batched:
            result["batched"] = run_batched(Z, rates)
            result["batched_vs_plain_speedup"] = result["batched"]["speedup"]
        torch.cuda.empty_cache()
        if not a.no_served:
            result["served"] = run_served(Z, a.served_requests, rates)
            result["p99_latency_ms"] = result["served"]["poisson"]["served_p99_ms"]
    if not a.no_cpu:
        result["cpu_baseline"] = cpu_baseline(C, a.cpu_lg_n, r, h, a.cpu_reps)
    print(json.dumps(result))


if __name__ == "__main__":
    main()
