"""Autotuner of the NTT plans of libzeknox.so (one MI355X).

For every transform size 2^lg it runs the coset LDE of the commit path (135
columns, blowup 8: 2^(lg-3) coefficients per column -> 2^lg rows) under each
candidate plan -- the split of the lg bits into 1..3 HBM passes of at most 10
bits, the groups per workgroup of the passes (the tile width: more groups make
longer contiguous row runs and use more LDS), and for a 10-bit pass the
16-element kernel (512 threads, 8 groups, a 16 x 4 x 16 split) or the
32-element kernel (two 32-point register DITs, more LDS per group) -- checks
that the output equals the default plan's bit for bit, times the passes with
the library's per-kernel HIP-event timings (median of --reps runs), and writes
the fastest plan of each size to the table the library loads when a device is
first used (ntt_tuning_gfx950.txt beside libzeknox.so).  A size whose best
plan is not at least --min-gain faster than the default keeps the default.

    python -m zeknox_amd.tune [--lg 11-24] [--columns 135] [--reps 5] [--out PATH]

Replaces the reference's fixed stage structure (okx/zeknox
native/ntt/ntt.cu:106-158: lg - 10 global radix-2 stages plus one 10-stage
shared-memory pass), which has nothing to tune.
"""
import argparse
import itertools
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
DEFAULT_OUT = os.path.join(HERE, "ntt_tuning_gfx950.txt")
RATE_BITS = 3
GBITS = (-1, 4, 3, 2, 1)  # -1: the most the pass allows (the library caps the others at that)


def default_bits(lg):
    """The library's default split (ntt_plan_bits): ceil(lg / 10) near-equal passes."""
    np_ = 1 if lg <= 10 else (2 if lg <= 20 else 3)
    bits, left = [], lg
    for p in range(np_):
        b = (left + (np_ - p) - 1) // (np_ - p)
        bits.append(b)
        left -= b
    return tuple(bits)


def splits(lg):
    """Pass splits to try: the fewest passes that fit and one more (at most 3),
    every pass of 3..10 bits (fewer bits only for tiny transforms)."""
    lo = 1 if lg < 6 else 3
    out = []
    nmin = (lg + 9) // 10
    for np_ in range(nmin, min(3, nmin + 1) + 1):
        for parts in itertools.product(range(lo, 11), repeat=np_):
            if sum(parts) == lg:
                out.append(parts)
    return out


def candidates(lg):
    seen = set()
    for bits in splits(lg):
        k16_opts = [(1,) * len(bits)]
        if 10 in bits:
            k16_opts.append(tuple(0 if b == 10 else 1 for b in bits))
        for g in GBITS:
            for k16 in k16_opts:
                plan = (bits, (g,) * len(bits), k16)
                if plan not in seen:
                    seen.add(plan)
                    yield plan


def lde_once(Z, buf, lg, C):
    cfg = Z.NTTConfig(batches=C, extension_rate_bits=RATE_BITS, with_coset=True, are_inputs_on_device=True,
                      are_outputs_on_device=True)
    Z.lde_batch(0, buf["lde"], buf["coeffs"], lg - RATE_BITS, cfg)


def time_plan(Z, buf, lg, C, reps):
    """Median over reps of the summed LDE pass kernel times (ms)."""
    import torch
    lde_once(Z, buf, lg, C)  # warm-up: kernel attributes, twiddle tables of the plan
    times = []
    for _ in range(reps):
        torch.cuda.synchronize()
        Z.profile_reset()
        Z.profile_enable(True)
        lde_once(Z, buf, lg, C)
        torch.cuda.synchronize()
        Z.profile_enable(False)
        prof = Z.profile_read()
        times.append(sum(v[0] for k, v in prof.items() if k.startswith("lde.pass")))
    times.sort()
    return times[len(times) // 2]


def tune(lgs, C, reps, min_gain, log=sys.stderr):
    import torch
    # torch's HIP runtime first (as the tests and bench.py do), then the library's
    assert torch.cuda.is_available(), "the tuner needs a HIP device"
    import zeknox_amd as Z
    Z.init_cuda_rs()
    results = []
    for lg in lgs:
        n, N = 1 << (lg - RATE_BITS), 1 << lg
        coeffs = torch.empty(C * n, dtype=torch.int64, device="cuda")
        Z.fill_random(0, coeffs, C * n, 77 + lg)
        buf = {"coeffs": coeffs, "lde": torch.empty(C * N, dtype=torch.int64, device="cuda")}
        Z.ntt_tune_set(lg)  # the default plan
        t_def = time_plan(Z, buf, lg, C, reps)
        ref = buf["lde"].clone()
        timed = []
        for bits, gbits, k16 in candidates(lg):
            Z.ntt_tune_set(lg, bits, gbits, k16)
            t = time_plan(Z, buf, lg, C, reps)
            if not torch.equal(buf["lde"], ref):
                Z.ntt_tune_set(lg)
                raise RuntimeError("plan %s %s %s of 2^%d differs from the default plan" % (bits, gbits, k16, lg))
            timed.append((t, (bits, gbits, k16)))
            if len(timed) % 50 == 0:
                print("2^%d: %d plans timed" % (lg, len(timed)), file=log, flush=True)
        # the fastest few against the default again, round-robin, so that
        # clock and cache states affect them alike: best of three medians each
        timed.sort(key=lambda x: x[0])
        finalists = [None] + [p for _, p in timed[:6]]
        t_fin = {p: float("inf") for p in finalists}
        for _ in range(3):
            for plan in finalists:
                Z.ntt_tune_set(lg) if plan is None else Z.ntt_tune_set(lg, *plan)
                t_fin[plan] = min(t_fin[plan], time_plan(Z, buf, lg, C, reps))
        Z.ntt_tune_set(lg)
        t_def = t_fin[None]
        best = min(((t, p) for p, t in t_fin.items() if p is not None), default=(t_def, None))
        chosen = best[1] if best[1] is not None and best[0] < t_def * (1 - min_gain) else None
        tried = len(timed)
        results.append((lg, t_def, best[0], chosen))
        print("2^%d: default %s %.4f ms, best %s %.4f ms over %d plans -> %s" % (
            lg, default_bits(lg), t_def, best[1], best[0], tried, "tuned" if chosen else "default"),
            file=log, flush=True)
        del buf, ref, coeffs
        torch.cuda.empty_cache()
    return results


def write_table(results, path, C, reps):
    lines = ["# NTT plans chosen by zeknox_amd/tune.py on an MI355X (gfx950): coset LDE of %d columns, blowup %d,"
             % (C, 1 << RATE_BITS),
             "# pass kernels timed with HIP events, median of %d runs; the six fastest plans re-timed against the"
             % reps,
             "# default three times round-robin (best median each); written %s." % time.strftime("%Y-%m-%d"),
             "# lg np b1 b2 b3 g1 g2 g3 k1 k2 k3   (b: bits per pass, g: log2 groups per workgroup, -1 the most",
             "# the pass allows, k: 1 = the 16-element kernel may run a 10-bit pass)"]
    for lg, t_def, t_best, chosen in results:
        if chosen is None:
            lines.append("# %d default %s %.4f ms (no plan measured %.0f%% faster)" % (
                lg, default_bits(lg), t_def, 100 * 0.0 if t_best >= t_def else 100 * (1 - t_best / t_def)))
            continue
        bits, gbits, k16 = chosen
        pad = lambda v, f: list(v) + [f] * (3 - len(v))  # noqa: E731
        lines.append("%d %d %s %s %s   # %.4f ms, default %s %.4f ms" % (
            lg, len(bits), " ".join(map(str, pad(bits, 0))), " ".join(map(str, pad(gbits, -1))),
            " ".join(map(str, pad(k16, 0))), t_best, default_bits(lg), t_def))
    with open(path + ".tmp", "w") as f:
        f.write("\n".join(lines) + "\n")
    os.replace(path + ".tmp", path)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lg", default="11-24", help="sizes: a-b or a comma list")
    ap.add_argument("--columns", type=int, default=135)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--min-gain", type=float, default=0.02)
    ap.add_argument("--out", default=DEFAULT_OUT)
    a = ap.parse_args()
    if "-" in a.lg:
        lo, hi = map(int, a.lg.split("-"))
        lgs = list(range(lo, hi + 1))
    else:
        lgs = [int(x) for x in a.lg.split(",")]
    if any(lg <= RATE_BITS or lg > 30 for lg in lgs):
        sys.exit("sizes must lie in %d..30" % (RATE_BITS + 1))
    results = tune(lgs, a.columns, a.reps, a.min_gain)
    write_table(results, a.out, a.columns, a.reps)
    print("wrote", a.out)


if __name__ == "__main__":
    main()
