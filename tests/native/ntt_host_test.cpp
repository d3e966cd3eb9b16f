// Host emulation of the NTT pass kernels (zeknox_amd/csrc/ntt.hip): the same
// per-thread step functions the kernels run, called for every thread of every
// tile in kernel order (a step's threads all finish before the next step, as
// the kernels' barriers make them), with the workgroup's LDS as a host array.
// The plan (pass split, tiles, blocked intermediate, twiddle tables) is the one
// ntt_launch makes.  Built with hipcc for the host only (no device code runs);
// reads the input columns from a file and writes the transform to another, for
// tests/test_host_ntt.py to compare with the oracle.
//
// usage: ntt_host_test lg lg_in ncols inverse coset force_exact table in.bin out.bin
//        [np b1 b2 b3 g1 g2 g3 r1 r2 r3]  (a tuned plan: pass bits, groups per
//        workgroup (log2, < 0 auto), 16-element kernel allowed per pass)
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../zeknox_amd/csrc/ntt_pass.hpp"

using namespace zk;

// kernel-order emulation of one 10-bit pass (ntt_r10_kernel)
template <bool INV>
static void run_r10(const PassArgs &a, u64 ntiles)
{
    std::vector<u64> lds(kR10Lds / sizeof(u64));
    for (u64 b = 0; b < ntiles; b++) {
        const u64 tile = xcd_tile(b, ntiles);
        u32 col;
        u64 jhi, fast0;
        Pass<10>::tile_decode(a, tile, col, jhi, fast0);
        const bool last = (a.S + 10) == a.L;
        std::vector<u64> out(16 * kR10Threads);
        bool any = false;
        std::vector<char> rare(kR10Threads, (a.flags & F_FORCE_EXACT) != 0);
        for (u32 t = 0; t < kR10Threads; t++) {
            bool r = rare[t];
            r10_step_a<INV, false>(a, col, jhi, fast0, t, lds.data(), r);
            rare[t] = r;
        }
        for (u32 t = 0; t < kR10Threads; t++) {
            bool r = rare[t];
            r10_step_b<INV, false>(a, t, lds.data(), r);
            rare[t] = r;
        }
        for (u32 t = 0; t < kR10Threads; t++) {
            bool r = rare[t];
            r10_step_c<INV, false>(a, t, lds.data(), last, &out[16 * t], r);
            any |= r;
        }
        if (any) {
            bool r = false;
            for (u32 t = 0; t < kR10Threads; t++) r10_step_a<INV, true>(a, col, jhi, fast0, t, lds.data(), r);
            for (u32 t = 0; t < kR10Threads; t++) r10_step_b<INV, true>(a, t, lds.data(), r);
            for (u32 t = 0; t < kR10Threads; t++) r10_step_c<INV, true>(a, t, lds.data(), last, &out[16 * t], r);
        }
        for (u32 t = 0; t < kR10Threads; t++) r10_store(a, col, jhi, fast0, t, last, &out[16 * t]);
    }
}

// kernel-order emulation of one <= 10-bit pass of ntt_fast_kernel
template <int B, bool INV>
static void run_fast(const PassArgs &a, u64 ntiles)
{
    constexpr int BA = (B + 1) / 2, BB = B / 2;
    std::vector<u64> lds(fast_lds_bytes<B>(1u << a.Gbits) / sizeof(u64) + 1);
    for (u64 b = 0; b < ntiles; b++) {
        const u64 tile = xcd_tile(b, ntiles);
        for (u32 t = 0; t < kThreads; t++) fast_phase<BA, BB, INV>(a, tile, 0, t, kThreads, lds.data());
        for (u32 t = 0; t < kThreads; t++) fast_phase<BA, BB, INV>(a, tile, 1, t, kThreads, lds.data());
    }
}

template <bool INV>
static void run_fast_b(u32 B, const PassArgs &a, u64 ntiles)
{
    switch (B) {
    case 1: run_fast<1, INV>(a, ntiles); break;
    case 2: run_fast<2, INV>(a, ntiles); break;
    case 3: run_fast<3, INV>(a, ntiles); break;
    case 4: run_fast<4, INV>(a, ntiles); break;
    case 5: run_fast<5, INV>(a, ntiles); break;
    case 6: run_fast<6, INV>(a, ntiles); break;
    case 7: run_fast<7, INV>(a, ntiles); break;
    case 8: run_fast<8, INV>(a, ntiles); break;
    case 9: run_fast<9, INV>(a, ntiles); break;
    case 10: run_fast<10, INV>(a, ntiles); break;
    default: fprintf(stderr, "pass of %u bits not emulated\n", B); exit(2);
    }
}

static void fill_powers(u64 x, u64 *row, int n)
{
    u64 v = 1;
    for (int i = 0; i < n; i++) {
        row[i] = v;
        v = gl_mul(v, x);
    }
}

int main(int argc, char **argv)
{
    if (argc != 10 && argc != 20) {
        fprintf(stderr, "usage: %s lg lg_in ncols inverse coset force_exact table in.bin out.bin "
                        "[np b1 b2 b3 g1 g2 g3 r1 r2 r3]\n", argv[0]);
        return 2;
    }
    const u32 lg = atoi(argv[1]), lg_in = atoi(argv[2]), ncols = atoi(argv[3]);
    const bool inv = atoi(argv[4]) != 0, coset = atoi(argv[5]) != 0, force = atoi(argv[6]) != 0;
    const bool table = atoi(argv[7]) != 0;
    const u64 N = 1ull << lg, n = 1ull << lg_in;
    std::vector<u64> in(ncols * n), out(ncols * N, 0xDEADBEEFull), scratch(ncols * N);
    FILE *f = fopen(argv[8], "rb");
    if (!f || fread(in.data(), 8, in.size(), f) != in.size()) return 3;
    fclose(f);

    // the device context's tables, on the host
    std::vector<u64> pw((size_t)(kMaxLg + 1) * kPowLen), gpw(3 * kPowLen), tb(16 * 64);
    for (int L = 0; L <= kMaxLg; L++) {
        u64 w = gl_omega((u32)L);
        if (inv) w = gl_inv(w);
        fill_powers(w, pw.data() + (size_t)L * kPowLen, kPowLen);
    }
    u64 base = 7;
    for (int k = 0; k < 3; k++) {
        fill_powers(base, gpw.data() + k * kPowLen, kPowLen);
        base = gl_pow(base, (u64)kPowLen);
    }
    const u64 w1024 = pw[(size_t)10 * kPowLen + 1];
    for (u32 v = 0; v < 16; v++) fill_powers(gl_pow(w1024, zk_bitrev(v, 4)), tb.data() + 64 * v, 64);

    NttJob job;
    job.src = in.data();
    job.dst = out.data();
    job.src_stride = n;
    job.dst_stride = N;
    job.ncols = ncols;
    job.lg = lg;
    job.lg_in = lg_in;
    job.inverse = inv;
    job.coset_in = coset && !inv;
    job.coset_out = coset && inv;
    job.scale_out = inv;
    if (argc > 10) {  // a tuned plan: np b1 b2 b3 g1 g2 g3 r1 r2 r3 (g < 0: auto)
        NttTune t{};
        t.np = (u32)atoi(argv[10]);
        for (int p = 0; p < 3; p++) {
            t.bits[p] = (u32)atoi(argv[11 + p]);
            const int g = atoi(argv[14 + p]);
            t.gbits[p] = g < 0 ? kTuneAuto : (u32)g;
            t.r10[p] = (u32)atoi(argv[17 + p]);
        }
        if (!ntt_tune_valid(lg, t)) return 6;
        ntt_tune_table()[lg] = t;
    }
    u32 bits[4];
    const int np = ntt_plan_bits(lg, bits);
    const bool blocked = np == 2 && bits[0] == 10 && bits[1] == 10;
    // ZK_HOST_IN_DST=1: later passes run in the output buffer, as an
    // out-of-place transform does on the device without a scratch buffer
    const char *in_dst = getenv("ZK_HOST_IN_DST");
    const bool use_scratch = np > 1 && !(in_dst && in_dst[0] == '1');
    auto passes = plan_passes(job, pw.data(), gpw.data(), use_scratch ? scratch.data() : nullptr, 66560);
    std::vector<std::vector<u64>> tw_tables;
    for (PassPlan &pp : passes) {
        pp.a.tb = tb.data();
        if (force) pp.a.flags |= F_FORCE_EXACT;
        if (pp.fast && pp.a.S > 0 && table) {
            const u64 cnt = 1ull << (pp.a.L - pp.a.Sprev);
            tw_tables.emplace_back(cnt);
            for (u64 i = 0; i < cnt; i++)
                tw_tables.back()[i] = twiddle_entry(pw.data(), pp.a.L, pp.a.S, pp.a.Sprev, i);
            pp.a.tw = tw_tables.back().data();
        }
    }
    if (blocked && passes[0].r10 && passes[1].r10 &&
        passes[0].a.Gbits == 3 && passes[1].a.Gbits == 3 && passes[1].a.Jbits == 0) {
        passes[0].a.flags |= F_BLOCK_OUT;
        passes[1].a.flags |= F_BLOCK_IN;
    }
    for (const PassPlan &pp : passes) {
        if (!pp.fast) {
            fprintf(stderr, "generic passes are not emulated\n");
            return 2;
        }
        if (pp.r10) {
            inv ? run_r10<true>(pp.a, pp.ntiles) : run_r10<false>(pp.a, pp.ntiles);
        } else {
            inv ? run_fast_b<true>(pp.B, pp.a, pp.ntiles) : run_fast_b<false>(pp.B, pp.a, pp.ntiles);
        }
        printf("pass S=%u B=%u tiles=%llu r10=%d flags=%u\n", pp.a.S, pp.B, (unsigned long long)pp.ntiles,
               (int)pp.r10, pp.a.flags);
    }
    f = fopen(argv[9], "wb");
    if (!f || fwrite(out.data(), 8, out.size(), f) != out.size()) return 4;
    fclose(f);
    return 0;
}
