// Batched Goldilocks NTT / coset LDE / transpose kernels for gfx950.
//
// Algorithm (replaces okx/zeknox native/ntt/ntt.cu:106-158 and the per-stage
// kernels of native/ntt/kernels.cu:160-392, which make one HBM pass per
// radix-2 stage plus separate extension, coset, bit-reverse passes):
//
// An N = 2^L point transform of a column is computed by Cooley-Tukey with
// bit-reversed input and natural output, split into P <= 3 HBM passes of
// B_1 + ... + B_P = L bits (four-step decomposition).  Pass p owns bit range
// [S, S+B) of the index: it loads groups of 2^B elements at stride 2^S,
// multiplies them by the inter-pass twiddle (omega of the sub-problem raised to
// rev(J) * b'), runs B radix-2 stages in LDS with radix-16/8 register blocks,
// and stores the group back.  Pass 1 reads the caller's natural-order input
// at bit-reversed positions directly (so no separate bit-reverse pass), with
// zero-extension (LDE) and the coset scaling g^n fused into the load; the
// last pass fuses the 2^-L scaling (inverse).  Output is natural order,
// identical to the reference's forward/inverse batch NTT and batch LDE.
//
// Each workgroup processes G adjacent groups so every HBM row access is a
// contiguous run of G * 8 bytes (pass >= 2) or a whole group (pass 1 stores).
#include <dlfcn.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <cstdlib>
#include <string>
#include <vector>

// per-thread code of the passes, the pass plan
#include "ntt_pass.hpp"

namespace zk {

// K radix-2 DIT stages s0 .. s0+K-1 on 2^K register values holding group
// elements j0 + m * 2^s0 (m < 2^K), j0 < 2^s0.  Stage s pairs (j, j + 2^s)
// with twiddle omega_{2^(s+1)}^(j mod 2^s) = omega_{2^11}^((j mod 2^s) << (10-s)).
template <int K>
ZK_DEV void dit_regs(u64 v[1 << K], u32 s0, u32 j0, const u64 *w11)
{
#pragma unroll
    for (int t = 0; t < K; t++) {
        const u32 s = s0 + t;
#pragma unroll
        for (int m = 0; m < (1 << K); m++) {
            if (m & (1 << t)) continue;
            const u32 e = j0 + ((u32)(m & ((1 << t) - 1)) << s0);
            u64 y = v[m + (1 << t)];
            if (e) y = gl_mul(y, w11[e << (10 - s)]);
            const u64 x = v[m];
            v[m] = gl_add(x, y);
            v[m + (1 << t)] = gl_sub(x, y);
        }
    }
}

template <int B>
__global__ __launch_bounds__(kThreads) void ntt_pass_kernel(PassArgs a)
{
    extern __shared__ u64 lds[];
    using P = Pass<B>;
    constexpr int NBLK = nblocks_for(B);
    const u32 G = 1u << a.Gbits;
    const u64 *w11 = a.pw + (size_t)11 * kPowLen;
    u32 col;
    u64 jhi, fast0;
    P::tile_decode(a, a.tile_base + xcd_tile(blockIdx.x, gridDim.x), col, jhi, fast0);
    const bool gfast_store = a.S > 0;

    if constexpr (NBLK == 1) {
        // whole group in registers: one set per group
        constexpr int K = B;
        for (u32 g = threadIdx.x; g < G; g += kThreads) {
            u64 v[1 << K];
#pragma unroll
            for (int m = 0; m < (1 << K); m++) v[m] = P::load(a, col, jhi, fast0, g, (u32)m);
            dit_regs<K>(v, 0, 0, w11);
#pragma unroll
            for (int m = 0; m < (1 << K); m++) P::store(a, col, jhi, fast0, g, (u32)m, v[m]);
        }
        return;
    } else {
        // block 0: global -> registers -> LDS, g-fast mapping (coalesced loads)
        {
            constexpr int K = block_k(B, 0);
            const u32 nsets = G << (B - K);
            for (u32 set = threadIdx.x; set < nsets; set += kThreads) {
                const u32 g = set & (G - 1), r = set >> a.Gbits;
                const u32 j0 = r << K;
                u64 v[1 << K];
#pragma unroll
                for (int m = 0; m < (1 << K); m++) v[m] = P::load(a, col, jhi, fast0, g, j0 + m);
                dit_regs<K>(v, 0, 0, w11);
#pragma unroll
                for (int m = 0; m < (1 << K); m++) lds[g * a.pitch + j0 + m] = v[m];
            }
        }
        __syncthreads();
        // middle blocks: LDS -> registers -> LDS, r-fast mapping
#pragma unroll
        for (int bi = 1; bi < NBLK - 1; bi++) {
            const int K = block_k(B, bi);
            const u32 s0 = block_s0(B, bi);
            const u32 nsets = G << (B - K);
            for (u32 set = threadIdx.x; set < nsets; set += kThreads) {
                const u32 r = set & ((1u << (B - K)) - 1), g = set >> (B - K);
                const u32 jl = r & ((1u << s0) - 1);
                const u32 j0 = jl | ((r >> s0) << (s0 + K));
                if (K == 4) {
                    u64 v[16];
#pragma unroll
                    for (int m = 0; m < 16; m++) v[m] = lds[g * a.pitch + j0 + (m << s0)];
                    dit_regs<4>(v, s0, jl, w11);
#pragma unroll
                    for (int m = 0; m < 16; m++) lds[g * a.pitch + j0 + (m << s0)] = v[m];
                } else {
                    u64 v[8];
#pragma unroll
                    for (int m = 0; m < 8; m++) v[m] = lds[g * a.pitch + j0 + (m << s0)];
                    dit_regs<3>(v, s0, jl, w11);
#pragma unroll
                    for (int m = 0; m < 8; m++) lds[g * a.pitch + j0 + (m << s0)] = v[m];
                }
            }
            __syncthreads();
        }
        // last block: LDS -> registers -> global
        {
            constexpr int K = block_k(B, NBLK - 1);
            constexpr u32 s0 = block_s0(B, NBLK - 1);
            const u32 nsets = G << (B - K);
            for (u32 set = threadIdx.x; set < nsets; set += kThreads) {
                u32 g, r;
                if (gfast_store) {
                    g = set & (G - 1);
                    r = set >> a.Gbits;
                } else {
                    r = set & ((1u << (B - K)) - 1);
                    g = set >> (B - K);
                }
                const u32 jl = r & ((1u << s0) - 1);
                const u32 j0 = jl | ((r >> s0) << (s0 + K));
                u64 v[1 << K];
#pragma unroll
                for (int m = 0; m < (1 << K); m++) v[m] = lds[g * a.pitch + j0 + (m << s0)];
                dit_regs<K>(v, s0, jl, w11);
#pragma unroll
                for (int m = 0; m < (1 << K); m++) P::store(a, col, jhi, fast0, g, j0 + (m << s0), v[m]);
            }
        }
    }
}

template <int BA, int BB, bool INV>
__global__ __launch_bounds__(kThreads) void ntt_fast_kernel(PassArgs a)
{
    extern __shared__ u64 lds[];
    const u64 tile = a.tile_base + xcd_tile(blockIdx.x, gridDim.x);
    fast_phase<BA, BB, INV>(a, tile, 0, threadIdx.x, kThreads, lds);
    __syncthreads();
    fast_phase<BA, BB, INV>(a, tile, 1, threadIdx.x, kThreads, lds);
}

template <bool INV>
__global__ __launch_bounds__(kR10Threads, 4) void ntt_r10_kernel(PassArgs a)
{
    extern __shared__ u64 lds[];
    u32 *votes = reinterpret_cast<u32 *>(lds + kR10Groups * kR10GP);
    const u64 tile = a.tile_base + xcd_tile(blockIdx.x, gridDim.x);
    u32 col;
    u64 jhi, fast0;
    Pass<10>::tile_decode(a, tile, col, jhi, fast0);
    const bool last = (a.S + 10) == a.L;
    const u32 tid = threadIdx.x;
    u64 out[16];
    bool rare = (a.flags & F_FORCE_EXACT) != 0;
    r10_step_a<INV, false>(a, col, jhi, fast0, tid, lds, rare);
    __syncthreads();
    r10_step_b<INV, false>(a, tid, lds, rare);
    __syncthreads();
    r10_step_c<INV, false>(a, tid, lds, last, out, rare);
    // one vote per tile: each wave writes whether any of its lanes met a rare case
    const bool wave_rare = __builtin_amdgcn_ballot_w64(rare) != 0;
    if ((tid & 63) == 0) votes[tid >> 6] = wave_rare;
    __syncthreads();
    bool redo = false;
#pragma unroll
    for (u32 w = 0; w < kR10Threads / 64; w++) redo |= votes[w] != 0;
    if (redo) {
        r10_step_a<INV, true>(a, col, jhi, fast0, tid, lds, rare);
        __syncthreads();
        r10_step_b<INV, true>(a, tid, lds, rare);
        __syncthreads();
        r10_step_c<INV, true>(a, tid, lds, last, out, rare);
    }
    r10_store(a, col, jhi, fast0, tid, last, out);
}

// Launches `kernel` over tiles [0, ntiles) in chunks whose grids stay below
// 2^32 threads; each chunk passes its first tile in a.tile_base.
template <typename K>
static void launch_tiles(K kernel, PassArgs a, u64 ntiles, u32 threads, size_t lds, hipStream_t st)
{
    const u64 max_blocks = (((u64)1 << 32) - 1) / threads;
    for (u64 t0 = 0; t0 < ntiles; t0 += max_blocks) {
        a.tile_base = t0;
        const u64 nb = std::min<u64>(max_blocks, ntiles - t0);
        hipLaunchKernelGGL(kernel, dim3((u32)nb), dim3(threads), lds, st, a);
        ZK_HIP(hipGetLastError());
    }
}

template <bool INV>
static void launch_r10(const PassArgs &a, u64 ntiles, hipStream_t st)
{
    ZK_HIP(hipFuncSetAttribute((const void *)ntt_r10_kernel<INV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)kR10Lds));
    launch_tiles(ntt_r10_kernel<INV>, a, ntiles, kR10Threads, kR10Lds, st);
}

static void launch_f10(const PassPlan &pp, bool inv, hipStream_t st)
{
    inv ? launch_r10<true>(pp.a, pp.ntiles, st) : launch_r10<false>(pp.a, pp.ntiles, st);
}

template <int B, bool INV>
static void launch_fast(const PassArgs &a, u64 ntiles, hipStream_t st)
{
    constexpr int BA = (B + 1) / 2, BB = B / 2;
    const size_t lds = fast_lds_bytes<B>(1u << a.Gbits);
    if (lds > 65536)
        ZK_HIP(hipFuncSetAttribute((const void *)ntt_fast_kernel<BA, BB, INV>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    launch_tiles(ntt_fast_kernel<BA, BB, INV>, a, ntiles, kThreads, lds, st);
}

static void dispatch_fast(u32 B, bool inv, const PassArgs &a, u64 ntiles, hipStream_t st)
{
    switch (B * 2 + (inv ? 1 : 0)) {
#define ZK_FCASE(b)                                    \
    case 2 * b: launch_fast<b, false>(a, ntiles, st); break; \
    case 2 * b + 1: launch_fast<b, true>(a, ntiles, st); break;
        ZK_FCASE(1) ZK_FCASE(2) ZK_FCASE(3) ZK_FCASE(4) ZK_FCASE(5)
        ZK_FCASE(6) ZK_FCASE(7) ZK_FCASE(8) ZK_FCASE(9) ZK_FCASE(10)
#undef ZK_FCASE
    default: throw Error((int)hipErrorInvalidValue, "unsupported NTT pass size");
    }
}

int ntt_plan(u32 lg, u32 bits[4]) { return ntt_plan_bits(lg, bits); }

bool ntt_tune_set(u32 lg, u32 np, const u32 bits[3], const int32_t gbits[3], const u32 r10[3])
{
    if (lg == 0 || lg > 30) return false;
    NttTune t{};
    t.np = np;
    for (u32 p = 0; p < 3 && p < np; p++) {
        t.bits[p] = bits[p];
        t.gbits[p] = gbits[p] < 0 ? kTuneAuto : (u32)gbits[p];
        t.r10[p] = r10[p];
    }
    if (np && !ntt_tune_valid(lg, t)) return false;
    std::lock_guard<std::mutex> lk(ntt_tune_mutex());
    ntt_tune_table()[lg] = t;
    return true;
}

void ntt_tune_read(u32 lg, u32 &np, u32 bits[3], int32_t gbits[3], u32 r10[3])
{
    const NttTune t = ntt_tune_get(lg);
    np = t.np;
    for (int p = 0; p < 3; p++) {
        bits[p] = t.bits[p];
        gbits[p] = t.gbits[p] == kTuneAuto ? -1 : (int32_t)t.gbits[p];
        r10[p] = t.r10[p];
    }
}

static void ntt_tune_load(const std::string &path, bool named)
{
    FILE *f = fopen(path.c_str(), "r");
    if (!f) {
        if (named) fprintf(stderr, "zeknox: cannot open NTT tuning file %s\n", path.c_str());
        return;
    }
    char line[512];
    int lineno = 0;
    while (fgets(line, sizeof line, f)) {
        lineno++;
        if (char *c = strchr(line, '#')) *c = 0;
        int v[11];
        const int n = sscanf(line, "%d %d %d %d %d %d %d %d %d %d %d", &v[0], &v[1], &v[2], &v[3], &v[4], &v[5],
                             &v[6], &v[7], &v[8], &v[9], &v[10]);
        if (n <= 0) continue;
        bool ok = n == 11 && v[0] > 0 && v[1] >= 0;
        for (int i = 2; ok && i < 11; i++) ok = i >= 5 && i < 8 ? v[i] >= -1 : v[i] >= 0;
        if (ok) {
            const u32 bits[3] = {(u32)v[2], (u32)v[3], (u32)v[4]}, r10[3] = {(u32)v[8], (u32)v[9], (u32)v[10]};
            const int32_t gb[3] = {v[5], v[6], v[7]};
            ok = ntt_tune_set((u32)v[0], (u32)v[1], bits, gb, r10);
        }
        if (!ok) fprintf(stderr, "zeknox: %s:%d: not a runnable NTT plan, skipped\n", path.c_str(), lineno);
    }
    fclose(f);
}

void ntt_tune_load_once()
{
    static std::once_flag once;
    std::call_once(once, [] {
        const char *env = getenv("ZK_NTT_TUNING");
        if (env && env[0]) {
            ntt_tune_load(env, true);
            return;
        }
        Dl_info info;
        if (dladdr((const void *)&ntt_tune_load_once, &info) && info.dli_fname) {
            std::string dir(info.dli_fname);
            const size_t slash = dir.find_last_of('/');
            dir = slash == std::string::npos ? std::string(".") : dir.substr(0, slash);
            ntt_tune_load(dir + "/ntt_tuning_gfx950.txt", false);
        }
    });
}

bool ntt_blocked_intermediate(u32 lg)
{
    u32 bits[4];
    return lg <= (u32)kMaxLg && ntt_plan_bits(lg, bits) == 2 && bits[0] == 10 && bits[1] == 10;
}

static const char *pass_name(const char *tag, int p)
{
    static const char *names[3][3] = {{"ntt.pass1", "ntt.pass2", "ntt.pass3"},
                                      {"intt.pass1", "intt.pass2", "intt.pass3"},
                                      {"lde.pass1", "lde.pass2", "lde.pass3"}};
    int t = tag[0] == 'i' ? 1 : (tag[0] == 'l' ? 2 : 0);
    return names[t][p];
}

template <int B>
static void launch_pass(const PassArgs &a, u64 ntiles, hipStream_t st)
{
    size_t lds = nblocks_for(B) > 1 ? ((size_t)(1u << a.Gbits) * a.pitch * sizeof(u64)) : 0;
    if (lds > 65536) ZK_HIP(hipFuncSetAttribute((const void *)ntt_pass_kernel<B>,
                                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    launch_tiles(ntt_pass_kernel<B>, a, ntiles, kThreads, lds, st);
}

static void dispatch_pass(u32 B, const PassArgs &a, u64 ntiles, hipStream_t st)
{
    switch (B) {
#define ZK_CASE(b) \
    case b: launch_pass<b>(a, ntiles, st); break;
        ZK_CASE(1) ZK_CASE(2) ZK_CASE(3) ZK_CASE(4) ZK_CASE(5) ZK_CASE(6)
        ZK_CASE(7) ZK_CASE(8) ZK_CASE(9) ZK_CASE(10) ZK_CASE(11)
#undef ZK_CASE
    default: throw Error((int)hipErrorInvalidValue, "unsupported NTT pass size");
    }
}

__global__ __launch_bounds__(256) void twiddle_table_kernel(u64 *out, const u64 *pw, u32 L, u32 S, u32 Sprev, u64 n)
{
    const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = twiddle_entry(pw, L, S, Sprev, i);
}

constexpr u32 kMaxTwiddleTableBits = 22;  // 32 MiB per table at most

// The table for (direction, L, S, Sprev), built on `st` at first use and kept
// for the process; null (the pass computes its twiddles) when it would exceed
// 2^kMaxTwiddleTableBits entries, would have to be built while `st` is being
// captured into a graph, or does not exist yet and `build` is false.
static const u64 *twiddle_table(DeviceCtx &c, bool inv, u32 L, u32 S, u32 Sprev, hipStream_t st, bool build)
{
    if (L - Sprev > kMaxTwiddleTableBits) return nullptr;
    const uint64_t key = ((uint64_t)inv << 24) | ((uint64_t)L << 16) | ((uint64_t)S << 8) | Sprev;
    std::lock_guard<std::mutex> lk(c.tw_mu);
    auto it = c.tw_tables.find(key);
    if (it != c.tw_tables.end()) return it->second;
    if (!build) return nullptr;
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    ZK_HIP(hipStreamIsCapturing(st, &cap));
    if (cap != hipStreamCaptureStatusNone) return nullptr;
    const u64 n = 1ull << (L - Sprev);
    u64 *t = nullptr;
    ZK_HIP(hipMalloc(&t, n * sizeof(u64)));
    hipLaunchKernelGGL(twiddle_table_kernel, dim3((u32)((n + 255) / 256)), dim3(256), 0, st, t, c.pw[inv ? 1 : 0],
                       L, S, Sprev, n);
    ZK_HIP(hipGetLastError());
    ZK_HIP(hipStreamSynchronize(st));  // complete before any other stream can use it
    c.tw_tables.emplace(key, t);
    return t;
}

void ntt_launch(DeviceCtx &c, const NttJob &job, hipStream_t st)
{
    if (job.lg == 0 || job.lg > (u32)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
    if (job.ncols == 0) return;
    u32 bits[4];
    const int np = ntt_plan(job.lg, bits);
    // Two 10-bit passes (2^20 points) on a caller that may use the library
    // scratch: pass 1 writes the intermediate blocked -- the 2^10 group
    // elements of each 8 consecutive second-pass fast indices together, as
    // one contiguous 64 KiB block -- so that every tile of pass 2 reads one
    // block instead of a 64-byte piece of each of 1024 rows 8 KiB apart.
    bool blocked = job.use_scratch && np == 2 && bits[0] == 10 && bits[1] == 10;
    u64 *inter = nullptr;
    const size_t scratch_bytes = (size_t)job.ncols * ((size_t)1 << job.lg) * sizeof(u64);
    if (job.no_alloc) {
        if (job.src == job.dst && np > 1)
            throw Error((int)hipErrorInvalidValue, "invalid args: an in-place transform needs scratch memory");
        blocked = job.inter && np == 2 && bits[0] == 10 && bits[1] == 10;
        if (blocked && scratch_bytes > job.inter_bytes)
            throw Error((int)hipErrorInvalidValue, "invalid args: the intermediate exceeds the given scratch");
        if (blocked) inter = job.inter;
    } else if (job.src == job.dst && np > 1) {
        inter = (u64 *)c.get_scratch(scratch_bytes);
    } else if (blocked) {
        // optional: without room for the scratch the passes run in place
        try {
            inter = (u64 *)c.get_scratch(scratch_bytes);
        } catch (const Error &) {
            (void)hipGetLastError();
            blocked = false;
        }
    }
    auto passes = plan_passes(job, c.pw[job.inverse ? 1 : 0], job.gpw ? job.gpw : c.gpw, inter, c.lds_budget);
    for (PassPlan &pp : passes) {
        pp.a.tb = c.r10tb[job.inverse ? 1 : 0];
        if (pp.fast && pp.a.S > 0)
            pp.a.tw = twiddle_table(c, job.inverse, pp.a.L, pp.a.S, pp.a.Sprev, st, !job.no_alloc);
    }
    if (blocked && passes[0].r10 && passes[1].r10 &&
        passes[0].a.Gbits == 3 && passes[1].a.Gbits == 3 && passes[1].a.Jbits == 0) {
        passes[0].a.flags |= F_BLOCK_OUT;
        passes[1].a.flags |= F_BLOCK_IN;
    }
    for (size_t p = 0; p < passes.size(); p++) {
        const PassPlan &pp = passes[p];
        ProfScope ps(pass_name(job.prof_tag, (int)p), st);
        if (pp.r10)
            launch_f10(pp, job.inverse, st);
        else if (pp.fast)
            dispatch_fast(pp.B, job.inverse, pp.a, pp.ntiles, st);
        else
            dispatch_pass(pp.B, pp.a, pp.ntiles, st);
    }
}

// ---------------------------------------------------------------- transpose
// 64 x 64 tiles through LDS (padded rows: conflict-free column reads), for
// more than 447 columns or fewer than 16 rows: the
// loads read 512-byte column runs, the stores write 512-byte runs of output
// row rev(i) (reference: kernels.cu:94-135, which uses 32 x 32 tiles).
// Tiles are taken in XCD-contiguous order like the NTT passes.
constexpr int TI = 64, TJ = 64;

__global__ __launch_bounds__(256) void transpose_rev_kernel(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                            u32 lg, u64 ncols, u64 tiles_i, u64 tile_base)
{
    __shared__ u64 tile[TJ][TI + 1];
    const u64 N = 1ull << lg;
    const u64 b = tile_base + xcd_tile(blockIdx.x, gridDim.x);
    const u64 ti = b % tiles_i, tj = b / tiles_i;
    const u64 i0 = ti * TI, j0 = tj * TJ;
    const u32 tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < TJ / 4; k++) {
        const u64 j = j0 + ty + 4 * k, i = i0 + tx;
        if (j < ncols && i < N) tile[ty + 4 * k][tx] = in[j * N + i];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < TI / 4; k++) {
        const u64 i = i0 + ty + 4 * k, j = j0 + tx;
        if (j < ncols && i < N) out[(u64)zk_bitrev((u32)i, lg) * ncols + j] = tile[tx][ty + 4 * k];
    }
}

// Output rows written in pairs.  Workgroup w owns the base positions
// i = B w .. B w + B - 1 (all below N/2) and, with each, i + N/2: positions i
// and i + N/2 become output rows rev(i) and rev(i) + 1 (rev(i) is even), so the
// workgroup stages B pairs of whole rows (all columns) in LDS and writes each
// pair as one contiguous run of 2 * ncols elements: only the two ends of a
// pair can share a 64-byte sector with another workgroup's rows, where the 64
// x 64 tiles above leave every row in three runs (64 + 64 + 7 columns for 135
// columns), each with partly written sectors at both ends.  The loads read B
// consecutive positions (8 B bytes) of a column and half per lane group.
template <u32 B, u32 W>
__global__ __launch_bounds__(256) void transpose_rev_pairs_kernel(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                                  u32 lg, u32 ncols, u32 pitch, u64 block_base)
{
    extern __shared__ u64 lds[];  // [B][2][pitch]
    const u64 N = 1ull << lg, half = N >> 1;
    const u64 i0 = (block_base + xcd_tile(blockIdx.x, gridDim.x)) * B;
    const u32 tid = threadIdx.x;
    {
        // a lane loads W consecutive positions of a column (one 16-byte load
        // for W = 2, when `in` is 16-byte aligned), 16 columns in flight before
        // their LDS writes (a load-store loop would wait for each load in turn)
        struct alignas(8 * W) Vec {
            u64 e[W];
        };
        constexpr u32 kLanes = 2 * B / W;     // lanes per column
        constexpr u32 kStep = 256 / kLanes;   // columns per sweep of the workgroup
        constexpr u32 kUnroll = 16;
        const u32 b = (tid % (B / W)) * W, h = (tid / (B / W)) & 1;
        const Vec *src = reinterpret_cast<const Vec *>(in + i0 + b + (h ? half : 0));
        const u64 colv = N / W;  // a column in Vec units
        u64 *row = lds + (2 * b + h) * pitch;
        u32 j = tid / kLanes;
        for (; j + (kUnroll - 1) * kStep < ncols; j += kUnroll * kStep) {
            Vec x[kUnroll];
#pragma unroll
            for (u32 u = 0; u < kUnroll; u++) x[u] = src[(u64)(j + u * kStep) * colv];
#pragma unroll
            for (u32 u = 0; u < kUnroll; u++)
#pragma unroll
                for (u32 w = 0; w < W; w++) row[2 * w * pitch + j + u * kStep] = x[u].e[w];
        }
        // the rest 8 at a time; loads past the last column re-read it
        for (; j < ncols; j += 8 * kStep) {
            Vec x[8];
#pragma unroll
            for (u32 u = 0; u < 8; u++) x[u] = src[(u64)min(j + u * kStep, ncols - 1) * colv];
#pragma unroll
            for (u32 u = 0; u < 8; u++)
                if (j + u * kStep < ncols) {
#pragma unroll
                    for (u32 w = 0; w < W; w++) row[2 * w * pitch + j + u * kStep] = x[u].e[w];
                }
        }
    }
    __syncthreads();
    // pair k: LDS rows 2k and 2k + 1 (positions i0 + k and i0 + k + N/2)
    // become output rows rev(i0 + k) and rev(i0 + k) + 1, one run of 2 ncols
    // elements; four LDS reads per lane before their stores
    const u32 wave = tid >> 6, lane = tid & 63;
    const u32 run = 2 * ncols;
    for (u32 k = wave; k < B; k += 4) {
        const u64 orow = (u64)zk_bitrev((u32)(i0 + k), lg);
        u64 *dst = out + orow * ncols;
        const u64 *r0 = lds + 2 * k * pitch;  // row 2k + 1 at r0 + pitch
        for (u32 e0 = 0; e0 < run; e0 += 4 * 64) {
            u64 x[4];
#pragma unroll
            for (u32 u = 0; u < 4; u++) {
                const u32 e = min(e0 + 64 * u + lane, run - 1);  // clamped: every lane reads
                x[u] = r0[e + (e < ncols ? 0 : pitch - ncols)];
            }
#pragma unroll
            for (u32 u = 0; u < 4; u++) {
                const u32 e = e0 + 64 * u + lane;
                if (e < run) dst[e] = x[u];
            }
        }
    }
}

// Positions per workgroup of the paired-row transpose: the largest B whose
// LDS (B pairs of rows of ncols elements, odd pitch) stays within 56 KiB
// (three or more workgroups per CU), and whose workgroup count fits N / 2;
// 0 when none does (the 64 x 64 tile kernel runs instead).
static u32 pairs_b(u32 lg, u64 ncols, u32 &pitch)
{
    pitch = (u32)(ncols | 1);
    if (lg < 1 || ncols > 0xFFFFFFFFull) return 0;
    const u64 half = 1ull << (lg - 1);
    for (u32 B : {16u, 8u}) {
        if (half >= B && (u64)2 * B * pitch * sizeof(u64) <= (56u << 10)) return B;
    }
    return 0;
}

void transpose_rev_launch(const u64 *in, u64 *out, u32 lg, u64 ncols, hipStream_t st)
{
    if (ncols == 0) return;
    const u64 N = 1ull << lg;
    ProfScope ps("transpose_rev", st);
    // HIP counts a grid in threads as a 32-bit number: launch in chunks of
    // fewer than 2^32 threads, each passing its first block
    const u64 max_blocks = (((u64)1 << 32) - 1) / 256;
    u32 pitch;
    const u32 B = pairs_b(lg, ncols, pitch);
    if (B) {
        const u64 blocks = (N >> 1) / B;
        const size_t lds = (size_t)2 * B * pitch * sizeof(u64);
        for (u64 b0 = 0; b0 < blocks; b0 += max_blocks) {
            const u32 nb = (u32)std::min<u64>(max_blocks, blocks - b0);
            const bool wide = ((uintptr_t)in & 15) == 0;  // 16-byte loads
            if (B == 16 && wide)
                hipLaunchKernelGGL((transpose_rev_pairs_kernel<16, 2>), dim3(nb), dim3(256), lds, st, in, out, lg,
                                   (u32)ncols, pitch, b0);
            else if (B == 16)
                hipLaunchKernelGGL((transpose_rev_pairs_kernel<16, 1>), dim3(nb), dim3(256), lds, st, in, out, lg,
                                   (u32)ncols, pitch, b0);
            else if (wide)
                hipLaunchKernelGGL((transpose_rev_pairs_kernel<8, 2>), dim3(nb), dim3(256), lds, st, in, out, lg,
                                   (u32)ncols, pitch, b0);
            else
                hipLaunchKernelGGL((transpose_rev_pairs_kernel<8, 1>), dim3(nb), dim3(256), lds, st, in, out, lg,
                                   (u32)ncols, pitch, b0);
            ZK_HIP(hipGetLastError());
        }
        return;
    }
    const u64 tiles_i = (N + TI - 1) / TI, tiles_j = (ncols + TJ - 1) / TJ, tiles = tiles_i * tiles_j;
    for (u64 t0 = 0; t0 < tiles; t0 += max_blocks) {
        const u32 nb = (u32)std::min<u64>(max_blocks, tiles - t0);
        hipLaunchKernelGGL(transpose_rev_kernel, dim3(nb), dim3(256), 0, st, in, out, lg, ncols, tiles_i, t0);
        ZK_HIP(hipGetLastError());
    }
}

// ------------------------------------------------------------------ inputs
ZK_DEV u64 mix64(u64 z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}

__global__ void fill_random_kernel(u64 *out, u64 n, u64 seed)
{
    const u64 stride = (u64)gridDim.x * blockDim.x;
    for (u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        u64 z = mix64(seed + (i + 1) * 0x9e3779b97f4a7c15ull);
        while (z >= ZK_P) z = mix64(z);
        out[i] = z;
    }
}

void fill_random_launch(u64 *out, u64 n, u64 seed, hipStream_t st)
{
    if (!n) return;
    u64 blocks = (n + 255) / 256;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(fill_random_kernel, dim3((u32)blocks), dim3(256), 0, st, out, n, seed);
    ZK_HIP(hipGetLastError());
}

}  // namespace zk
