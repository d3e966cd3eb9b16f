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
#include "ntt.hpp"

namespace zk {

enum : u32 { F_COSET_IN = 1, F_COSET_OUT = 2, F_SCALE_OUT = 4 };

struct PassArgs {
    const u64 *src;
    u64 *dst;
    u64 src_stride, dst_stride;
    u32 ncols;
    u32 L, S, Sprev, Jbits, Gbits;
    u32 lg_in;
    u32 flags;
    u32 pitch;  // LDS row pitch in elements (2^B, +1 when the padding fits)
    u64 scale;
    const u64 *pw;   // (kMaxLg+1) rows of kPowLen powers for this direction
    const u64 *gpw;  // 3 rows of coset powers
};

constexpr int kThreads = 256;

// omega_{2^L}^e for e < 2^L (L <= 32) from three 2^11-entry power rows.
ZK_DEV u64 root_pow(const u64 *pw, u32 L, u64 e)
{
    u64 r = pw[(size_t)L * kPowLen + (e & (kPowLen - 1))];
    if (L > 11) {
        u32 e1 = (u32)(e >> 11) & (kPowLen - 1);
        if (e1) r = gl_mul(r, pw[(size_t)(L - 11) * kPowLen + e1]);
        if (L > 22) {
            u32 e2 = (u32)(e >> 22);
            if (e2) r = gl_mul(r, pw[(size_t)(L - 22) * kPowLen + e2]);
        }
    }
    return r;
}

ZK_DEV u64 coset_pow(const u64 *gpw, u64 n)
{
    u64 r = gpw[n & (kPowLen - 1)];
    u32 n1 = (u32)(n >> 11) & (kPowLen - 1);
    if (n1) r = gl_mul(r, gpw[kPowLen + n1]);
    u32 n2 = (u32)(n >> 22);
    if (n2) r = gl_mul(r, gpw[2 * kPowLen + n2]);
    return r;
}

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

// Register block sizes for a B-bit group: ceil(B/4) blocks of <= 4 stages.
__host__ __device__ constexpr int nblocks_for(int B) { return (B + 3) / 4; }
__host__ __device__ constexpr int block_k(int B, int i)
{
    return B / nblocks_for(B) + (i < B % nblocks_for(B) ? 1 : 0);
}
__host__ __device__ constexpr int block_s0(int B, int i)
{
    int s = 0;
    for (int t = 0; t < i; t++) s += block_k(B, t);
    return s;
}

template <int B>
struct Pass {
    static constexpr u32 NB = 1u << B;

    // tile -> (column, J_hi, first fast index)
    ZK_DEV static void tile_decode(const PassArgs &a, u32 &col, u64 &jhi, u64 &fast0)
    {
        u64 t = blockIdx.x;
        if (a.S > 0) {
            u64 tiles_b = (1ull << a.S) >> a.Gbits;
            fast0 = (t % tiles_b) << a.Gbits;
            t /= tiles_b;
            jhi = t & ((1ull << a.Jbits) - 1);
            col = (u32)(t >> a.Jbits);
        } else if (a.Jbits > 0) {
            u64 tiles_q = (1ull << a.Jbits) >> a.Gbits;
            fast0 = (t % tiles_q) << a.Gbits;
            col = (u32)(t / tiles_q);
            jhi = 0;
        } else {
            fast0 = 0;
            jhi = 0;
            col = (u32)(t << a.Gbits);
        }
    }

    // Loads element j of group g (with every fused input transform applied).
    ZK_DEV static u64 load(const PassArgs &a, u32 col0, u64 jhi, u64 fast0, u32 g, u32 j)
    {
        if (a.S > 0) {
            const u64 b = fast0 + g;
            const u64 J = (u64)j + (jhi << B);
            const u64 pos = b + (J << a.S);
            u64 v = a.src[(u64)col0 * a.src_stride + pos];
            const u64 bp = b >> a.Sprev;
            if (bp) {
                const u32 jb = B + a.Jbits;  // bits of J
                const u64 rj = jb > 32 ? 0 : (u64)zk_bitrev((u32)J, jb);
                const u64 e = (rj * bp) << a.Sprev;  // < 2^L
                if (e) v = gl_mul(v, root_pow(a.pw, a.L, e));
            }
            return v;
        }
        u32 col = col0;
        u64 n;
        if (a.Jbits > 0) {
            n = ((u64)zk_bitrev(j, B) << a.Jbits) + fast0 + g;
        } else {
            col = col0 + g;
            n = zk_bitrev(j, B);
        }
        if (col >= a.ncols || n >= (1ull << a.lg_in)) return 0;
        u64 v = a.src[(u64)col * a.src_stride + n];
        if ((a.flags & F_COSET_IN) && n) v = gl_mul(v, coset_pow(a.gpw, n));
        return v;
    }

    ZK_DEV static void store(const PassArgs &a, u32 col0, u64 jhi, u64 fast0, u32 g, u32 j, u64 v)
    {
        u32 col = col0;
        u64 pos;
        if (a.S > 0) {
            pos = fast0 + g + (((u64)j + (jhi << B)) << a.S);
        } else if (a.Jbits > 0) {
            pos = (u64)j + ((u64)zk_bitrev((u32)(fast0 + g), a.Jbits) << B);
        } else {
            col = col0 + g;
            pos = j;
            if (col >= a.ncols) return;
        }
        const bool last = (a.S + B) == a.L;
        if (last) {
            if (a.flags & F_SCALE_OUT) v = gl_mul(v, a.scale);
            if ((a.flags & F_COSET_OUT) && pos) v = gl_mul(v, coset_pow(a.gpw, pos));
        }
        a.dst[(u64)col * a.dst_stride + pos] = v;
    }
};

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
    P::tile_decode(a, col, jhi, fast0);
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

int ntt_plan(u32 lg, u32 bits[4])
{
    if (lg <= 11) {
        bits[0] = lg;
        return 1;
    }
    if (lg <= 22) {
        bits[0] = (lg + 1) / 2;
        bits[1] = lg - bits[0];
        return 2;
    }
    bits[0] = (lg + 2) / 3;
    bits[1] = (lg - bits[0] + 1) / 2;
    bits[2] = lg - bits[0] - bits[1];
    return 3;
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
    hipLaunchKernelGGL(ntt_pass_kernel<B>, dim3((u32)ntiles), dim3(kThreads), lds, st, a);
    ZK_HIP(hipGetLastError());
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

// Groups per workgroup (log2) and LDS pitch: as many groups as the fast
// dimension and the LDS budget allow (up to 64), rows padded by one element
// when the padding still fits.
static u32 pick_gbits(u32 B, u64 fast_extent, size_t budget, u32 &pitch)
{
    u32 gb = 0;
    while (gb < 6 && ((size_t)(2u << gb) << B) * sizeof(u64) <= budget && (2ull << gb) <= fast_extent) gb++;
    pitch = ((size_t)(1u << gb) * ((1u << B) + 1) * sizeof(u64) <= budget) ? (1u << B) + 1 : (1u << B);
    return gb;
}

void ntt_launch(DeviceCtx &c, const NttJob &job, hipStream_t st)
{
    if (job.lg == 0 || job.lg > (u32)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
    if (job.ncols == 0) return;
    u32 bits[4];
    const int np = ntt_plan(job.lg, bits);
    const bool inplace = job.src == job.dst;
    u64 *scratch = nullptr;
    if (inplace && np > 1)
        scratch = (u64 *)c.get_scratch((size_t)job.ncols * ((size_t)1 << job.lg) * sizeof(u64));
    const u64 N = 1ull << job.lg;
    u32 S = 0, Sprev = 0;
    for (int p = 0; p < np; p++) {
        const u32 B = bits[p];
        PassArgs a{};
        a.ncols = job.ncols;
        a.L = job.lg;
        a.S = S;
        a.Sprev = Sprev;
        a.Jbits = job.lg - S - B;
        a.lg_in = job.lg_in;
        a.pw = c.pw[job.inverse ? 1 : 0];
        a.gpw = c.gpw;
        a.flags = 0;
        if (p == 0 && job.coset_in) a.flags |= F_COSET_IN;
        if (p == np - 1) {
            if (job.coset_out) a.flags |= F_COSET_OUT;
            if (job.scale_out) {
                a.flags |= F_SCALE_OUT;
                a.scale = gl_inv(N % ZK_P);
            }
        }
        // buffers: pass 0 reads the caller's input; the last pass writes dst
        const u64 *src;
        u64 src_stride;
        if (p == 0) {
            src = job.src;
            src_stride = job.src_stride;
        } else {
            src = scratch ? scratch : job.dst;
            src_stride = scratch ? N : job.dst_stride;
        }
        u64 *dst;
        u64 dst_stride;
        if (p == np - 1) {
            dst = job.dst;
            dst_stride = job.dst_stride;
        } else {
            dst = scratch ? scratch : job.dst;
            dst_stride = scratch ? N : job.dst_stride;
        }
        a.src = src;
        a.src_stride = src_stride;
        a.dst = dst;
        a.dst_stride = dst_stride;
        u64 fast_extent, ntiles;
        if (S > 0) {
            fast_extent = 1ull << S;
            a.Gbits = pick_gbits(B, fast_extent, c.lds_budget, a.pitch);
            ntiles = (u64)job.ncols * (1ull << a.Jbits) * (fast_extent >> a.Gbits);
        } else if (a.Jbits > 0) {
            fast_extent = 1ull << a.Jbits;
            a.Gbits = pick_gbits(B, fast_extent, c.lds_budget, a.pitch);
            ntiles = (u64)job.ncols * (fast_extent >> a.Gbits);
        } else {
            fast_extent = job.ncols;
            a.Gbits = pick_gbits(B, fast_extent, c.lds_budget, a.pitch);
            ntiles = (job.ncols + (1u << a.Gbits) - 1) >> a.Gbits;
        }
        {
            ProfScope ps(pass_name(job.prof_tag, p), st);
            dispatch_pass(B, a, ntiles, st);
        }
        Sprev = S;
        S += B;
    }
}

// ---------------------------------------------------------------- transpose
// 64 (rows i) x 32 (columns j) tiles through LDS; loads read 512 B column runs,
// stores write 256 B runs of output row rev(i)  (reference: kernels.cu:94-135).
constexpr int TI = 64, TJ = 32;

__global__ __launch_bounds__(256) void transpose_rev_kernel(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                            u32 lg, u64 ncols, u64 tiles_i)
{
    __shared__ u64 tile[TJ][TI + 1];
    const u64 N = 1ull << lg;
    const u64 ti = blockIdx.x % tiles_i, tj = blockIdx.x / tiles_i;
    const u64 i0 = ti * TI, j0 = tj * TJ;
    const u32 tx = threadIdx.x & 63, ty = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < TJ / 4; k++) {
        const u64 j = j0 + ty + 4 * k, i = i0 + tx;
        if (j < ncols && i < N) tile[ty + 4 * k][tx] = in[j * N + i];
    }
    __syncthreads();
    const u32 sx = threadIdx.x & 31, sy = threadIdx.x >> 5;
#pragma unroll
    for (int k = 0; k < TI / 8; k++) {
        const u64 i = i0 + sy + 8 * k, j = j0 + sx;
        if (j < ncols && i < N) out[(u64)zk_bitrev((u32)i, lg) * ncols + j] = tile[sx][sy + 8 * k];
    }
}

void transpose_rev_launch(const u64 *in, u64 *out, u32 lg, u64 ncols, hipStream_t st)
{
    if (ncols == 0) return;
    const u64 N = 1ull << lg;
    const u64 tiles_i = (N + TI - 1) / TI, tiles_j = (ncols + TJ - 1) / TJ;
    ProfScope ps("transpose_rev", st);
    hipLaunchKernelGGL(transpose_rev_kernel, dim3((u32)(tiles_i * tiles_j)), dim3(256), 0, st, in, out, lg,
                       ncols, tiles_i);
    ZK_HIP(hipGetLastError());
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
