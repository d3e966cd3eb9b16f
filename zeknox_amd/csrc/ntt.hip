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
#include <algorithm>
#include <cstdlib>
#include <vector>

#include "ntt.hpp"

namespace zk {

// F_BLOCK_OUT / F_BLOCK_IN: the intermediate of a two-pass 2^20 transform is
// kept blocked (see ntt_launch) instead of in its natural place.
// F_FORCE_EXACT: every tile takes the canonical recompute path, as if its
// relaxed pass had reported a rare case (set when the environment variable
// ZK_NTT_FORCE_EXACT is 1: tests use it to check that path against the
// oracle; random inputs reach it with probability about 2^-32 per operation).
enum : u32 { F_COSET_IN = 1, F_COSET_OUT = 2, F_SCALE_OUT = 4, F_BLOCK_OUT = 8, F_BLOCK_IN = 16, F_FORCE_EXACT = 32 };

struct PassArgs {
    const u64 *src;
    u64 *dst;
    u64 src_stride, dst_stride;
    u32 ncols;
    u32 L, S, Sprev, Jbits, Gbits;
    u32 lg_in;
    u32 flags;
    u32 skip;   // first pass: leading DIT stages that only replicate (LDE zeros)
    u32 pitch;  // LDS row pitch in elements (2^B, +1 when the padding fits)
    u64 scale;
    const u64 *pw;   // (kMaxLg+1) rows of kPowLen powers for this direction
    const u64 *gpw;  // 3 rows of coset powers
    // passes after the first: the inter-pass twiddles as a table
    // tw[J * 2^(S - Sprev) + bp] (see twiddle_table), or null to compute them
    const u64 *tw;
};

struct PassPlan {
    PassArgs a;
    u32 B;
    u64 ntiles;
    bool fast;
};

// Kernel arguments of every pass of a transform (no launch): pw is the
// direction's power table, scratch the ping-pong buffer of in-place plans.
std::vector<PassPlan> plan_passes(const NttJob &job, const u64 *pw, const u64 *gpw, u64 *scratch, size_t lds_budget);

constexpr int kThreads = 256;

// omega_{2^L}^e for e < 2^L (L <= 32) from three 2^11-entry power rows.
ZK_HD u64 root_pow(const u64 *pw, u32 L, u64 e)
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

ZK_HD u64 coset_pow(const u64 *gpw, u64 n)
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

// Workgroups are dispatched to the 8 XCDs round-robin (block b on XCD b % 8),
// each XCD with its own L2.  Adjacent tiles share 128-byte lines (a tile's
// rows are runs of G * 8 = 64 bytes), so block b gets the tile that makes
// each XCD walk one contiguous tile range: x = b % 8 takes tiles
// [start_x, start_x + count_x) in order.  A bijection on [0, ntiles).
constexpr u32 kXcds = 8;
ZK_HD u64 xcd_tile(u64 b, u64 ntiles)
{
    const u64 q = ntiles / kXcds, r = ntiles % kXcds, x = b % kXcds, i = b / kXcds;
    return x * q + (x < r ? x : r) + i;
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

    // tile -> (column, J_hi, first fast index).  Passes after the first with
    // at most 8 groups per tile take superblocks of up to 8 adjacent fast
    // blocks across all columns: the tiles that run together on one XCD
    // (about 64) cover 8 columns x 8 adjacent fast blocks, so their rows are
    // read and written in 512-byte runs and they share one slice of the
    // twiddle table in L2.  Wider tiles (whose rows are already 512-byte
    // runs) take the fast index fastest.
    ZK_HD static void tile_decode(const PassArgs &a, u64 t, u32 &col, u64 &jhi, u64 &fast0)
    {
        if (a.S > 0 && a.Gbits > 3) {
            const u64 tiles_b = (1ull << a.S) >> a.Gbits;
            fast0 = (t % tiles_b) << a.Gbits;
            t /= tiles_b;
            jhi = t & ((1ull << a.Jbits) - 1);
            col = (u32)(t >> a.Jbits);
        } else if (a.S > 0) {
            const u32 tiles_b = (u32)((1ull << a.S) >> a.Gbits);
            const u32 sb = tiles_b < 8 ? tiles_b : 8;  // divides tiles_b (powers of two)
            u32 r = (u32)t;
            const u32 bb = r % sb;
            r /= sb;
            col = r % a.ncols;
            r /= a.ncols;
            const u32 nsb = tiles_b / sb;
            fast0 = (u64)((r % nsb) * sb + bb) << a.Gbits;
            jhi = r / nsb;
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
    P::tile_decode(a, xcd_tile(blockIdx.x, gridDim.x), col, jhi, fast0);
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

// ---------------------------------------------------------------------------
// Fast pass (B <= 10): a group of 2^B = 2^BA * 2^BB elements is transformed as
// 2^BB register-resident DITs of 2^BA points (step A), a twiddle
// omega_{2^B}^(rev_BB(q) * k), an LDS transpose, and 2^BA register-resident
// DITs of 2^BB points (step B).  BA, BB <= 5, so every twiddle inside the
// register DITs is a power of omega_32, i.e. a power of two in Goldilocks
// (omega_64 = 2^39, 2^96 = -1): those multiplications are shifts plus one
// reduction.  Arithmetic is relaxed (any u64) until the last pass's store.

// omega_{2^k} = 2^E for k <= 6: E = 39 * 2^(6-k) mod 192 (inverse: -E).
__host__ __device__ constexpr int root_exp2(int k, bool inv)
{
    return inv ? (192 - ((39 << (6 - k)) % 192)) % 192 : (39 << (6 - k)) % 192;
}

// Arithmetic of one block, relaxed (SAFE = false: any u64, rare cases ORed
// into `rare`) or canonical (SAFE = true: inputs < p, exact, never rare).
template <bool SAFE>
ZK_HD u64 f_add(u64 a, u64 b, bool &rare) { return SAFE ? gl_add(a, b) : gl_add_nc(a, b, rare); }
template <bool SAFE>
ZK_HD u64 f_sub(u64 a, u64 b, bool &rare) { return SAFE ? gl_sub(a, b) : gl_sub_nc(a, b, rare); }
template <bool SAFE>
ZK_HD u64 f_mul(u64 a, u64 b, bool &rare) { return SAFE ? gl_mul(a, b) : gl_mul_nc(a, b, rare); }
template <bool SAFE>
ZK_HD u64 f_mul_pow2(u64 x, int e, bool &rare)
{
    if (SAFE) return e ? gl_mul(x, pow2_mod_p(e)) : x;
    return gl_mul_pow2_nc(x, e, rare);
}

// K radix-2 DIT stages on 2^K registers (bit-reversed in, natural out) with
// power-of-two twiddles; the first `skip` stages only replicate (their odd
// inputs are known zeros).
template <int K, bool INV, bool SAFE>
ZK_HD void dit_pow2(u64 v[1 << K], int skip, bool &rare)
{
    static_for<0, K>([&](auto sc) {
        constexpr int s = decltype(sc)::value;
        if (s < skip) {
            static_for<0, (1 << K)>([&](auto mc) {
                constexpr int m = decltype(mc)::value;
                if constexpr (!(m & (1 << s))) v[m + (1 << s)] = v[m];
            });
            return;
        }
        static_for<0, (1 << K)>([&](auto mc) {
            constexpr int m = decltype(mc)::value;
            if constexpr (!(m & (1 << s))) {
                constexpr int j = m & ((1 << s) - 1);
                constexpr int E = (j * root_exp2(s + 1, INV)) % 192;
                const u64 t = f_mul_pow2<SAFE>(v[m + (1 << s)], E % 96, rare);
                const u64 x = v[m];
                const u64 a = f_add<SAFE>(x, t, rare), b = f_sub<SAFE>(x, t, rare);
                if constexpr (E >= 96) {
                    v[m] = b;
                    v[m + (1 << s)] = a;
                } else {
                    v[m] = a;
                    v[m + (1 << s)] = b;
                }
            }
        });
    });
}

template <int BA, int BB, bool INV>
struct FastPass {
    static constexpr int B = BA + BB, NA = 1 << BA, NB = 1 << BB;
    static constexpr u32 ROW = NA + 1;      // LDS row (one step-A block) pitch
    static constexpr u32 GP = NB * ROW + 2;  // LDS pitch of one group

    // step A of block (g, q): load (+ inter-pass twiddle), 2^BA-point DIT,
    // middle twiddle omega_{2^B}^(rev_BB(q) k)
    template <bool SAFE>
    ZK_HD static void step_a(const PassArgs &a, u32 col, u64 jhi, u64 fast0, u32 g, u32 q, u64 v[NA], bool &rare)
    {
        if (a.S > 0) {
            const u64 b = fast0 + g;
            const u64 *src = a.src + (u64)col * a.src_stride + b;
#pragma unroll
            for (int t = 0; t < NA; t++) {
                const u64 x = src[((u64)(q * NA + t) + (jhi << B)) << a.S];
                v[t] = SAFE ? gl_canon(x) : x;
            }
            const u64 bp = b >> a.Sprev;
            if (a.tw) {
                // all twiddle loads are issued together with the data loads
                // (the scheduler would otherwise interleave each load with
                // its multiplication and wait on it)
                const u64 *tw = a.tw + bp;
                const u32 sh = a.S - a.Sprev;
                u64 w[NA];
#pragma unroll
                for (int t = 0; t < NA; t++) w[t] = tw[((u64)(q * NA + t) + (jhi << B)) << sh];
#ifdef __HIP_DEVICE_COMPILE__
                __builtin_amdgcn_sched_barrier(0);
#endif
#pragma unroll
                for (int t = 0; t < NA; t++) v[t] = f_mul<SAFE>(v[t], w[t], rare);
            } else if (bp) {
                // omega_N^(bp 2^Sprev rev(J)), rev(J) = rev(t) 2^(BB+Jbits) + rev(q) 2^Jbits + rev(jhi)
                const u64 mask = (1ull << a.L) - 1;
                const u64 ub = (bp << a.Sprev) & mask;
                const u64 rq = ((u64)zk_bitrev(q, BB) << a.Jbits) + zk_bitrev((u32)jhi, a.Jbits);
                const u64 base = root_pow(a.pw, a.L, (ub * rq) & mask);
                const u64 step = root_pow(a.pw, a.L, (ub << (BB + a.Jbits)) & mask);
                const u64 step2 = f_mul<SAFE>(step, step, rare);
                u64 m0 = base, m1 = f_mul<SAFE>(base, step, rare);
                static_for<0, NA / 2>([&](auto ic) {
                    constexpr int m = 2 * decltype(ic)::value;
                    constexpr int r0 = ct_bitrev(m, BA), r1 = ct_bitrev(m + 1, BA);
                    v[r0] = f_mul<SAFE>(v[r0], m0, rare);
                    v[r1] = f_mul<SAFE>(v[r1], m1, rare);
                    if constexpr (m + 2 < NA) {
                        m0 = f_mul<SAFE>(m0, step2, rare);
                        m1 = f_mul<SAFE>(m1, step2, rare);
                    }
                });
            }
        } else {
            // first pass: group element J = q NA + t is natural index
            // n = rev_B(J) 2^Jbits + (group's fast index); single pass: n = rev_B(J)
            u32 c = col;
            u64 nlow;
            if (a.Jbits > 0) {
                nlow = fast0 + g;
            } else {
                c = col + g;
                nlow = 0;
            }
            const u64 *src = a.src + (u64)c * a.src_stride;
            const u64 lim = 1ull << a.lg_in;
            const u32 skipmask = (1u << a.skip) - 1;
            const bool colok = c < a.ncols;
#pragma unroll
            for (int t = 0; t < NA; t++) {
                const u64 n = ((u64)zk_bitrev((u32)(q * NA + t), B) << a.Jbits) + nlow;
                const bool ok = colok && !(t & skipmask) && n < lim;
                u64 x = 0;
                if (ok) x = src[n];
                v[t] = SAFE ? gl_canon(x) : x;
            }
            if (a.flags & F_COSET_IN) {
                // g^n from the three 2^11-entry rows (entry 0 of each row is 1,
                // so no digit needs a branch); n < 2^32
#pragma unroll
                for (int t = 0; t < NA; t++) {
                    if (t & skipmask) continue;
                    const u64 n = ((u64)zk_bitrev((u32)(q * NA + t), B) << a.Jbits) + nlow;
                    const u64 c0 = a.gpw[n & (kPowLen - 1)];
                    const u64 c1 = a.gpw[kPowLen + ((n >> kPowBits) & (kPowLen - 1))];
                    const u64 c2 = a.gpw[2 * kPowLen + (n >> (2 * kPowBits))];
                    v[t] = f_mul<SAFE>(v[t], f_mul<SAFE>(f_mul<SAFE>(c0, c1, rare), c2, rare), rare);
                }
            }
        }
        dit_pow2<BA, INV, SAFE>(v, a.S > 0 ? 0 : (int)a.skip, rare);
        const u32 rq = zk_bitrev(q, BB);
        if (rq) {
            const u64 *wB = a.pw + (size_t)B * kPowLen;  // powers of omega_{2^B}
#pragma unroll
            for (int k = 1; k < NA; k++) v[k] = f_mul<SAFE>(v[k], wB[rq * k], rare);
        }
    }

    // step B of column ka of group g: 2^BB-point DIT over the step-A blocks,
    // plus the inverse scaling on the last pass
    template <bool SAFE>
    ZK_HD static void step_b(const PassArgs &a, const u64 *lds, u32 g, u32 ka, bool last, u64 v[NB], bool &rare)
    {
        const u64 *colp = lds + g * GP + ka;
#pragma unroll
        for (int qq = 0; qq < NB; qq++) v[qq] = SAFE ? gl_canon(colp[qq * ROW]) : colp[qq * ROW];
        dit_pow2<BB, INV, SAFE>(v, 0, rare);
        if (last && (a.flags & F_SCALE_OUT)) {
#pragma unroll
            for (int kb = 0; kb < NB; kb++) v[kb] = f_mul<SAFE>(v[kb], a.scale, rare);
        }
    }
};

// One phase of one workgroup tile: phase 0 = step A of every block of the
// tile into LDS, phase 1 = step B and the stores; threads tid, tid + nthr, ...
// (the kernel runs both phases with a barrier in between).
template <int BA, int BB, bool INV>
ZK_HD void fast_phase(const PassArgs &a, u64 tile, int phase, u32 tid, u32 nthr, u64 *lds)
{
    using F = FastPass<BA, BB, INV>;
    constexpr int B = F::B, NA = F::NA, NB = F::NB;
    const u32 G = 1u << a.Gbits;
    u32 col;
    u64 jhi, fast0;
    Pass<B>::tile_decode(a, tile, col, jhi, fast0);
    const bool last = (a.S + B) == a.L;
    if (phase == 0) {
        for (u32 set = tid; set < (G << BB); set += nthr) {
            const u32 g = set & (G - 1), q = set >> a.Gbits;
            u64 v[NA];
            bool rare = (a.flags & F_FORCE_EXACT) != 0;
            F::template step_a<false>(a, col, jhi, fast0, g, q, v, rare);
            if (rare) F::template step_a<true>(a, col, jhi, fast0, g, q, v, rare);
            u64 *row = lds + g * F::GP + q * F::ROW;
#pragma unroll
            for (int k = 0; k < NA; k++) row[k] = v[k];
        }
        return;
    }
    for (u32 set = tid; set < (G << BA); set += nthr) {
        u32 g, ka;
        if (a.S > 0) {  // consecutive lanes: consecutive b (coalesced rows)
            g = set & (G - 1);
            ka = set >> a.Gbits;
        } else {        // consecutive lanes: consecutive output index
            ka = set & (NA - 1);
            g = set >> BA;
        }
        u64 v[NB];
        bool rare = (a.flags & F_FORCE_EXACT) != 0;
        F::template step_b<false>(a, lds, g, ka, last, v, rare);
        if (rare) F::template step_b<true>(a, lds, g, ka, last, v, rare);
        u32 c = col;
        u64 base, stride;
        if (a.S > 0) {
            base = fast0 + g + ((u64)ka << a.S) + ((jhi << B) << a.S);
            stride = (u64)NA << a.S;
        } else if (a.Jbits > 0) {
            base = ka + ((u64)zk_bitrev((u32)(fast0 + g), a.Jbits) << B);
            stride = NA;
        } else {
            c = col + g;
            base = ka;
            stride = NA;
        }
        if (c >= a.ncols) continue;
        u64 *dst = a.dst + (u64)c * a.dst_stride;
#pragma unroll
        for (int kb = 0; kb < NB; kb++) {
            const u64 pos = base + (u64)kb * stride;
            u64 x = v[kb];
            if (last) {
                x = gl_canon(x);
                if ((a.flags & F_COSET_OUT) && pos) x = gl_mul(x, coset_pow(a.gpw, pos));
            }
            dst[pos] = x;
        }
    }
}

template <int BA, int BB, bool INV>
__global__ __launch_bounds__(kThreads) void ntt_fast_kernel(PassArgs a)
{
    extern __shared__ u64 lds[];
    const u64 tile = xcd_tile(blockIdx.x, gridDim.x);
    fast_phase<BA, BB, INV>(a, tile, 0, threadIdx.x, kThreads, lds);
    __syncthreads();
    fast_phase<BA, BB, INV>(a, tile, 1, threadIdx.x, kThreads, lds);
}

// ---------------------------------------------------------------------------
// 10-bit pass with 16 elements per thread (the 2^10-point passes of LDEs and
// NTTs of 2^19 .. 2^20 points and of 2^28 .. 2^30).  A group of 1024 = 16 x 16
// x 4 elements is transformed in three register blocks: element
// J = t + 16 m + 256 h of the group (bit-reversed input order) gives
//   X[kA + 16 kB + 256 kC] = sum_h w4^(rev2(h) kC) w1024^(rev2(h)(kA + 16 kB))
//       sum_m w16^(rev4(m) kB) w256^(rev4(m) kA) sum_t w16^(rev4(t) kA) x[J]
// (step A over t, twiddle w256, LDS, step B over m, twiddle w1024, LDS, step C
// over h).  A workgroup of 64 NG threads holds NG groups (one wave per group
// in step B); 16 elements per thread keep the registers under 128, so four
// waves per SIMD fit, twice the occupancy of the 32-element pass: a single
// wave cannot issue its dependent 64-bit integer instructions back to back.
// Pass layout (tiles, loads, stores, fused coset / twiddle / scaling) is that
// of FastPass.
constexpr u32 kF10Row = 17;                 // one (g, h, kA) row of 16 m, padded
constexpr u32 kF10H = 16 * kF10Row;         // one (g, h) block
// One group; pitch = 4 mod 32 elements.  Pitches 8, 4, 2 and 1 mod 32 measured
// (LDS bank-conflict cycles of the second LDE pass 66 M, 27 M, 13 M, 13 M):
// 4 gives the fastest passes, 8 the slowest (pass 2 0.739 against 0.767 ms).
constexpr u32 kF10G = 4 * kF10H + 4;
template <int NG>
constexpr size_t f10_lds_bytes() { return (size_t)NG * kF10G * sizeof(u64); }
template <int NG>
constexpr int f10_lg() { return NG == 16 ? 4 : 3; }

// Step A of thread tid: load 16 elements of one group, inter-pass twiddle or
// coset factor, 16-point DIT, w256 twiddle, into LDS.
template <int NG, bool INV, bool SAFE>
ZK_HD void f10_step_a(const PassArgs &a, u32 col, u64 jhi, u64 fast0, u32 tid, u64 *lds, bool &rare)
{
    constexpr int LG = f10_lg<NG>();
    u32 g = tid & (NG - 1), m = (tid >> LG) & 15, h = tid >> (LG + 4);
    if (a.flags & F_BLOCK_IN) {
        // h's high bit on lane bit LG: J and J + 512 (slots 2i, 2i+1 of the
        // blocked layout, one 128-byte line) are read by the same instruction
        h = ((tid >> LG) & 1) * 2 + (tid >> (LG + 5));
        m = (tid >> (LG + 1)) & 15;
    }
    const u32 J0 = 16 * m + 256 * h;
    u64 v[16];
    if (a.S > 0) {
        const u64 b = fast0 + g;
        const u64 Jhi = jhi << 10;
        // per-thread base pointers, per-element offsets uniform across the
        // workgroup (J0 has its low 4 bits clear: rev10(J0 + t) = rev10(J0) +
        // 64 rev4(t))
        if (a.flags & F_BLOCK_IN) {
            // blocked intermediate (Jbits = 0): element (b, J) at
            // (b >> 3) 2^(B+3) + rev_B(J) 8 + (b & 7)
            const u64 *src = a.src + (u64)col * a.src_stride + ((b >> 3) << 13) + (b & 7) +
                             ((u64)zk_bitrev(J0, 10) << 3);
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const u64 x = src[ct_bitrev(t, 4) << 9];
                v[t] = SAFE ? gl_canon(x) : x;
            }
        } else {
            const u64 *src = a.src + (u64)col * a.src_stride + b + (((u64)J0 + Jhi) << a.S);
#pragma unroll
            for (int t = 0; t < 16; t++) {
                const u64 x = src[(u64)t << a.S];
                v[t] = SAFE ? gl_canon(x) : x;
            }
        }
        const u64 bp = b >> a.Sprev;
        if (a.tw) {
            const u32 sh = a.S - a.Sprev;
            const u64 *tw = a.tw + bp + (((u64)J0 + Jhi) << sh);
            u64 w[16];
#pragma unroll
            for (int t = 0; t < 16; t++) w[t] = tw[(u64)t << sh];
#ifdef __HIP_DEVICE_COMPILE__
            __builtin_amdgcn_sched_barrier(0);
#endif
#pragma unroll
            for (int t = 0; t < 16; t++) v[t] = f_mul<SAFE>(v[t], w[t], rare);
        } else if (bp) {
            // omega_N^(bp 2^Sprev rev(J)), rev(J) = rev4(t) 2^(jb-4) + rev_(jb-4)(J >> 4)
            const u32 jb = a.L - a.S;
            const u64 mask = (1ull << a.L) - 1;
            const u64 ub = (bp << a.Sprev) & mask;
            const u64 rhi = zk_bitrev((u32)((J0 >> 4) + (jhi << 6)), jb - 4);
            const u64 base = root_pow(a.pw, a.L, (ub * rhi) & mask);
            const u64 step = root_pow(a.pw, a.L, (ub << (jb - 4)) & mask);
            const u64 step2 = f_mul<SAFE>(step, step, rare);
            u64 m0 = base, m1 = f_mul<SAFE>(base, step, rare);
            static_for<0, 8>([&](auto ic) {
                constexpr int k = 2 * decltype(ic)::value;
                constexpr int r0 = ct_bitrev(k, 4), r1 = ct_bitrev(k + 1, 4);
                v[r0] = f_mul<SAFE>(v[r0], m0, rare);
                v[r1] = f_mul<SAFE>(v[r1], m1, rare);
                if constexpr (k + 2 < 16) {
                    m0 = f_mul<SAFE>(m0, step2, rare);
                    m1 = f_mul<SAFE>(m1, step2, rare);
                }
            });
        }
    } else {
        // first pass of a multi-pass transform: element J is natural index
        // n = rev10(J) 2^Jbits + (the group's fast index)
        const u64 nlow = fast0 + g;
        const u64 n0 = ((u64)zk_bitrev(J0, 10) << a.Jbits) + nlow;  // + (64 rev4(t) << Jbits)
        const u64 *src = a.src + (u64)col * a.src_stride + n0;
        const u64 lim = 1ull << a.lg_in;
        const u32 skipmask = (1u << a.skip) - 1;
        const bool colok = col < a.ncols;
#pragma unroll
        for (int t = 0; t < 16; t++) {
            const u64 dn = (u64)(64 * ct_bitrev(t, 4)) << a.Jbits;
            const bool ok = colok && !(t & skipmask) && n0 + dn < lim;
            u64 x = 0;
            if (ok) x = src[dn];
            v[t] = SAFE ? gl_canon(x) : x;
        }
        if (a.flags & F_COSET_IN) {
#pragma unroll
            for (int t = 0; t < 16; t++) {
                if (t & skipmask) continue;
                const u64 n = ((u64)zk_bitrev(J0 + t, 10) << a.Jbits) + nlow;
                const u64 c0 = a.gpw[n & (kPowLen - 1)];
                const u64 c1 = a.gpw[kPowLen + ((n >> kPowBits) & (kPowLen - 1))];
                const u64 c2 = a.gpw[2 * kPowLen + (n >> (2 * kPowBits))];
                v[t] = f_mul<SAFE>(v[t], f_mul<SAFE>(f_mul<SAFE>(c0, c1, rare), c2, rare), rare);
            }
        }
    }
    dit_pow2<4, INV, SAFE>(v, a.S > 0 ? 0 : (int)a.skip, rare);
    const u32 rm = zk_bitrev(m, 4);
    if (rm) {
        const u64 *w8 = a.pw + (size_t)8 * kPowLen;  // powers of omega_256
#pragma unroll
        for (int k = 1; k < 16; k++) v[k] = f_mul<SAFE>(v[k], w8[rm * k], rare);
    }
    u64 *row = lds + g * kF10G + h * kF10H + m;
#pragma unroll
    for (int k = 0; k < 16; k++) row[k * kF10Row] = v[k];
}

// Step B of thread tid (wave = group): 16-point DIT over m, w1024 twiddle, in
// place (each thread rewrites exactly the 16 entries it read).
template <int NG, bool INV, bool SAFE>
ZK_HD void f10_step_b(const PassArgs &a, u32 tid, u64 *lds, bool &rare)
{
    const u32 g = tid >> 6, kA = tid & 15, h = (tid >> 4) & 3;
    u64 *row = lds + g * kF10G + h * kF10H + kA * kF10Row;
    u64 u[16];
#pragma unroll
    for (int m = 0; m < 16; m++) u[m] = SAFE ? gl_canon(row[m]) : row[m];
    dit_pow2<4, INV, SAFE>(u, 0, rare);
    const u32 rh = zk_bitrev(h, 2);
    if (rh) {
        const u64 *w10 = a.pw + (size_t)10 * kPowLen;  // powers of omega_1024
#pragma unroll
        for (int kB = 0; kB < 16; kB++) u[kB] = f_mul<SAFE>(u[kB], w10[rh * (kA + 16 * kB)], rare);
    }
#pragma unroll
    for (int kB = 0; kB < 16; kB++) row[kB] = u[kB];
}

// The (g, kA, kBq) of a step-C thread: passes after the first put consecutive
// groups (fast indices) on consecutive lanes, the first pass consecutive
// output indices.
template <int NG>
ZK_HD void f10_c_roles(const PassArgs &a, u32 tid, u32 &g, u32 &kA, u32 &kBq)
{
    constexpr int LG = f10_lg<NG>();
    if (a.S > 0) {
        g = tid & (NG - 1);
        kA = (tid >> LG) & 15;
        kBq = tid >> (LG + 4);
    } else {
        g = tid >> 6;
        kA = tid & 15;
        kBq = (tid >> 4) & 3;
    }
}

// Step C: four 4-point DITs over h (kB = kBq + 4 i); out[4 i + kC] is output
// index kA + 16 kB + 256 kC.
template <int NG, bool INV, bool SAFE>
ZK_HD void f10_step_c(const PassArgs &a, u32 tid, const u64 *lds, bool last, u64 out[16], bool &rare)
{
    u32 g, kA, kBq;
    f10_c_roles<NG>(a, tid, g, kA, kBq);
    const u64 *base = lds + g * kF10G + kA * kF10Row + kBq;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u64 w[4];
#pragma unroll
        for (int hh = 0; hh < 4; hh++) {
            const u64 x = base[hh * kF10H + 4 * i];
            w[hh] = SAFE ? gl_canon(x) : x;
        }
        dit_pow2<2, INV, SAFE>(w, 0, rare);
        if (last && (a.flags & F_SCALE_OUT)) {
#pragma unroll
            for (int kC = 0; kC < 4; kC++) w[kC] = f_mul<SAFE>(w[kC], a.scale, rare);
        }
#pragma unroll
        for (int kC = 0; kC < 4; kC++) out[4 * i + kC] = w[kC];
    }
}

template <int NG>
ZK_HD void f10_store(const PassArgs &a, u32 col, u64 jhi, u64 fast0, u32 tid, bool last, const u64 out[16])
{
    u32 g, kA, kBq;
    f10_c_roles<NG>(a, tid, g, kA, kBq);
    if (col >= a.ncols) return;
    // output index k = k0 + 64 i + 256 kC; position = pos0 + (uniform offset)
    const u64 k0 = kA + 16 * kBq;
    u64 pos0;
    u32 sh;
    if (a.S > 0) {
        pos0 = fast0 + g + ((k0 + (jhi << 10)) << a.S);
        sh = a.S;
    } else if (a.flags & F_BLOCK_OUT) {  // blocked for the next pass (its b = k, rev(J) = f)
        pos0 = ((k0 >> 3) << (a.Jbits + 3)) + ((fast0 + g) << 3) + (k0 & 7);
        sh = a.Jbits;  // (64 i + 256 kC) >> 3 blocks of 2^(Jbits+3)
    } else {
        pos0 = k0 + ((u64)zk_bitrev((u32)(fast0 + g), a.Jbits) << 10);
        sh = 0;
    }
    u64 *dst = a.dst + (u64)col * a.dst_stride + pos0;
#pragma unroll
    for (int i = 0; i < 4; i++) {
#pragma unroll
        for (int kC = 0; kC < 4; kC++) {
            const u64 d = (u64)(64 * i + 256 * kC) << sh;
            u64 x = out[4 * i + kC];
            if (last) {
                x = gl_canon(x);
                const u64 pos = pos0 + d;
                if ((a.flags & F_COSET_OUT) && pos) x = gl_mul(x, coset_pow(a.gpw, pos));
            }
            dst[d] = x;
        }
    }
}

// One tile: relaxed steps with a workgroup-wide rare-case check after each;
// a flagged tile is recomputed on the canonical path before anything is
// stored (the tile's inputs are still intact, also in place).
template <int NG, bool INV>
__global__ __launch_bounds__(64 * NG, 4) void ntt_fast10_kernel(PassArgs a)
{
    extern __shared__ u64 lds[];
    const u64 tile = xcd_tile(blockIdx.x, gridDim.x);
    u32 col;
    u64 jhi, fast0;
    Pass<10>::tile_decode(a, tile, col, jhi, fast0);
    const bool last = (a.S + 10) == a.L;
    const u32 tid = threadIdx.x;
    u64 out[16];
    bool rare = (a.flags & F_FORCE_EXACT) != 0;
    f10_step_a<NG, INV, false>(a, col, jhi, fast0, tid, lds, rare);
    bool redo = __syncthreads_or(rare);
    if (!redo) {
        f10_step_b<NG, INV, false>(a, tid, lds, rare);
        redo = __syncthreads_or(rare);
    }
    if (!redo) {
        f10_step_c<NG, INV, false>(a, tid, lds, last, out, rare);
        redo = __syncthreads_or(rare);
    }
    if (redo) {
        f10_step_a<NG, INV, true>(a, col, jhi, fast0, tid, lds, rare);
        __syncthreads();
        f10_step_b<NG, INV, true>(a, tid, lds, rare);
        __syncthreads();
        f10_step_c<NG, INV, true>(a, tid, lds, last, out, rare);
    }
    f10_store<NG>(a, col, jhi, fast0, tid, last, out);
}

template <int NG, bool INV>
static void launch_fast10(const PassArgs &a, u64 ntiles, hipStream_t st)
{
    ZK_HIP(hipFuncSetAttribute((const void *)ntt_fast10_kernel<NG, INV>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)f10_lds_bytes<NG>()));
    hipLaunchKernelGGL((ntt_fast10_kernel<NG, INV>), dim3((u32)ntiles), dim3(64 * NG), f10_lds_bytes<NG>(), st, a);
    ZK_HIP(hipGetLastError());
}

// Groups per workgroup of the 16-element pass (log2): 8 groups, two
// workgroups per CU.  (16 groups in a 1024-thread workgroup make every row
// access a whole line but leave one workgroup per CU, whose barriers then
// stall the CU: measured slower for both passes.)
static u32 f10_gbits(const PassArgs &) { return 3; }

// The 16-element pass applies to 10-bit passes whose workgroup holds the
// groups of one column: passes after the first, and first passes of
// multi-pass transforms (a single 10-bit pass keeps the 32-element kernel),
// when the fast dimension has at least that many groups.
static bool use_fast10(u32 B, const PassArgs &a)
{
    const u64 extent = a.S > 0 ? (1ull << a.S) : (a.Jbits > 0 ? (1ull << a.Jbits) : 0);
    return B == 10 && extent >= (1ull << f10_gbits(a));
}

// Re-tiles a pass for the 16-element kernel: its groups per workgroup.
static void f10_retile(PassPlan &pp)
{
    PassArgs &a = pp.a;
    a.Gbits = f10_gbits(a);
    if (a.S > 0)
        pp.ntiles = (u64)a.ncols * (1ull << a.Jbits) * ((1ull << a.S) >> a.Gbits);
    else
        pp.ntiles = (u64)a.ncols * ((1ull << a.Jbits) >> a.Gbits);
}

static void launch_f10(const PassPlan &pp, bool inv, hipStream_t st)
{
    inv ? launch_fast10<8, true>(pp.a, pp.ntiles, st) : launch_fast10<8, false>(pp.a, pp.ntiles, st);
}

template <int B>
constexpr size_t fast_lds_bytes(u32 G)
{
    return (size_t)G * ((1u << (B / 2)) * ((1u << ((B + 1) / 2)) + 1) + 2) * sizeof(u64);
}

template <int B, bool INV>
static void launch_fast(const PassArgs &a, u64 ntiles, hipStream_t st)
{
    constexpr int BA = (B + 1) / 2, BB = B / 2;
    const size_t lds = fast_lds_bytes<B>(1u << a.Gbits);
    if (lds > 65536)
        ZK_HIP(hipFuncSetAttribute((const void *)ntt_fast_kernel<BA, BB, INV>,
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((ntt_fast_kernel<BA, BB, INV>), dim3((u32)ntiles), dim3(kThreads), lds, st, a);
    ZK_HIP(hipGetLastError());
}

static size_t fast_lds_for(u32 B, u32 G)
{
    const u32 NA = 1u << ((B + 1) / 2), NB = 1u << (B / 2);
    return (size_t)G * (NB * (NA + 1) + 2) * sizeof(u64);
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

int ntt_plan(u32 lg, u32 bits[4])
{
    if (lg <= 30) {  // fast passes of <= 10 bits
        const int np = lg <= 10 ? 1 : (lg <= 20 ? 2 : 3);
        u32 left = lg;
        for (int p = 0; p < np; p++) {
            bits[p] = (left + (np - p) - 1) / (np - p);
            left -= bits[p];
        }
        return np;
    }
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

std::vector<PassPlan> plan_passes(const NttJob &job, const u64 *pw, const u64 *gpw, u64 *scratch, size_t lds_budget)
{
    std::vector<PassPlan> out;
    u32 bits[4];
    const int np = ntt_plan(job.lg, bits);
    const char *fe = getenv("ZK_NTT_FORCE_EXACT");
    const bool force_exact = fe && fe[0] == '1';
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
        a.pw = pw;
        a.gpw = gpw;
        a.flags = force_exact ? F_FORCE_EXACT : 0;
        if (p == 0 && job.coset_in) a.flags |= F_COSET_IN;
        if (p == np - 1) {
            if (job.coset_out) a.flags |= F_COSET_OUT;
            if (job.scale_out) {
                a.flags |= F_SCALE_OUT;
                a.scale = gl_inv(N % ZK_P);
            }
        }
        // buffers: pass 0 reads the caller's input; the last pass writes dst
        if (p == 0) {
            a.src = job.src;
            a.src_stride = job.src_stride;
        } else {
            a.src = scratch ? scratch : job.dst;
            a.src_stride = scratch ? N : job.dst_stride;
        }
        if (p == np - 1) {
            a.dst = job.dst;
            a.dst_stride = job.dst_stride;
        } else {
            a.dst = scratch ? scratch : job.dst;
            a.dst_stride = scratch ? N : job.dst_stride;
        }
        const bool fast = job.lg <= 30;  // plan made of <= 10-bit passes
        auto gbits = [&](u64 extent) -> u32 {
            if (!fast) return pick_gbits(B, extent, lds_budget, a.pitch);
            u32 gb = 0;
            while (gb < 6 && fast_lds_for(B, 2u << gb) <= 69632 && (2ull << gb) <= extent) gb++;
            return gb;
        };
        u64 fast_extent, ntiles;
        if (S > 0) {
            fast_extent = 1ull << S;
            a.Gbits = gbits(fast_extent);
            ntiles = (u64)job.ncols * (1ull << a.Jbits) * (fast_extent >> a.Gbits);
        } else if (a.Jbits > 0) {
            fast_extent = 1ull << a.Jbits;
            a.Gbits = gbits(fast_extent);
            ntiles = (u64)job.ncols * (fast_extent >> a.Gbits);
        } else {
            fast_extent = job.ncols;
            a.Gbits = gbits(fast_extent);
            ntiles = (job.ncols + (1u << a.Gbits) - 1) >> a.Gbits;
        }
        // first pass of a zero-extended input: the first L - lg_in DIT stages
        // of every step-A block only replicate
        a.skip = (S == 0 && fast) ? std::min<u32>(job.lg - job.lg_in, (B + 1) / 2) : 0;
        out.push_back(PassPlan{a, B, ntiles, fast});
        Sprev = S;
        S += B;
    }
    return out;
}

// Inter-pass twiddles of a pass after the first (bits [S, S+B) of the index,
// previous pass starting at bit Sprev): element (b, J) of the pass's input,
// b < 2^S the fast index and J < 2^(L-S) the group index, is multiplied by
// omega_{2^L}^((b >> Sprev) 2^Sprev rev_{L-S}(J)).  Table entry
// J * 2^(S - Sprev) + (b >> Sprev): 2^(L - Sprev) canonical values.
ZK_HD u64 twiddle_entry(const u64 *pw, u32 L, u32 S, u32 Sprev, u64 i)
{
    const u32 bpbits = S - Sprev;
    const u64 J = i >> bpbits, bp = i & ((1ull << bpbits) - 1);
    const u64 e = ((bp << Sprev) * zk_bitrev((u32)J, L - S)) & ((1ull << L) - 1);
    return root_pow(pw, L, e);
}

__global__ __launch_bounds__(256) void twiddle_table_kernel(u64 *out, const u64 *pw, u32 L, u32 S, u32 Sprev, u64 n)
{
    const u64 i = (u64)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = twiddle_entry(pw, L, S, Sprev, i);
}

constexpr u32 kMaxTwiddleTableBits = 22;  // 32 MiB per table at most

// The table for (direction, L, S, Sprev), built on `st` at first use and kept
// for the process; null (the pass computes its twiddles) when it would exceed
// 2^kMaxTwiddleTableBits entries or would have to be built while `st` is being
// captured into a graph.
static const u64 *twiddle_table(DeviceCtx &c, bool inv, u32 L, u32 S, u32 Sprev, hipStream_t st)
{
    if (L - Sprev > kMaxTwiddleTableBits) return nullptr;
    const uint64_t key = ((uint64_t)inv << 24) | ((uint64_t)L << 16) | ((uint64_t)S << 8) | Sprev;
    std::lock_guard<std::mutex> lk(c.tw_mu);
    auto it = c.tw_tables.find(key);
    if (it != c.tw_tables.end()) return it->second;
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
    u64 *scratch = nullptr;
    const size_t scratch_bytes = (size_t)job.ncols * ((size_t)1 << job.lg) * sizeof(u64);
    if (job.src == job.dst && np > 1) {
        scratch = (u64 *)c.get_scratch(scratch_bytes);
    } else if (blocked) {
        // optional: without room for the scratch the passes run in place
        try {
            scratch = (u64 *)c.get_scratch(scratch_bytes);
        } catch (const Error &) {
            (void)hipGetLastError();
            blocked = false;
        }
    }
    auto passes = plan_passes(job, c.pw[job.inverse ? 1 : 0], c.gpw, scratch, c.lds_budget);
    for (PassPlan &pp : passes) {
        if (pp.fast && pp.a.S > 0) pp.a.tw = twiddle_table(c, job.inverse, pp.a.L, pp.a.S, pp.a.Sprev, st);
        if (pp.fast && use_fast10(pp.B, pp.a)) f10_retile(pp);
    }
    if (blocked && use_fast10(passes[0].B, passes[0].a) && use_fast10(passes[1].B, passes[1].a) &&
        passes[0].a.Gbits == 3 && passes[1].a.Gbits == 3 && passes[1].a.Jbits == 0) {
        passes[0].a.flags |= F_BLOCK_OUT;
        passes[1].a.flags |= F_BLOCK_IN;
    }
    for (size_t p = 0; p < passes.size(); p++) {
        const PassPlan &pp = passes[p];
        ProfScope ps(pass_name(job.prof_tag, (int)p), st);
        if (pp.fast && use_fast10(pp.B, pp.a))
            launch_f10(pp, job.inverse, st);
        else if (pp.fast)
            dispatch_fast(pp.B, job.inverse, pp.a, pp.ntiles, st);
        else
            dispatch_pass(pp.B, pp.a, pp.ntiles, st);
    }
}

// ---------------------------------------------------------------- transpose
// 64 x 64 tiles through LDS (padded rows: conflict-free column reads): the
// loads read 512-byte column runs, the stores write 512-byte runs of output
// row rev(i) (reference: kernels.cu:94-135, which uses 32 x 32 tiles).
// Tiles are taken in XCD-contiguous order like the NTT passes.
constexpr int TI = 64, TJ = 64;

__global__ __launch_bounds__(256) void transpose_rev_kernel(const u64 *__restrict__ in, u64 *__restrict__ out,
                                                            u32 lg, u64 ncols, u64 tiles_i)
{
    __shared__ u64 tile[TJ][TI + 1];
    const u64 N = 1ull << lg;
    const u64 b = xcd_tile(blockIdx.x, gridDim.x);
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
