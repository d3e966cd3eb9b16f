// Per-thread code of the NTT / coset LDE passes (ntt.hip): kernel arguments,
// tile decoding, the register DITs with power-of-two twiddles, the step
// functions of the fast passes and of the 10-bit pass, and the pass plan.
// Everything here is host-and-device code, so a host build can run a pass
// thread by thread (tests/native/ntt_host_test.cpp); the kernels that call
// it, their launches and the device tables are in ntt.hip.
#pragma once
#include <algorithm>
#include <cstdlib>
#include <mutex>
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
    // 10-bit passes: tb[64 v + j] = omega_1024^(rev4(v) j) for this direction
    const u64 *tb;
    // first tile of this launch: a pass whose grid would exceed 2^32 threads
    // (HIP counts a grid in threads, as a 32-bit number) is launched in chunks
    u64 tile_base;
};

struct PassPlan {
    PassArgs a;
    u32 B;
    u64 ntiles;
    bool fast;
    bool r10;  // runs the 16-element 10-bit kernel (ntt_r10_kernel)
};

// Tuned plan of the transforms of one size 2^lg (lg <= 30), chosen by the
// autotuner (zeknox_amd/tune.py) and loaded from its table when a device is
// first used: the bits of each pass, each pass's groups per workgroup
// (log2; kTuneAuto keeps the most the pass's extent and LDS allow) and, for
// a 10-bit pass, whether the 16-element kernel may run it (else the
// 32-element kernel does).  np == 0: no entry, the default plan.
constexpr u32 kTuneAuto = 0xFFFFFFFFu;
struct NttTune {
    u32 np;
    u32 bits[3];
    u32 gbits[3];
    u32 r10[3];
};

inline std::mutex &ntt_tune_mutex()
{
    static std::mutex m;
    return m;
}

inline NttTune *ntt_tune_table()
{
    static NttTune t[31] = {};
    return t;
}

// The entry for lg (np == 0 when there is none).
inline NttTune ntt_tune_get(u32 lg)
{
    if (lg > 30) return NttTune{};
    std::lock_guard<std::mutex> lk(ntt_tune_mutex());
    return ntt_tune_table()[lg];
}

// Whether an entry is a plan the passes can run: 1..3 passes of 1..10 bits
// summing to lg, groups per workgroup auto or <= 6, 10-bit kernel flags 0/1.
inline bool ntt_tune_valid(u32 lg, const NttTune &t)
{
    if (lg == 0 || lg > 30 || t.np < 1 || t.np > 3) return false;
    u32 sum = 0;
    for (u32 p = 0; p < t.np; p++) {
        if (t.bits[p] < 1 || t.bits[p] > 10 || (t.gbits[p] != kTuneAuto && t.gbits[p] > 6) || t.r10[p] > 1)
            return false;
        sum += t.bits[p];
    }
    return sum == lg;
}

// Kernel arguments of every pass of a transform (no launch): pw is the
// direction's power table, scratch the ping-pong buffer of in-place plans
// (plan_passes, below).

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
    // no branch on the digits (entry 0 of each row is 1): the three loads of
    // every element of a thread can be in flight together
    const u64 r0 = gpw[n & (kPowLen - 1)];
    const u64 r1 = gpw[kPowLen + ((u32)(n >> 11) & (kPowLen - 1))];
    const u64 r2 = gpw[2 * kPowLen + (u32)(n >> 22)];
    return gl_mul(gl_mul(r0, r1), r2);
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
        if (a.flags & F_COSET_IN) v = gl_mul(v, coset_pow(a.gpw, n));
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
            if (a.flags & F_COSET_OUT) v = gl_mul(v, coset_pow(a.gpw, pos));
        }
        a.dst[(u64)col * a.dst_stride + pos] = v;
    }
};

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

// y * 2^E up to sign, for a constant E in [0, 192): returns t and sets `neg`
// with y 2^E == (neg ? -t : t).  2^96 == -1 folds E into [0, 96); on the
// relaxed path 2^e for e in [64, 96) is -2^-(96-e), a cheaper shift down.
template <int E, bool SAFE>
ZK_HD u64 f_mul_root2_signed(u64 y, bool &neg, bool &rare)
{
    constexpr int e = E % 96;
    if constexpr (!SAFE && e >= 64) {
        neg = E < 96;
        return gl_shr_nc(y, 96 - e);
    } else {
        neg = E >= 96;
        return f_mul_pow2<SAFE>(y, e, rare);
    }
}

// y * 2^E for a constant E in [0, 192), sign applied.
template <int E, bool SAFE>
ZK_HD u64 f_mul_root2(u64 y, bool &rare)
{
    if constexpr (E == 0) {
        return y;
    } else {
        bool neg;
        const u64 t = f_mul_root2_signed<E, SAFE>(y, neg, rare);
        if (!neg) return t;
        return SAFE ? gl_neg(t) : gl_neg_nc(t, rare);
    }
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
                bool neg;
                const u64 t = f_mul_root2_signed<E, SAFE>(v[m + (1 << s)], neg, rare);
                const u64 x = v[m];
                const u64 a = f_add<SAFE>(x, t, rare), b = f_sub<SAFE>(x, t, rare);
                if (neg) {
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

    // First-pass loads of block q: element t is zero when its low SKIP bits
    // are set (the LDE's replication stages) or past the input; the others
    // are loaded with no lane-dependent branch (a lane past the input reads
    // a.src[0], always present, and keeps 0), so all are in flight together.
    template <int SKIP, bool SAFE>
    ZK_HD static void load_first(const PassArgs &a, const u64 *src, u32 q, u64 nlow, u64 lim, bool colok, u64 v[NA])
    {
        static_for<0, NA>([&](auto tc) {
            constexpr int t = decltype(tc)::value;
            if constexpr ((t & ((1 << SKIP) - 1)) != 0) {
                v[t] = 0;
            } else {
                const u64 n = ((u64)zk_bitrev((u32)(q * NA + t), B) << a.Jbits) + nlow;
                const bool ok = colok && n < lim;
                const u64 x = *(ok ? src + n : a.src);
                v[t] = ok ? (SAFE ? gl_canon(x) : x) : 0;
            }
        });
    }

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
            switch (a.skip) {  // uniform
            case 0: load_first<0, SAFE>(a, src, q, nlow, lim, colok, v); break;
            case 1: load_first<1, SAFE>(a, src, q, nlow, lim, colok, v); break;
            case 2: load_first<2, SAFE>(a, src, q, nlow, lim, colok, v); break;
            case 3: load_first<3, SAFE>(a, src, q, nlow, lim, colok, v); break;
            case 4: load_first<4, SAFE>(a, src, q, nlow, lim, colok, v); break;
            default: load_first<5, SAFE>(a, src, q, nlow, lim, colok, v); break;
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
                if (a.flags & F_COSET_OUT) x = gl_mul(x, coset_pow(a.gpw, pos));
            }
            dst[pos] = x;
        }
    }
}

// ---------------------------------------------------------------------------
// 10-bit pass with 16 elements per thread (the 2^10-point passes of LDEs and
// NTTs of 2^19 .. 2^20 points and of 2^28 .. 2^30).  Element J = t + 16 u +
// 64 v of a group (bit-reversed input order; t, v < 16, u < 4) and output
// index k = kA + 16 kB + 64 kC (kA, kC < 16, kB < 4), with rt = rev4(t),
// ru = rev2(u), rv = rev4(v):
//   X[k] = sum_v w16^(rv kC) w1024^(rv (kA + 16 kB))
//            sum_u w4^(ru kB) w64^(ru kA) sum_t w16^(rt kA) x[J]
// Step A: a 16-point DIT over t, then w64^(ru kA), a power of two (w64 =
// 2^39): u is the same across a wave, so each of the four values of ru has
// its own straight-line code with constant shifts.  LDS.  Step B: four
// 4-point DITs over u, then w1024^(rv (kA + 16 kB)), a full multiplication
// by an entry of a per-device 16 x 64 table.  LDS.  Step C: a 16-point DIT
// over v.  Only one layer of full multiplications besides the inter-pass
// twiddle.  A workgroup of 512 threads holds 8 groups (adjacent fast
// indices); 16 elements per thread keep the registers under 128 and two
// workgroups fit a CU's LDS, so four waves per SIMD.  Arithmetic is relaxed;
// each wave records whether any lane met a rare case, and after the last step
// a flagged tile is recomputed on the canonical path before anything is stored
// (the tile's inputs are still intact, also in place).  Pass layout (tiles,
// loads, stores, fused coset / twiddle / scaling) is that of FastPass.
//
// LDS: entry (g, kA, X, v) (X = u after step A, kB after step B) at
// g R10_GP + kA R10_KP + 16 X + v.  Thread roles: g = tid & 7, y = (tid >> 3)
// & 15 (v in steps A and B, kA in step C), w = tid >> 7 (u, kA / 4, kB).  A
// 64-bit LDS access is served 16 lanes at a time; those are 8 g x one bit of
// y, at word offsets 2 g + y0 (mod 16) for every step (R10_GP == 2 mod 16,
// R10_KP odd): no bank conflicts.
constexpr u32 kR10KP = 65;
constexpr u32 kR10GP = 16 * kR10KP + 2;
constexpr u32 kR10Groups = 8;
constexpr u32 kR10Threads = 512;
constexpr size_t kR10Lds = ((size_t)kR10Groups * kR10GP + kR10Threads / 64) * sizeof(u64);

// w64^(ru kA) of step A for the wave's ru, kA = 1..15 (constant exponents
// per ru branch).
template <int RU, bool INV, bool SAFE>
ZK_HD void r10_w64(u64 v[16], bool &rare)
{
    static_for<1, 16>([&](auto kc) {
        constexpr int kA = decltype(kc)::value;
        constexpr int E = (root_exp2(6, INV) * RU * kA) % 192;
        v[kA] = f_mul_root2<E, SAFE>(v[kA], rare);
    });
}

// First-pass loads of step A: element t is zero when its low SKIP bits are set
// (the LDE's replication stages) or past the input; the others are loaded
// with no lane-dependent branch (a lane past the input reads a.src[0], always
// present, and keeps 0), so all of them are in flight together.
template <int SKIP, bool SAFE>
ZK_HD void r10_load_first(const PassArgs &a, const u64 *src, u64 n0, u64 lim, bool colok, u64 v[16])
{
    static_for<0, 16>([&](auto tc) {
        constexpr int t = decltype(tc)::value;
        if constexpr ((t & ((1 << SKIP) - 1)) != 0) {
            v[t] = 0;
        } else {
            const u64 dn = (u64)(64 * ct_bitrev(t, 4)) << a.Jbits;
            const bool ok = colok && n0 + dn < lim;
            const u64 x = *(ok ? src + dn : a.src);
            v[t] = ok ? (SAFE ? gl_canon(x) : x) : 0;
        }
    });
}

// Step A of thread tid: load 16 elements of one group, inter-pass twiddle or
// coset factor, 16-point DIT, w64 twiddle, into LDS.
template <bool INV, bool SAFE>
ZK_HD void r10_step_a(const PassArgs &a, u32 col, u64 jhi, u64 fast0, u32 tid, u64 *lds, bool &rare)
{
    const u32 g = tid & 7, v_ = (tid >> 3) & 15, u = tid >> 7;
    const u32 J0 = 16 * u + 64 * v_;
    u64 v[16];
    if (a.S > 0) {
        const u64 b = fast0 + g;
        const u64 Jhi = jhi << 10;
        // per-thread base pointers, per-element offsets uniform across the
        // workgroup (rev10(J0 + t) = rev10(J0) + 64 rev4(t))
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
            // all twiddle loads are issued together with the data loads
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
        switch (a.skip) {  // uniform
        case 0: r10_load_first<0, SAFE>(a, src, n0, lim, colok, v); break;
        case 1: r10_load_first<1, SAFE>(a, src, n0, lim, colok, v); break;
        case 2: r10_load_first<2, SAFE>(a, src, n0, lim, colok, v); break;
        case 3: r10_load_first<3, SAFE>(a, src, n0, lim, colok, v); break;
        default: r10_load_first<4, SAFE>(a, src, n0, lim, colok, v); break;
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
    switch (zk_bitrev(u, 2)) {  // uniform across the wave
    case 1: r10_w64<1, INV, SAFE>(v, rare); break;
    case 2: r10_w64<2, INV, SAFE>(v, rare); break;
    case 3: r10_w64<3, INV, SAFE>(v, rare); break;
    default: break;
    }
    u64 *row = lds + g * kR10GP + 16 * u + v_;
#pragma unroll
    for (int kA = 0; kA < 16; kA++) row[kA * kR10KP] = v[kA];
}

// Step B of thread tid: four 4-point DITs over u (kA = 4 (tid >> 7) + i),
// w1024 twiddle, in place (each thread rewrites exactly the 16 entries it read).
template <bool INV, bool SAFE>
ZK_HD void r10_step_b(const PassArgs &a, u32 tid, u64 *lds, bool &rare)
{
    const u32 g = tid & 7, v_ = (tid >> 3) & 15, kAh = tid >> 7;
    u64 *base = lds + g * kR10GP + (4 * kAh) * kR10KP + v_;
    const u64 *tb = a.tb + 64 * v_ + 4 * kAh;  // w1024^(rv j), j = kA + 16 kB
#pragma unroll
    for (int i = 0; i < 4; i++) {
        u64 y[4];
#pragma unroll
        for (int x = 0; x < 4; x++) y[x] = SAFE ? gl_canon(base[i * kR10KP + 16 * x]) : base[i * kR10KP + 16 * x];
        dit_pow2<2, INV, SAFE>(y, 0, rare);
#pragma unroll
        for (int kB = 0; kB < 4; kB++) base[i * kR10KP + 16 * kB] = f_mul<SAFE>(y[kB], tb[i + 16 * kB], rare);
    }
}

// Step C of thread tid (kA = (tid >> 3) & 15, kB = tid >> 7): 16-point DIT
// over v; out[kC] is output index kA + 16 kB + 64 kC.
template <bool INV, bool SAFE>
ZK_HD void r10_step_c(const PassArgs &a, u32 tid, const u64 *lds, bool last, u64 out[16], bool &rare)
{
    const u32 g = tid & 7, kA = (tid >> 3) & 15, kB = tid >> 7;
    const u64 *row = lds + g * kR10GP + kA * kR10KP + 16 * kB;
#pragma unroll
    for (int v = 0; v < 16; v++) out[v] = SAFE ? gl_canon(row[v]) : row[v];
    dit_pow2<4, INV, SAFE>(out, 0, rare);
    if (last && (a.flags & F_SCALE_OUT)) {
#pragma unroll
        for (int kC = 0; kC < 16; kC++) out[kC] = f_mul<SAFE>(out[kC], a.scale, rare);
    }
}

ZK_HD void r10_store(const PassArgs &a, u32 col, u64 jhi, u64 fast0, u32 tid, bool last, const u64 out[16])
{
    const u32 g = tid & 7, kA = (tid >> 3) & 15, kB = tid >> 7;
    if (col >= a.ncols) return;
    // output index k = k0 + 64 kC; position = pos0 + (uniform offset)
    const u64 k0 = kA + 16 * kB;
    u64 pos0;
    u32 sh;
    if (a.S > 0) {
        pos0 = fast0 + g + ((k0 + (jhi << 10)) << a.S);
        sh = a.S + 6;
    } else if (a.flags & F_BLOCK_OUT) {  // blocked for the next pass (its b = k, rev(J) = f)
        pos0 = ((k0 >> 3) << (a.Jbits + 3)) + ((fast0 + g) << 3) + (k0 & 7);
        sh = a.Jbits + 6;  // 64 kC >> 3 = 8 kC blocks of 2^(Jbits+3)
    } else {
        pos0 = k0 + ((u64)zk_bitrev((u32)(fast0 + g), a.Jbits) << 10);
        sh = 6;
    }
    u64 *dst = a.dst + (u64)col * a.dst_stride + pos0;
#pragma unroll
    for (int kC = 0; kC < 16; kC++) {
        const u64 d = (u64)kC << sh;
        u64 x = out[kC];
        if (last) {
            x = gl_canon(x);
            const u64 pos = pos0 + d;
            if (a.flags & F_COSET_OUT) x = gl_mul(x, coset_pow(a.gpw, pos));
        }
        dst[d] = x;
    }
}

// Groups per workgroup of the 16-element pass (log2): 8 groups, two
// workgroups per CU.
inline u32 f10_gbits(const PassArgs &) { return 3; }

// The 16-element pass applies to 10-bit passes whose workgroup holds the
// groups of one column: passes after the first, and first passes of
// multi-pass transforms (a single 10-bit pass keeps the 32-element kernel),
// when the fast dimension has at least that many groups.
inline bool use_fast10(u32 B, const PassArgs &a)
{
    const u64 extent = a.S > 0 ? (1ull << a.S) : (a.Jbits > 0 ? (1ull << a.Jbits) : 0);
    return B == 10 && extent >= (1ull << f10_gbits(a));
}

// Re-tiles a pass for the 16-element kernel: its groups per workgroup.
inline void f10_retile(PassPlan &pp)
{
    PassArgs &a = pp.a;
    a.Gbits = f10_gbits(a);
    if (a.S > 0)
        pp.ntiles = (u64)a.ncols * (1ull << a.Jbits) * ((1ull << a.S) >> a.Gbits);
    else
        pp.ntiles = (u64)a.ncols * ((1ull << a.Jbits) >> a.Gbits);
}

template <int B>
constexpr size_t fast_lds_bytes(u32 G)
{
    return (size_t)G * ((1u << (B / 2)) * ((1u << ((B + 1) / 2)) + 1) + 2) * sizeof(u64);
}

inline size_t fast_lds_for(u32 B, u32 G)
{
    const u32 NA = 1u << ((B + 1) / 2), NB = 1u << (B / 2);
    return (size_t)G * (NB * (NA + 1) + 2) * sizeof(u64);
}

inline int ntt_plan_bits(u32 lg, u32 bits[4])
{
    const NttTune t = ntt_tune_get(lg);
    if (t.np) {
        for (u32 p = 0; p < t.np; p++) bits[p] = t.bits[p];
        return (int)t.np;
    }
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


// Groups per workgroup (log2) and LDS pitch: as many groups as the fast
// dimension and the LDS budget allow (up to 64), rows padded by one element
// when the padding still fits.
inline u32 pick_gbits(u32 B, u64 fast_extent, size_t budget, u32 &pitch)
{
    u32 gb = 0;
    while (gb < 6 && ((size_t)(2u << gb) << B) * sizeof(u64) <= budget && (2ull << gb) <= fast_extent) gb++;
    pitch = ((size_t)(1u << gb) * ((1u << B) + 1) * sizeof(u64) <= budget) ? (1u << B) + 1 : (1u << B);
    return gb;
}

inline std::vector<PassPlan> plan_passes(const NttJob &job, const u64 *pw, const u64 *gpw, u64 *scratch,
                                        size_t lds_budget)
{
    std::vector<PassPlan> out;
    u32 bits[4];
    const int np = ntt_plan_bits(job.lg, bits);
    const NttTune tune = ntt_tune_get(job.lg);
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
        // a tuned pass takes fewer groups per workgroup than the most it can
        if (tune.np && tune.gbits[p] != kTuneAuto && tune.gbits[p] < a.Gbits) {
            a.Gbits = tune.gbits[p];
            if (S > 0)
                ntiles = (u64)job.ncols * (1ull << a.Jbits) * (fast_extent >> a.Gbits);
            else if (a.Jbits > 0)
                ntiles = (u64)job.ncols * (fast_extent >> a.Gbits);
            else
                ntiles = (job.ncols + (1u << a.Gbits) - 1) >> a.Gbits;
        }
        // first pass of a zero-extended input: the first L - lg_in DIT stages
        // of every step-A block only replicate
        a.skip = (S == 0 && fast) ? std::min<u32>(job.lg - job.lg_in, (B + 1) / 2) : 0;
        PassPlan pp{a, B, ntiles, fast, false};
        pp.r10 = fast && (!tune.np || tune.r10[p]) && use_fast10(B, a);
        if (pp.r10) f10_retile(pp);
        out.push_back(pp);
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


}  // namespace zk
