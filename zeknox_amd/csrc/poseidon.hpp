// Poseidon permutation over Goldilocks, width 12 (rate 8, capacity 4),
// 8 full rounds + 22 partial rounds, S-box x^7, and the Plonky2 sponge
// ("hash_or_noop" and two-to-one compression) used by okx/zeknox's
// PoseidonHasher (native/poseidon/poseidon.cu:771-780,834-838,
// native/poseidon/poseidon_permutation.cuh:83-134).
//
// One permutation per lane: the 12-element state lives in 24 VGPRs; MDS
// coefficients are immediates, round constants come from constant memory.
// Inside the permutation values are kept "relaxed" (any u64 congruent mod p)
// and canonicalised once at the digest, which keeps every Goldilocks
// multiplication at 4 v_mad_u64_u32 for the product plus 2 for the reduction.
#pragma once
#include "gl64.hpp"
#include "poseidon_gl_constants.h"

namespace zk {

struct PoseidonConsts {
    static constexpr u64 circ[12] = ZK_POSEIDON_MDS_CIRC;
    static constexpr u64 diag[12] = ZK_POSEIDON_MDS_DIAG;
};

// Round constants (30 rounds x 12, then 12 zeros so the last round's MDS can
// fold "the next round's constants" uniformly): scalar loads on the device.
#define ZK_RC_WITH_TAIL { ZK_POSEIDON_RC_TABLE_BODY 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0 }
static __constant__ const u64 kRoundConstDev[372] = ZK_RC_WITH_TAIL;
static const u64 kRoundConstHost[372] = ZK_RC_WITH_TAIL;

ZK_HD const u64 *round_constants()
{
#ifdef __HIP_DEVICE_COMPILE__
    return kRoundConstDev;
#else
    return kRoundConstHost;
#endif
}

// x^7 = x^3 * x^4 with relaxed (any u64) operands and result.  The squares
// use the general product: the three-product square (gl_sqr_nc) issues fewer
// instructions but its carry chain measured 1.27x slower per square on gfx950.
ZK_HD u64 sbox7(u64 x, bool &rare)
{
    const u64 x2 = gl_mul_nc<true>(x, x, rare);
    const u64 x4 = gl_mul_nc<true>(x2, x2, rare);
    const u64 x3 = gl_mul_nc<true>(x, x2, rare);
    return gl_mul_nc<true>(x3, x4, rare);
}

// a + 2^32 b mod p (relaxed) for a, b < 2^42: with b = b0 + 2^32 b1,
// a + 2^32 b == (a + b1 (2^32 - 1)) + 2^32 b0.  The first sum stays below
// 2^43 (one multiply-add), adding 2^32 b0 wraps past 2^64 at most once and
// the wrap is worth 2^32 - 1, which cannot wrap again (the wrapped value is
// below 2^43).
ZK_HD u64 reduce_lohi42(u64 a, u64 b)
{
    const u64 t = (u64)(u32)(b >> 32) * ZK_EPS + a;
    unsigned c;
    const u32 hi = __builtin_addc((u32)(t >> 32), (u32)b, 0u, &c);
    const u64 r = ((u64)hi << 32) | (u32)t;
    return r + (c ? ZK_EPS : 0);
}

// state <- M state + rcn (the next round's constants), M circulant (first row
// circ) plus diag.  Row r is rcn[r] + sum of 13 products of a state element
// with a constant below 2^6: the low and high 32-bit halves are accumulated
// separately (the sums start at the constant's halves and stay below 2^42),
// then reduced once by reduce_lohi42.  Inputs and outputs
// are relaxed (any u64).  The coefficients come from mds_coeff:
//
// the MDS coefficients as multiplier operands.  The powers of two among them
// (16 and 2) are passed through an opaque scalar register on the device: as
// literals the compiler turns c * x into a 64-bit shift of x, which first
// needs x zero-extended into a register pair (two moves) and so costs more
// than the multiply-add it replaces.
ZK_HD u32 mds_coeff(int i)
{
    const u32 c = (u32)PoseidonConsts::circ[i];
#ifdef __HIP_DEVICE_COMPILE__
    if ((c & (c - 1)) == 0) {
        u32 r;
        asm("s_mov_b32 %0, %1" : "=s"(r) : "i"(c));
        return r;
    }
#endif
    return c;
}

ZK_HD void mds_fold(u64 s[12], const u64 *rcn)
{
    u32 lo[12], hi[12];
#pragma unroll
    for (int j = 0; j < 12; j++) {
        lo[j] = (u32)s[j];
        hi[j] = (u32)(s[j] >> 32);
    }
#pragma unroll
    for (int r = 0; r < 12; r++) {
        const u64 k = rcn[r];
        u64 a = (u32)k, b = k >> 32;
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const u64 c = i == 0 ? PoseidonConsts::circ[0] + PoseidonConsts::diag[r] : mds_coeff(i);
            a += c * lo[(i + r) % 12];
            b += c * hi[(i + r) % 12];
        }
        s[r] = reduce_lohi42(a, b);
    }
}

ZK_HD void full_round(u64 s[12], const u64 *rcn, bool &rare)
{
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = sbox7(s[i], rare);
    mds_fold(s, rcn);
}

ZK_HD void partial_round(u64 s[12], const u64 *rcn, bool &rare)
{
    s[0] = sbox7(s[0], rare);
    mds_fold(s, rcn);
}

// 4 full rounds, 22 partial rounds (S-box on lane 0), 4 full rounds; round
// r's constants are added by round r-1's MDS (round 0's explicitly).  The
// round loops are not unrolled: one round body is ~1-2k instructions, so the
// whole permutation stays resident in the instruction cache.  Input: relaxed;
// output: relaxed (canonicalise with gl_canon).
ZK_HD void poseidon_permute(u64 s[12], bool &rare)
{
    const u64 *rc = round_constants();
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = gl_add_nc(s[i], rc[i], rare);
#pragma unroll 1
    for (int r = 0; r < 4; r++) full_round(s, rc + 12 * (r + 1), rare);
#pragma unroll 1
    for (int r = 4; r < 26; r++) partial_round(s, rc + 12 * (r + 1), rare);
#pragma unroll 1
    for (int r = 26; r < 30; r++) full_round(s, rc + 12 * (r + 1), rare);
}

// The same permutation with canonical arithmetic only (no rare cases): the
// recomputation path when a relaxed permutation reported a rare case.
ZK_HD void poseidon_permute_safe(u64 s[12])
{
    const u64 *rc = round_constants();
#pragma unroll 1
    for (int r = 0; r < 30; r++) {
        const bool full = r < 4 || r >= 26;
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = gl_add(gl_canon(s[i]), rc[12 * r + i]);
#pragma unroll
        for (int i = 0; i < 12; i++) {
            if (full || i == 0) {
                const u64 x2 = gl_mul(s[i], s[i]), x4 = gl_mul(x2, x2);
                s[i] = gl_mul(gl_mul(s[i], x2), x4);
            }
        }
        u64 o[12];
#pragma unroll
        for (int row = 0; row < 12; row++) {
            u64 lo = 0, hi = 0;
#pragma unroll
            for (int i = 0; i < 12; i++) {
                const u64 c = PoseidonConsts::circ[i] + (i == 0 ? PoseidonConsts::diag[row] : 0);
                u64 pl, ph;
                gl_mul_wide(s[(i + row) % 12], c, pl, ph);
                lo += pl;
                hi += ph + (lo < pl ? 1 : 0);
            }
            o[row] = gl_reduce128(lo, hi);
        }
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = o[i];
    }
}

// Two-to-one compression: state = [a, b, 0, 0, 0, 0] -> permute -> state[0..4).
ZK_HD void hash_two(const u64 a[4], const u64 b[4], u64 out[4])
{
    u64 s[12];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        s[i] = a[i];
        s[4 + i] = b[i];
        s[8 + i] = 0;
    }
    bool rare = false;
    poseidon_permute(s, rare);
    if (rare) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s[i] = a[i];
            s[4 + i] = b[i];
            s[8 + i] = 0;
        }
        poseidon_permute_safe(s);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl_canon(s[i]);
}

// hash_or_noop over n elements fetched by `get(i)`: n <= 4 copies the input,
// zero-padded; otherwise the overwrite-mode sponge (a final short chunk leaves
// the stale rate lanes in place), digest = state[0..4).  `chunk(idx, m, v)` sees
// each chunk of input once, right after it is fetched: elements idx .. idx+m-1
// in v[0..m) (the exact recompute of a rare case fetches again, unseen).
template <typename Get, typename Chunk>
ZK_HD void hash_one_chunks(Get get, u32 n, u64 out[4], Chunk chunk)
{
    if (n <= 4) {
#pragma unroll
        for (u32 i = 0; i < 4; i++) out[i] = i < n ? get(i) : 0;
        chunk(0u, n, (const u64 *)out);
        return;
    }
    u64 s[12];
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = 0;
    bool rare = false;
    for (u32 idx = 0; idx < n; idx += 8) {
        const u32 rem = n - idx;
#pragma unroll
        for (u32 i = 0; i < 8; i++)
            if (i < rem) s[i] = get(idx + i);
        chunk(idx, rem < 8 ? rem : 8u, (const u64 *)s);
        poseidon_permute(s, rare);
    }
    if (rare) {  // some lane hit a rare relaxed case: redo the sponge exactly
#pragma unroll
        for (int i = 0; i < 12; i++) s[i] = 0;
        for (u32 idx = 0; idx < n; idx += 8) {
            const u32 rem = n - idx;
#pragma unroll
            for (u32 i = 0; i < 8; i++)
                if (i < rem) s[i] = get(idx + i);
            poseidon_permute_safe(s);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = gl_canon(s[i]);
}

template <typename Get>
ZK_HD void hash_one(Get get, u32 n, u64 out[4])
{
    hash_one_chunks(get, n, out, [](u32, u32, const u64 *) {});
}

}  // namespace zk
