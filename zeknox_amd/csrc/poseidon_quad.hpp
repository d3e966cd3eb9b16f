// Poseidon permutation spread over a quad of lanes, for latency-bound callers
// (the top levels of a Merkle tree, small trees): same function as
// poseidon_permute, about a third of the latency of one lane per state.
//
// Lane j (= lane & 3) of a quad holds state elements 3j, 3j+1, 3j+2.  A round
// is: S-box on the lane's elements (full rounds) or on element 0 only (partial
// rounds, lane 0), then the lane gathers the other nine elements of its quad
// with DPP quad permutes, t[e] = s[(3j + e) mod 12], and computes its three
// MDS rows: row 3j + k = sum_i c_i t[(i + k) mod 12] (+ 8 t[0] for row 0) +
// the next round's constant.  Because 3 * 4 = 12, the rotated index needs no
// per-lane fix-up, so the MDS coefficients stay immediates.
//
// The S-box and MDS phases are host-callable (a host emulation runs the four
// lanes in lock step); only the gather is device code.
#pragma once
#include "poseidon.hpp"

namespace zk {

// Phase A of a round: S-box on the lane's elements (all three in a full
// round, element 0 of lane 0 in a partial round).
template <bool SAFE>
ZK_HD void quad_sbox(u64 s[3], u32 j, bool full, bool &rare)
{
#pragma unroll
    for (int k = 0; k < 3; k++) {
        if (full || k == 0) {
            const u64 x = s[k];
            u64 y;
            if (SAFE) {
                const u64 x2 = gl_mul(x, x), x4 = gl_mul(x2, x2);
                y = gl_mul(gl_mul(x, x2), x4);
            } else {
                bool r = false;
                y = sbox7(x, r);
                if (full || j == 0) rare |= r;
            }
            if (full || j == 0) s[k] = y;
        }
    }
}

// Phase B: the lane's three MDS rows from the gathered state t[e] =
// s[(3j + e) mod 12], plus the next round's constants rk[k] (= constant
// 3j + k of that round).  Exact
// (the limb sums stay below 2^42 and the reduction has no rare case); SAFE
// canonicalises the outputs.
template <bool SAFE>
ZK_HD void quad_mds(const u64 t[12], u32 j, const u64 rk[3], u64 s[3])
{
    u32 lo[12], hi[12];
#pragma unroll
    for (int e = 0; e < 12; e++) {
        lo[e] = (u32)t[e];
        hi[e] = (u32)(t[e] >> 32);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) {
        u64 a = (u32)rk[k], b = rk[k] >> 32;
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const u64 c = mds_coeff(i);
            a += c * lo[(i + k) % 12];
            b += c * hi[(i + k) % 12];
        }
        if (k == 0) {  // diag[0] = 8 on row 0 (lane 0), 0 elsewhere
            const u64 dg = j == 0 ? PoseidonConsts::diag[0] : 0;
            a += dg * lo[0];
            b += dg * hi[0];
        }
        const u64 v = reduce_lohi42(a, b);
        s[k] = SAFE ? gl_canon(v) : v;
    }
}

// Value of x held by lane (j + D) & 3 of the same quad (DPP quad_perm).
template <int D>
ZK_DEV u64 quad_fetch(u64 x)
{
#ifdef __HIP_DEVICE_COMPILE__
    constexpr int ctrl = ((0 + D) & 3) | (((1 + D) & 3) << 2) | (((2 + D) & 3) << 4) | (((3 + D) & 3) << 6);
    const u32 lo = (u32)__builtin_amdgcn_mov_dpp((int)(u32)x, ctrl, 0xF, 0xF, false);
    const u32 hi = (u32)__builtin_amdgcn_mov_dpp((int)(u32)(x >> 32), ctrl, 0xF, 0xF, false);
    return ((u64)hi << 32) | lo;
#else
    return x;  // device-only function; the host pass never runs it
#endif
}

// The whole permutation for lane j of a quad (all four lanes must execute
// it together).  Relaxed in and out unless SAFE (canonical in and out).
template <bool SAFE>
ZK_DEV void poseidon_permute_quad(u64 s[3], u32 j, bool &rare)
{
    const u64 *rc = round_constants();
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] = SAFE ? gl_add(s[k], rc[3 * j + k]) : gl_add_nc(s[k], rc[3 * j + k], rare);
    // this lane's constants are vector loads (lane-dependent index): fetch
    // them one round ahead so their latency hides behind the S-box
    u64 rk[3];
#pragma unroll
    for (int k = 0; k < 3; k++) rk[k] = rc[12 + 3 * j + k];
#pragma unroll 1
    for (int r = 0; r < 30; r++) {
        const int rn = r + 2 < 31 ? r + 2 : 30;  // row 30 is the all-zero tail
        u64 rk_next[3];
#pragma unroll
        for (int k = 0; k < 3; k++) rk_next[k] = rc[12 * rn + 3 * j + k];
        quad_sbox<SAFE>(s, j, r < 4 || r >= 26, rare);
        u64 t[12];
        t[0] = s[0];
        t[1] = s[1];
        t[2] = s[2];
        t[3] = quad_fetch<1>(s[0]);
        t[4] = quad_fetch<1>(s[1]);
        t[5] = quad_fetch<1>(s[2]);
        t[6] = quad_fetch<2>(s[0]);
        t[7] = quad_fetch<2>(s[1]);
        t[8] = quad_fetch<2>(s[2]);
        t[9] = quad_fetch<3>(s[0]);
        t[10] = quad_fetch<3>(s[1]);
        t[11] = quad_fetch<3>(s[2]);
        quad_mds<SAFE>(t, j, rk, s);
#pragma unroll
        for (int k = 0; k < 3; k++) rk[k] = rk_next[k];
    }
}

// State element 3j + k of a two-to-one compression [a, b, 0, 0, 0, 0].
ZK_DEV u64 quad_two_input(const u64 *a, const u64 *b, u32 j, int k)
{
    const u32 e = 3 * j + (u32)k;
    return e < 4 ? a[e] : (e < 8 ? b[e - 4] : 0);
}

// H(a, b) by a quad: every lane of the quad calls it (a, b read by the lanes
// that need them); lane 0 returns digest elements 0..2 in out[0..2], lane 1
// element 3 in out[0].  All lanes of the wave must call it together (the
// rare-case recomputation is wave-uniform).
ZK_DEV void hash_two_quad(const u64 *a, const u64 *b, u32 j, u64 out[3])
{
    u64 s[3];
#pragma unroll
    for (int k = 0; k < 3; k++) s[k] = quad_two_input(a, b, j, k);
    bool rare = false;
    poseidon_permute_quad<false>(s, j, rare);
    if (__any(rare)) {
#pragma unroll
        for (int k = 0; k < 3; k++) s[k] = gl_canon(quad_two_input(a, b, j, k));
        bool unused = false;
        poseidon_permute_quad<true>(s, j, unused);
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_canon(s[k]);
}

// hash_or_noop of n elements fetched by get(i), by a quad: lane j absorbs
// the rate elements 3j..3j+2 of each chunk of 8 (lane 2 only 6 and 7, lane 3
// none), overwrite mode.  Lane 0 returns digest elements 0..2 in out[0..2],
// lane 1 element 3 in out[0].  n <= 4 copies the input (zero padded).  All
// lanes of the wave must call it together with the same n.
template <typename Get>
ZK_DEV void hash_one_quad(Get get, u32 n, u32 j, u64 out[3])
{
    if (n <= 4) {
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const u32 e = 3 * j + (u32)k;
            out[k] = e < n ? get(e) : 0;
        }
        return;
    }
    u64 s[3] = {0, 0, 0};
    bool rare = false;
    for (u32 idx = 0; idx < n; idx += 8) {
        const u32 rem = n - idx;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const u32 e = 3 * j + (u32)k;
            if (e < 8 && e < rem) s[k] = get(idx + e);
        }
        poseidon_permute_quad<false>(s, j, rare);
    }
    if (__any(rare)) {
        s[0] = s[1] = s[2] = 0;
        for (u32 idx = 0; idx < n; idx += 8) {
            const u32 rem = n - idx;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const u32 e = 3 * j + (u32)k;
                if (e < 8 && e < rem) s[k] = gl_canon(get(idx + e));
            }
            bool unused = false;
            poseidon_permute_quad<true>(s, j, unused);
        }
    }
#pragma unroll
    for (int k = 0; k < 3; k++) out[k] = gl_canon(s[k]);
}

}  // namespace zk
