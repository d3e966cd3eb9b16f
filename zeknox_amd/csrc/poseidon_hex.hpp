// Poseidon permutation spread over 16 lanes, one state element per lane, for
// the latency-bound narrow levels of a Merkle tree: same function as
// poseidon_permute.
//
// Lane j (= lane & 15) of a group holds state element j for j < 12; lanes 12..15
// hold a copy of element j - 12 and their results are discarded.  A round is:
// the S-box on every lane (full rounds) or on lane 0 (partial rounds: every
// lane computes it and lane 0 keeps it, so the group does not diverge), then
// each lane gathers t[k] = s[(j + k) mod 12] for k = 1..11 from the other lanes
// of its group (ds_bpermute, two 32-bit moves per element) and computes its
// MDS row: sum_k c_k t[k] (+ 8 t[0] on row 0) + the next round's constant.
// The circulant makes the coefficient of t[k] the same on every lane, so the
// MDS coefficients stay immediates.  A lane's round is one S-box (four
// products) and one row (13 multiply-adds per 32-bit half) against three of
// each in the quad form, plus 22 lane exchanges; measured on levels of
// 2^11-2^12 nodes, 14-17 us per level against 20 us for the quad form (the
// exchanges' latency keeps it from the threefold the instruction count
// suggests).
//
// The S-box and MDS phases are host-callable, so a host emulation can run the
// 16 lanes of a group in lock step (tests/native/poseidon_lanes_test.cpp);
// only the gather is device code.
#pragma once
#include "poseidon.hpp"

namespace zk {

constexpr u32 kHexLanes = 16;

// Phase A of a round for the lane holding element j: the S-box (every lane
// in a full round, element 0 in a partial round).
template <bool SAFE>
ZK_HD u64 hex_sbox(u64 x, u32 j, bool full, bool &rare)
{
    u64 y;
    if (SAFE) {
        const u64 x2 = gl_mul(x, x), x4 = gl_mul(x2, x2);
        y = gl_mul(gl_mul(x, x2), x4);
    } else {
        bool r = false;
        y = sbox7(x, r);
        if (full || j == 0) rare |= r;
    }
    return (full || j == 0) ? y : x;
}

// Phase B: row j of the MDS from the gathered state t[k] = s[(j + k) mod 12]
// plus the next round's constant rk.  Exact (the half sums stay below 2^42 and
// the reduction has no rare case); SAFE canonicalises the output.
template <bool SAFE>
ZK_HD u64 hex_mds(const u64 t[12], u32 j, u64 rk)
{
    u64 a = (u32)rk, b = rk >> 32;
#pragma unroll
    for (int k = 0; k < 12; k++) {
        const u64 c = mds_coeff(k);
        a += c * (u32)t[k];
        b += c * (u32)(t[k] >> 32);
    }
    const u64 dg = j == 0 ? PoseidonConsts::diag[0] : 0;  // diag[0] = 8, 0 on every other row
    a += dg * (u32)t[0];
    b += dg * (u32)(t[0] >> 32);
    const u64 v = reduce_lohi42(a, b);
    return SAFE ? gl_canon(v) : v;
}

// x of lane `src` (a lane of the same wave), through the LDS crossbar.
ZK_DEV u64 hex_fetch(u64 x, u32 src)
{
#ifdef __HIP_DEVICE_COMPILE__
    const int addr = (int)(src << 2);
    const u32 lo = (u32)__builtin_amdgcn_ds_bpermute(addr, (int)(u32)x);
    const u32 hi = (u32)__builtin_amdgcn_ds_bpermute(addr, (int)(u32)(x >> 32));
    return ((u64)hi << 32) | lo;
#else
    (void)src;
    return x;  // device-only function; the host emulation gathers itself
#endif
}

// The whole permutation for lane j of a 16-lane group (all 16 lanes must run it
// together; `base` is the wave lane of the group's lane 0).  Relaxed in and
// out unless SAFE (canonical in and out).
template <bool SAFE>
ZK_DEV void poseidon_permute_hex(u64 &s, u32 j, u32 base, bool &rare)
{
    const u64 *rc = round_constants();
    const u32 e = j < 12 ? j : j - 12;  // the element this lane holds
    s = SAFE ? gl_add(s, rc[e]) : gl_add_nc(s, rc[e], rare);
    u32 src[12];
#pragma unroll
    for (int k = 1; k < 12; k++) src[k] = base + (e + (u32)k) % 12;
    // this lane's constant is a vector load (lane-dependent index): fetch it
    // one round ahead so its latency hides behind the S-box
    u64 rk = rc[12 + e];
#pragma unroll 1
    for (int r = 0; r < 30; r++) {
        const int rn = r + 2 < 31 ? r + 2 : 30;  // row 30 is the all-zero tail
        const u64 rk_next = rc[12 * rn + e];
        s = hex_sbox<SAFE>(s, e, r < 4 || r >= 26, rare);
        u64 t[12];
        t[0] = s;
#pragma unroll
        for (int k = 1; k < 12; k++) t[k] = hex_fetch(s, src[k]);
        s = hex_mds<SAFE>(t, e, rk);
        rk = rk_next;
    }
}

// H(a, b) by a 16-lane group: every lane calls it (lane j reads the element
// it holds); lanes 0..3 return digest element j.  All lanes of the wave must
// call it together (the rare-case recomputation is wave-uniform).
ZK_DEV u64 hash_two_hex(const u64 *a, const u64 *b, u32 j, u32 base)
{
    const u32 e = j < 12 ? j : j - 12;
    const u64 in = e < 4 ? a[e] : (e < 8 ? b[e - 4] : 0);
    u64 s = in;
    bool rare = false;
    poseidon_permute_hex<false>(s, j, base, rare);
    if (__any(rare)) {
        s = gl_canon(in);
        bool unused = false;
        poseidon_permute_hex<true>(s, j, base, unused);
    }
    return gl_canon(s);
}

}  // namespace zk
