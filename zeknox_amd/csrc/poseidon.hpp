// Poseidon permutation over Goldilocks, width 12 (rate 8, capacity 4),
// 8 full rounds + 22 partial rounds, S-box x^7, and the Plonky2 sponge
// ("hash_or_noop" and two-to-one compression) used by okx/zeknox's
// PoseidonHasher (native/poseidon/poseidon.cu:771-780,834-838,
// native/poseidon/poseidon_permutation.cuh:83-134).
//
// One permutation per lane: the 12-element state lives in 24 VGPRs; MDS
// coefficients are immediates, round constants come from constant memory.
#pragma once
#include "gl64.hpp"
#include "poseidon_gl_constants.h"

namespace zk {

struct PoseidonConsts {
    static constexpr u64 circ[12] = ZK_POSEIDON_MDS_CIRC;
    static constexpr u64 diag[12] = ZK_POSEIDON_MDS_DIAG;
};

// Round constants: scalar (uniform) loads on the device, a plain table on the host.
static __constant__ const u64 kRoundConstDev[360] = ZK_POSEIDON_RC_TABLE;
static const u64 kRoundConstHost[360] = ZK_POSEIDON_RC_TABLE;

ZK_HD const u64 *round_constants()
{
#ifdef __HIP_DEVICE_COMPILE__
    return kRoundConstDev;
#else
    return kRoundConstHost;
#endif
}

ZK_HD u64 sbox7(u64 x)
{
    u64 x2 = gl_sqr(x);
    u64 x4 = gl_sqr(x2);
    u64 x3 = gl_mul(x, x2);
    return gl_mul(x3, x4);
}

// state <- M state, M circulant (first row circ) plus diag.  Each row is a sum
// of 13 products of a u64 with a constant below 2^6; the low and high 32-bit
// halves are accumulated separately (each sum < 2^41) and the 105-bit total
// is reduced once.
ZK_HD void mds_layer(u64 s[12])
{
    u32 lo[12], hi[12];
#pragma unroll
    for (int j = 0; j < 12; j++) {
        lo[j] = (u32)s[j];
        hi[j] = (u32)(s[j] >> 32);
    }
#pragma unroll
    for (int r = 0; r < 12; r++) {
        u64 a = 0, b = 0;
#pragma unroll
        for (int i = 0; i < 12; i++) {
            const u64 c = PoseidonConsts::circ[i] + (i == 0 ? PoseidonConsts::diag[r] : 0);
            a += c * lo[(i + r) % 12];
            b += c * hi[(i + r) % 12];
        }
        u64 l = a + (b << 32);
        u64 h = (b >> 32) + (l < a ? 1 : 0);
        s[r] = gl_reduce128(l, h);
    }
}

ZK_HD void full_round(u64 s[12], const u64 *rc)
{
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = sbox7(gl_add(s[i], rc[i]));
    mds_layer(s);
}

ZK_HD void partial_round(u64 s[12], const u64 *rc)
{
#pragma unroll
    for (int i = 1; i < 12; i++) s[i] = gl_add(s[i], rc[i]);
    s[0] = sbox7(gl_add(s[0], rc[0]));
    mds_layer(s);
}

// 4 full rounds, 22 partial rounds (S-box on lane 0), 4 full rounds.  The
// round loops are not unrolled: one round body is ~1-2k instructions, so the
// whole permutation stays resident in the instruction cache.
ZK_HD void poseidon_permute(u64 s[12])
{
    const u64 *rc = round_constants();
#pragma unroll 1
    for (int r = 0; r < 4; r++) full_round(s, rc + 12 * r);
#pragma unroll 1
    for (int r = 4; r < 26; r++) partial_round(s, rc + 12 * r);
#pragma unroll 1
    for (int r = 26; r < 30; r++) full_round(s, rc + 12 * r);
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
    poseidon_permute(s);
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = s[i];
}

// hash_or_noop over n elements fetched by `get(i)`: n <= 4 copies the input,
// zero-padded; otherwise the overwrite-mode sponge (a final short chunk leaves
// the stale rate lanes in place), digest = state[0..4).
template <typename Get>
ZK_HD void hash_one(Get get, u32 n, u64 out[4])
{
    if (n <= 4) {
#pragma unroll
        for (u32 i = 0; i < 4; i++) out[i] = i < n ? get(i) : 0;
        return;
    }
    u64 s[12];
#pragma unroll
    for (int i = 0; i < 12; i++) s[i] = 0;
    for (u32 idx = 0; idx < n; idx += 8) {
        const u32 rem = n - idx;
#pragma unroll
        for (u32 i = 0; i < 8; i++)
            if (i < rem) s[i] = get(idx + i);
        poseidon_permute(s);
    }
#pragma unroll
    for (int i = 0; i < 4; i++) out[i] = s[i];
}

}  // namespace zk
