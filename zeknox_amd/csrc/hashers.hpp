// The Merkle hashers other than Poseidon, for the HashType argument of the
// Merkle entry points (okx/zeknox native/merkle/merkle.h:21-28): Keccak (1)
// and Poseidon2 (3).  Host- and device-callable, one hash per lane.
//
// Each hasher has one(get, n, out) for a leaf of n elements fetched by
// get(i) and two(a, b, out) for an internal node; digests are 4 u64.
#pragma once
#include "poseidon.hpp"
#include "poseidon2_gl_constants.h"

namespace zk {

// ------------------------------------------------------------------ Keccak
// Keccak-f[1600] (FIPS 202 section 3.2) on 25 lanes A[x + 5y].  The round
// constants come from the degree-8 LFSR rc(t) and the rotation offsets from
// the (x, y) -> (y, 2x + 3y) walk, both evaluated at compile time.
struct KeccakTables {
    u64 rc[24];
    int rot[25];
};

__host__ __device__ constexpr int keccak_rc_bit(int t)
{
    if (t % 255 == 0) return 1;
    unsigned r = 1;
    for (int i = 1; i <= t % 255; i++) {
        r <<= 1;
        if (r & 0x100) r ^= 0x171;  // x^8 + x^6 + x^5 + x^4 + 1
    }
    return (int)(r & 1);
}

__host__ __device__ constexpr KeccakTables make_keccak_tables()
{
    KeccakTables t{};
    for (int ir = 0; ir < 24; ir++) {
        u64 c = 0;
        for (int j = 0; j <= 6; j++)
            if (keccak_rc_bit(j + 7 * ir)) c |= 1ull << ((1 << j) - 1);
        t.rc[ir] = c;
    }
    t.rot[0] = 0;
    int x = 1, y = 0;
    for (int s = 0; s < 24; s++) {
        t.rot[x + 5 * y] = ((s + 1) * (s + 2) / 2) % 64;
        const int nx = y, ny = (2 * x + 3 * y) % 5;
        x = nx;
        y = ny;
    }
    return t;
}

ZK_HD u64 rotl64(u64 v, int n) { return n ? (v << n) | (v >> (64 - n)) : v; }

ZK_HD void keccak_f1600(u64 A[25])
{
    constexpr KeccakTables T = make_keccak_tables();
    static_for<0, 24>([&](auto rc) {
        constexpr int ir = decltype(rc)::value;
        u64 C[5], B[25];
        static_for<0, 5>([&](auto xc) {
            constexpr int x = decltype(xc)::value;
            C[x] = A[x] ^ A[x + 5] ^ A[x + 10] ^ A[x + 15] ^ A[x + 20];
        });
        static_for<0, 5>([&](auto xc) {
            constexpr int x = decltype(xc)::value;
            const u64 D = C[(x + 4) % 5] ^ rotl64(C[(x + 1) % 5], 1);
            static_for<0, 5>([&](auto yc) { A[x + 5 * decltype(yc)::value] ^= D; });
        });
        // rho and pi: B[y, 2x + 3y] = rot(A[x, y])
        static_for<0, 25>([&](auto ic) {
            constexpr int i = decltype(ic)::value, x = i % 5, y = i / 5;
            B[y + 5 * ((2 * x + 3 * y) % 5)] = rotl64(A[i], T.rot[i]);
        });
        static_for<0, 25>([&](auto ic) {
            constexpr int i = decltype(ic)::value, x = i % 5, y = i / 5;
            A[i] = B[i] ^ (~B[(x + 1) % 5 + 5 * y] & B[(x + 2) % 5 + 5 * y]);
        });
        A[0] ^= T.rc[ir];
    });
}

// Plonky2's Keccak hasher as zeknox defines it (native/keccak/keccak.cu:
// 139-158): fewer than 4 elements are copied (zero padded); otherwise
// Keccak-256 (original 0x01 padding, rate 136 bytes = 17 lanes) of the
// elements' little-endian bytes, keeping 25 bytes.  A node hashes the first
// 25 bytes of each child.
struct KeccakHasher {
    template <typename Get>
    ZK_HD static void one(Get get, u32 n, u64 out[4])
    {
        if (n < 4) {
#pragma unroll
            for (u32 i = 0; i < 4; i++) out[i] = i < n ? get(i) : 0;
            return;
        }
        u64 A[25];
#pragma unroll
        for (int i = 0; i < 25; i++) A[i] = 0;
        u32 idx = 0;
        for (; idx + 17 <= n; idx += 17) {
#pragma unroll
            for (u32 w = 0; w < 17; w++) A[w] ^= get(idx + w);
            keccak_f1600(A);
        }
        const u32 m = n - idx;  // 0 .. 16 words left, then the padding
#pragma unroll
        for (u32 w = 0; w < 17; w++) {
            if (w < m) A[w] ^= get(idx + w);
            if (w == m) A[w] ^= 0x01;
        }
        A[16] ^= 0x8000000000000000ull;
        keccak_f1600(A);
        out[0] = A[0];
        out[1] = A[1];
        out[2] = A[2];
        out[3] = A[3] & 0xFF;
    }

    ZK_HD static void two(const u64 a[4], const u64 b[4], u64 out[4])
    {
        u64 A[25];
#pragma unroll
        for (int i = 0; i < 25; i++) A[i] = 0;
        // message bytes: a[0..25) then b[0..25), then the 0x01 pad at byte 50
        A[0] = a[0];
        A[1] = a[1];
        A[2] = a[2];
        A[3] = (a[3] & 0xFF) | (b[0] << 8);
        A[4] = (b[0] >> 56) | (b[1] << 8);
        A[5] = (b[1] >> 56) | (b[2] << 8);
        A[6] = (b[2] >> 56) | ((b[3] & 0xFF) << 8) | (0x01ull << 16);
        A[16] = 0x8000000000000000ull;
        keccak_f1600(A);
        out[0] = A[0];
        out[1] = A[1];
        out[2] = A[2];
        out[3] = A[3] & 0xFF;
    }
};

// --------------------------------------------------------------- Poseidon2
// Poseidon2 over Goldilocks, width 12 (native/poseidon2/poseidon2.cu:
// 51-201): an external linear layer, 4 external rounds (constants and S-box
// on every element, external layer), 22 internal rounds (constant and S-box
// on element 0, internal layer), 4 external rounds.  External layer: the
// 4x4 block [[5,7,1,3],[4,6,1,1],[1,3,5,7],[1,1,4,6]] on each quarter, then
// every element adds the sum of its column over the three quarters.
// Internal layer: s_i <- s_i * DIAG[i] + sum(s).  Canonical arithmetic.
static __constant__ const u64 kP2RoundConstDev[360] = ZK_POSEIDON2_RC_TABLE;
static const u64 kP2RoundConstHost[360] = ZK_POSEIDON2_RC_TABLE;
static __constant__ const u64 kP2DiagDev[12] = ZK_POSEIDON2_DIAG;
static const u64 kP2DiagHost[12] = ZK_POSEIDON2_DIAG;

ZK_HD const u64 *p2_round_constants()
{
#ifdef __HIP_DEVICE_COMPILE__
    return kP2RoundConstDev;
#else
    return kP2RoundConstHost;
#endif
}

ZK_HD const u64 *p2_diag()
{
#ifdef __HIP_DEVICE_COMPILE__
    return kP2DiagDev;
#else
    return kP2DiagHost;
#endif
}

ZK_HD void p2_m4(u64 *x)
{
    const u64 t0 = gl_add(x[0], x[1]), t1 = gl_add(x[2], x[3]);
    const u64 t2 = gl_add(gl_add(x[1], x[1]), t1), t3 = gl_add(gl_add(x[3], x[3]), t0);
    const u64 t1x2 = gl_add(t1, t1), t0x2 = gl_add(t0, t0);
    const u64 t4 = gl_add(gl_add(t1x2, t1x2), t3), t5 = gl_add(gl_add(t0x2, t0x2), t2);
    x[0] = gl_add(t3, t5);
    x[1] = t5;
    x[2] = gl_add(t2, t4);
    x[3] = t4;
}

ZK_HD void p2_external(u64 s[12])
{
    p2_m4(s);
    p2_m4(s + 4);
    p2_m4(s + 8);
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const u64 col = gl_add(gl_add(s[k], s[4 + k]), s[8 + k]);
        s[k] = gl_add(s[k], col);
        s[4 + k] = gl_add(s[4 + k], col);
        s[8 + k] = gl_add(s[8 + k], col);
    }
}

ZK_HD u64 p2_sbox(u64 x)
{
    const u64 x2 = gl_mul(x, x), x4 = gl_mul(x2, x2);
    return gl_mul(gl_mul(x, x2), x4);
}

ZK_HD void poseidon2_permute(u64 s[12])
{
    const u64 *rc = p2_round_constants();
    const u64 *dg = p2_diag();
    p2_external(s);
#pragma unroll 1
    for (int r = 0; r < 30; r++) {
        if (r < 4 || r >= 26) {
#pragma unroll
            for (int i = 0; i < 12; i++) s[i] = p2_sbox(gl_add(s[i], rc[12 * r + i]));
            p2_external(s);
        } else {
            s[0] = p2_sbox(gl_add(s[0], rc[12 * r]));
            u64 sum = s[0];
#pragma unroll
            for (int i = 1; i < 12; i++) sum = gl_add(sum, s[i]);
#pragma unroll
            for (int i = 0; i < 12; i++) s[i] = gl_add(gl_mul(s[i], dg[i]), sum);
        }
    }
}

// The Plonky2 sponge of the Poseidon hasher (hash_or_noop, overwrite mode,
// two-to-one over [a, b, 0, 0, 0, 0]) around Poseidon2
// (native/poseidon2/poseidon2.cu:215-243).
struct Poseidon2Hasher {
    template <typename Get>
    ZK_HD static void one(Get get, u32 n, u64 out[4])
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
            poseidon2_permute(s);
        }
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = s[i];
    }

    ZK_HD static void two(const u64 a[4], const u64 b[4], u64 out[4])
    {
        u64 s[12];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            s[i] = a[i];
            s[4 + i] = b[i];
            s[8 + i] = 0;
        }
        poseidon2_permute(s);
#pragma unroll
        for (int i = 0; i < 4; i++) out[i] = s[i];
    }
};

// Poseidon under the same interface (the hashing of the optimised kernels).
struct PoseidonHasher {
    template <typename Get>
    ZK_HD static void one(Get get, u32 n, u64 out[4])
    {
        hash_one(get, n, out);
    }
    ZK_HD static void two(const u64 a[4], const u64 b[4], u64 out[4]) { hash_two(a, b, out); }
};

}  // namespace zk
