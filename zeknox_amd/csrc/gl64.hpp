// Goldilocks field arithmetic, p = 2^64 - 2^32 + 1, for CDNA4 (gfx950) wave64.
//
// Values are u64.  Every function takes canonical inputs (< p) and returns a
// canonical result, matching the reference device type's behaviour
// (okx/zeknox native/ff/gl64_t.cuh:93-121 add, :175-198 sub, :312-369 mul,
// :485-506 canonicalisation) so outputs are bit-identical.  The reduction is
// derived from 2^64 == 2^32 - 1 (mod p) and 2^96 == -1 (mod p); the 64x64
// product is built from four 32x32->64 partial products, which the compiler
// maps onto v_mad_u64_u32.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#define ZK_P 0xFFFFFFFF00000001ull
#define ZK_EPS 0x00000000FFFFFFFFull  // 2^64 mod p

#define ZK_DEV __device__ __forceinline__
#define ZK_HD __host__ __device__ __forceinline__

typedef uint64_t u64;
typedef uint32_t u32;

ZK_HD u64 gl_add(u64 a, u64 b)
{
    u64 s = a + b;
    u64 r = s + ZK_EPS;  // s - p (mod 2^64)
    return (s < a || s >= ZK_P) ? r : s;
}

ZK_HD u64 gl_sub(u64 a, u64 b)
{
    u64 d = a - b;
    return a < b ? d - ZK_EPS : d;  // d + p (mod 2^64)
}

ZK_HD u64 gl_neg(u64 a) { return a ? ZK_P - a : 0; }

// x = lo + 2^64 * (h0 + 2^32 * h1) == lo + h0 * (2^32 - 1) - h1  (mod p)
ZK_HD u64 gl_reduce128(u64 lo, u64 hi)
{
    u32 h0 = (u32)hi, h1 = (u32)(hi >> 32);
    u64 t0 = lo - h1;
    if (lo < h1) t0 -= ZK_EPS;            // borrow: + p == - (2^32 - 1) mod 2^64
    u64 t1 = ((u64)h0 << 32) - h0;         // h0 * (2^32 - 1)
    u64 r = t0 + t1;
    if (r < t1) r += ZK_EPS;               // carry: + 2^64 == + (2^32 - 1)
    return r >= ZK_P ? r - ZK_P : r;
}

// 64x64 -> 128 product from 32-bit partial products; no intermediate overflows:
// a_i*b_j + c <= (2^32-1)^2 + 2(2^32-1) = 2^64 - 1.
ZK_HD void gl_mul_wide(u64 a, u64 b, u64 &lo, u64 &hi)
{
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 p01 = (u64)a0 * b1 + (p00 >> 32);
    u64 p10 = (u64)a1 * b0 + (u32)p01;
    u64 p11 = (u64)a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (u32)p00;
    hi = p11;
}

ZK_HD u64 gl_mul(u64 a, u64 b)
{
    u64 lo, hi;
    gl_mul_wide(a, b, lo, hi);
    return gl_reduce128(lo, hi);
}

ZK_HD u64 gl_sqr(u64 a) { return gl_mul(a, a); }

// ---- relaxed forms: inputs and results are any u64 congruent mod p --------
// Used inside the Poseidon permutation and the NTT passes; canonicalised once
// at the output.  Each operation has a rare case (probability ~2^-32 for
// random data, reachable only by values near 2^64 or below 2^32) in which one
// more +-(2^32 - 1) correction would be needed.  The operations do not branch
// on it: they OR it into `rare` (a lane mask on the device), and the caller
// recomputes the whole block with the canonical operations when any lane of
// the wave has it set.  Carries come from the carry builtins (32-bit add and
// subtract chains), so the compiler sees every carry mask and inserts the wait
// states gfx950 requires between a VALU write of a mask and its use.

// x = lo + 2^64 h for a small h (< 2^32): lo + h (2^32 - 1) (no rare case:
// the sum wraps at most once, and the wrapped value plus 2^32 - 1 cannot).
ZK_HD u64 gl_reduce_small_nc(u64 lo, u32 h)
{
    u64 s;
    const bool c = __builtin_add_overflow(lo, (u64)h * ZK_EPS, &s);
    return s + (c ? ZK_EPS : 0);
}

// The same sum as one v_mad_u64_u32 (h (2^32 - 1) + lo) whose carry-out (bit
// 64) selects the correction, where the compiler adds the product separately
// and rebuilds the carry with a 64-bit compare (two instructions fewer).  A
// VALU write of an SGPR needs two wait states before a VALU reads it
// (s_nop 1).  Faster in the Poseidon permutation, whose dependency chains are
// long; the NTT passes, which interleave many short chains, ran slower with it
// (the asm block is a unit the scheduler cannot split), so they keep the form
// above.
ZK_HD u64 gl_reduce_small_mad_nc(u64 lo, u32 h)
{
#ifdef __HIP_DEVICE_COMPILE__
    u64 t, k;
    u32 e;
    asm volatile("v_mad_u64_u32 %0, %2, %3, -1, %4\n\ts_nop 1\n\tv_cndmask_b32 %1, 0, -1, %2"
                 : "=&v"(t), "=v"(e), "=&s"(k)
                 : "v"(h), "v"(lo));
    return t + e;
#else
    return gl_reduce_small_nc(lo, h);
#endif
}

// A zero the compiler cannot see through (a scalar register on the device):
// as the second operand of a subtract-with-borrow it keeps the instruction's
// own borrow-out usable, where a literal 0 makes the compiler rebuild the
// borrow with a compare.
ZK_HD u32 opaque_zero()
{
#ifdef __HIP_DEVICE_COMPILE__
    u32 r;
    asm("s_mov_b32 %0, 0" : "=s"(r));
    return r;
#else
    return 0;
#endif
}

// lo + 2^64 (h0 + 2^32 h1) == lo + h0 (2^32 - 1) - h1.  The borrow of "- h1"
// (value < 2^32) is the rare case, worth -EPS.
template <bool MAD = false>
ZK_HD u64 gl_reduce128_nc(u64 lo, u64 hi, bool &rare)
{
    const u64 s = MAD ? gl_reduce_small_mad_nc(lo, (u32)hi) : gl_reduce_small_nc(lo, (u32)hi);
    unsigned c0, c1;
    const u32 d0 = __builtin_subc((u32)s, (u32)(hi >> 32), 0u, &c0);
    const u32 d1 = __builtin_subc((u32)(s >> 32), opaque_zero(), c0, &c1);
    rare |= c1 != 0;
    return ((u64)d1 << 32) | d0;
}

// MAD: reduce with gl_reduce_small_mad_nc (see there).
template <bool MAD = false>
ZK_HD u64 gl_mul_nc(u64 a, u64 b, bool &rare)
{
    u64 lo, hi;
    gl_mul_wide(a, b, lo, hi);
    return gl_reduce128_nc<MAD>(lo, hi, rare);
}

// a^2 = a0^2 + 2^33 a0 a1 + 2^64 a1^2: three 32x32 products instead of four,
// and the cross term enters through a 32-bit carry chain, so no partial
// product has to be moved into a register pair to serve as an addend.
ZK_HD void gl_sqr_wide(u64 a, u64 &lo, u64 &hi)
{
    const u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    const u64 p00 = (u64)a0 * a0, m = (u64)a0 * a1, p11 = (u64)a1 * a1;
    const u64 m2 = m << 1;  // 2 a0 a1 mod 2^64; bit 64 is m >> 63
    unsigned k1, k2, k3;
    const u32 l1 = __builtin_addc((u32)(p00 >> 32), (u32)m2, 0u, &k1);
    const u32 h0 = __builtin_addc((u32)p11, (u32)(m2 >> 32), k1, &k2);
    const u32 h1 = __builtin_addc((u32)(p11 >> 32), (u32)(m >> 63), k2, &k3);
    lo = ((u64)l1 << 32) | (u32)p00;
    hi = ((u64)h1 << 32) | h0;
}

ZK_HD u64 gl_sqr_nc(u64 a, bool &rare)
{
    u64 lo, hi;
    gl_sqr_wide(a, lo, hi);
    return gl_reduce128_nc(lo, hi, rare);
}

// a + b: a wrap is worth +EPS; a second wrap (both inputs above 2^64 - 2^32)
// is the rare case.  Written as 32-bit carry chains so that every carry is
// the carry-out of an add (v_add_co/v_addc_co) rather than a 64-bit compare.
ZK_HD u64 gl_add_nc(u64 a, u64 b, bool &rare)
{
    unsigned c0, c1, c2, c3;
    const u32 s0 = __builtin_addc((u32)a, (u32)b, 0u, &c0);
    const u32 s1 = __builtin_addc((u32)(a >> 32), (u32)(b >> 32), c0, &c1);
    const u32 t0 = __builtin_addc(s0, c1 ? 0xFFFFFFFFu : 0u, 0u, &c2);
    const u32 t1 = __builtin_addc(s1, 0u, c2, &c3);
    rare |= c3 != 0;
    return ((u64)t1 << 32) | t0;
}

// a - b: a borrow is worth -EPS (i.e. + p); a second borrow (b > p + a) is
// the rare case.
ZK_HD u64 gl_sub_nc(u64 a, u64 b, bool &rare)
{
    unsigned c0, c1, c2, c3;
    const u32 d0 = __builtin_subc((u32)a, (u32)b, 0u, &c0);
    const u32 d1 = __builtin_subc((u32)(a >> 32), (u32)(b >> 32), c0, &c1);
    const u32 t0 = __builtin_subc(d0, c1 ? 0xFFFFFFFFu : 0u, 0u, &c2);
    const u32 t1 = __builtin_subc(d1, opaque_zero(), c2, &c3);
    rare |= c3 != 0;
    return ((u64)t1 << 32) | t0;
}

// 2^e mod p for 0 <= e < 96 (2^(e-32) - 2^(e-64) when e >= 64).
__host__ __device__ constexpr u64 pow2_mod_p(int e)
{
    return e < 64 ? (1ull << e) : ((1ull << (e - 32)) - (1ull << (e - 64)));
}

// x * 2^e mod p (relaxed) for a constant e in [0, 96): 2^96 == -1, so callers
// fold e >= 96 into a sign.  e < 64: the 128-bit shift is reduced directly;
// 64 <= e < 96: multiply by the constant 2^e mod p.
ZK_HD u64 gl_mul_pow2_nc(u64 x, int e, bool &rare)
{
    if (e == 0) return x;
    if (e < 64) {
        const u64 lo = x << e, hi = x >> (64 - e);
        if (e <= 32) return gl_reduce_small_nc(lo, (u32)hi);
        return gl_reduce128_nc(lo, hi, rare);
    }
    return gl_mul_nc(x, pow2_mod_p(e), rare);
}

// x * 2^-k mod p (relaxed) for a constant 0 < k <= 32.  With x = q 2^k + r:
// x 2^-k == q + r 2^-k and 2^-k == -2^(96-k), r 2^(96-k) == s (2^32 - 1) for
// s = r 2^(32-k) < 2^32, so x 2^-k == q + s - s 2^32.  q + s cannot wrap
// (q < 2^63); subtracting s 2^32 wraps at most once, by 2^64 == 2^32 - 1,
// and the wrapped value is at least 2^32, so removing 2^32 - 1 cannot wrap
// again: no rare case.
ZK_HD u64 gl_shr_nc(u64 x, int k)
{
    const u64 q = x >> k;
    const u32 s = k == 32 ? (u32)x : ((u32)x << (32 - k));
    unsigned c0, c1, b0, b1, b2;
    const u32 t0 = __builtin_addc((u32)q, s, 0u, &c0);
    const u32 t1 = __builtin_addc((u32)(q >> 32), 0u, c0, &c1);
    const u32 d1 = __builtin_subc(t1, s, 0u, &b0);
    const u32 e0 = __builtin_subc(t0, b0 ? 0xFFFFFFFFu : 0u, 0u, &b1);
    const u32 e1 = __builtin_subc(d1, opaque_zero(), b1, &b2);
    return ((u64)e1 << 32) | e0;
}

// p - x (relaxed): x above p is the rare case.
ZK_HD u64 gl_neg_nc(u64 x, bool &rare)
{
    unsigned b0, b1;
    const u32 d0 = __builtin_subc(1u, (u32)x, 0u, &b0);
    const u32 d1 = __builtin_subc(0xFFFFFFFFu, (u32)(x >> 32), b0, &b1);
    rare |= b1 != 0;
    return ((u64)d1 << 32) | d0;
}

ZK_HD u64 gl_pow(u64 a, u64 e)
{
    u64 r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        a = gl_mul(a, a);
        e >>= 1;
    }
    return r;
}

ZK_HD u64 gl_inv(u64 a) { return gl_pow(a, ZK_P - 2); }

// Canonical form of an arbitrary u64 (values in [p, 2^64) map to x - p).
ZK_HD u64 gl_canon(u64 a) { return a >= ZK_P ? a - ZK_P : a; }

ZK_HD u32 zk_bitrev(u32 x, u32 bits)
{
#ifdef __HIP_DEVICE_COMPILE__
    return bits ? (__builtin_bitreverse32(x) >> (32 - bits)) : 0;
#else
    u32 r = 0;
    for (u32 i = 0; i < bits; i++) { r = (r << 1) | (x & 1); x >>= 1; }
    return r;
#endif
}

// Compile-time loop: f(std::integral_constant<int, I>) for I in [BEGIN, END),
// so register-array indices are constants even in large unrolled bodies.
template <int I, int END, typename F>
ZK_HD void static_for(F &&f)
{
    if constexpr (I < END) {
        f(std::integral_constant<int, I>{});
        static_for<I + 1, END>(f);
    }
}

// Bit reversal usable in constant expressions (register-array indices).
__host__ __device__ constexpr u32 ct_bitrev(u32 x, int bits)
{
    u32 r = 0;
    for (int i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

// root of unity of order 2^lg: 7^((p-1)/2^lg)  (7 generates the multiplicative group)
ZK_HD u64 gl_omega(u32 lg) { return gl_pow(7, (ZK_P - 1) >> lg); }
