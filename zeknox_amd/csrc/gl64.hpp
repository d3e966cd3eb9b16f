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
// Used inside the Poseidon permutation, canonicalised once at the digest.

// lo + 2^64 (h0 + 2^32 h1) == lo + h0 (2^32 - 1) - h1; result in [0, 2^64).
// lo + h0 (2^32-1) wraps at most once and the wrap is worth +EPS, which cannot
// wrap again (the wrapped value is < h0 (2^32-1) <= 2^64 - 2^33 + 1); the
// borrow of "- h1" (only when the value is < 2^32) is worth -EPS.
ZK_HD u64 gl_reduce128_nc(u64 lo, u64 hi)
{
    const u32 h0 = (u32)hi, h1 = (u32)(hi >> 32);
#ifdef __HIP_DEVICE_COMPILE__
    // gfx950: the carry-out of v_mad_u64_u32 and the borrow of the 64-bit
    // subtraction feed the corrections directly (no 64-bit compares).
    u64 s, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(s), "=s"(c) : "v"(h0), "v"(lo));
    u32 e;
    asm("v_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(e) : "s"(c));
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(s), "=s"(c) : "v"(e), "v"(s));
    u32 tl, th;
    u64 bmask;
    asm("v_sub_co_u32 %0, %2, %3, %4\n\tv_subbrev_co_u32 %1, %2, 0, %5, %2"
        : "=&v"(tl), "=v"(th), "=&s"(bmask)
        : "v"((u32)s), "v"(h1), "v"((u32)(s >> 32)));
    u64 t = ((u64)th << 32) | tl;
    if (__builtin_expect(bmask != 0, 0)) {  // some lane borrowed (value < 2^32): -2^64 == -EPS
        if ((bmask >> __lane_id()) & 1) t -= ZK_EPS;
    }
    return t;
#else
    u64 s = lo + (u64)h0 * ZK_EPS;
    if (s < lo) s += ZK_EPS;
    u64 t = s - h1;
    if (s < h1) t -= ZK_EPS;
    return t;
#endif
}

// x = lo + 2^64 h for a small h (< 2^32): lo + h (2^32 - 1), result any u64.
ZK_HD u64 gl_reduce_small_nc(u64 lo, u32 h)
{
#ifdef __HIP_DEVICE_COMPILE__
    u64 s, c;
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(s), "=s"(c) : "v"(h), "v"(lo));
    u32 e;
    asm("v_cndmask_b32_e64 %0, 0, 1, %1" : "=v"(e) : "s"(c));
    asm("v_mad_u64_u32 %0, %1, %2, -1, %3" : "=v"(s), "=s"(c) : "v"(e), "v"(s));
    return s;
#else
    u64 s = lo + (u64)h * ZK_EPS;
    if (s < lo) s += ZK_EPS;
    return s;
#endif
}

ZK_HD u64 gl_mul_nc(u64 a, u64 b)
{
    u64 lo, hi;
    gl_mul_wide(a, b, lo, hi);
    return gl_reduce128_nc(lo, hi);
}

// a + b for any u64 a, b: a wrap is worth +EPS; a second wrap (both inputs
// near 2^64) is worth another +EPS and cannot wrap again.
ZK_HD u64 gl_add_nc(u64 a, u64 b)
{
    u64 s = a + b;
    if (s < a) {
        s += ZK_EPS;
        if (s < ZK_EPS) s += ZK_EPS;
    }
    return s;
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

// root of unity of order 2^lg: 7^((p-1)/2^lg)  (7 generates the multiplicative group)
ZK_HD u64 gl_omega(u32 lg) { return gl_pow(7, (ZK_P - 1) >> lg); }
