// Issue cost of gfx950 VALU instructions: SIMD-cycles per wave64 instruction
// when one instruction kind runs in 8 independent chains per lane over 32
// waves per SIMD (256 CUs x 4 SIMDs at 2.4 GHz).  The leaf-hash roofline of
// bench.py weights the permutation's instruction mix with these costs.
//
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o valu_costs tests/native/valu_costs.hip && ./valu_costs
#include <hip/hip_runtime.h>
#include <stdio.h>
#include <stdint.h>

#define ITERS 4096
typedef uint64_t u64; typedef uint32_t u32;

#define DEF(NAME, BODY) \
__global__ __launch_bounds__(256) void NAME(u32 *out, u32 seed) { \
  u32 a0 = threadIdx.x ^ seed, a1 = a0 * 3, a2 = a0 * 5, a3 = a0 * 7, a4 = a0 * 11, a5 = a0 * 13, a6 = a0 * 17, a7 = a0 * 19; \
  u32 b0 = a0 + 1, b1 = a1 + 1, b2 = a2 + 1, b3 = a3 + 1, b4 = a4 + 1, b5 = a5 + 1, b6 = a6 + 1, b7 = a7 + 1; \
  u32 k = seed | 1; \
  for (int i = 0; i < ITERS; i++) { BODY } \
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7 ^ b0 ^ b1 ^ b2 ^ b3 ^ b4 ^ b5 ^ b6 ^ b7; }

// 8 instructions per iteration in each body
#define X8(S) S(a0,b0) S(a1,b1) S(a2,b2) S(a3,b3) S(a4,b4) S(a5,b5) S(a6,b6) S(a7,b7)
#define ADD32(a,b) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define XOR32(a,b) asm volatile("v_xor_b32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MOV32(a,b) asm volatile("v_mov_b32 %0, %1" : "=v"(a) : "v"(b));
#define ADDCO(a,b) asm volatile("v_add_co_u32 %0, vcc, %0, %1" : "+v"(a) : "v"(b) : "vcc");
#define ADD3(a,b) asm volatile("v_add3_u32 %0, %0, %1, %1" : "+v"(a) : "v"(b));
#define ALIGNB(a,b) asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(a) : "v"(b));
#define PERM(a,b) asm volatile("v_perm_b32 %0, %0, %1, %2" : "+v"(a) : "v"(b), "v"(k));
#define MULLO(a,b) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MULHI(a,b) asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MAD24(a,b) asm volatile("v_mad_u32_u24 %0, %0, %1, %0" : "+v"(a) : "v"(b));
#define DOT2(a,b) asm volatile("v_dot2_u32_u16 %0, %0, %1, %0" : "+v"(a) : "v"(b));
#define DOT4(a,b) asm volatile("v_dot4_u32_u8 %0, %0, %1, %0" : "+v"(a) : "v"(b));
#define PKADD(a,b) asm volatile("v_pk_add_u16 %0, %0, %1" : "+v"(a) : "v"(b));
#define PKMAD(a,b) asm volatile("v_pk_mad_u16 %0, %0, %1, %0" : "+v"(a) : "v"(b));
#define LSHLADD32(a,b) asm volatile("v_lshl_add_u32 %0, %0, 3, %1" : "+v"(a) : "v"(b));

#define SUB32(a,b) asm volatile("v_sub_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MUL24(a,b) asm volatile("v_mul_u32_u24 %0, %0, %1" : "+v"(a) : "v"(b));
#define MULHI24(a,b) asm volatile("v_mul_hi_u32_u24 %0, %0, %1" : "+v"(a) : "v"(b));
#define LSHL32(a,b) asm volatile("v_lshlrev_b32 %0, %1, %0" : "+v"(a) : "v"(b));
#define LSHR32(a,b) asm volatile("v_lshrrev_b32 %0, %1, %0" : "+v"(a) : "v"(b));
#define AND32(a,b) asm volatile("v_and_b32 %0, %0, %1" : "+v"(a) : "v"(b));
#define OR32(a,b) asm volatile("v_or_b32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MAX32(a,b) asm volatile("v_max_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define ADDCOE64(a,b) asm volatile("v_add_co_u32_e64 %0, s[40:41], %0, %1" : "+v"(a) : "v"(b) : "s40", "s41");
#define ADD32E64(a,b) asm volatile("v_add_u32_e64 %0, %0, %1" : "+v"(a) : "v"(b));
#define MOVB32(a,b) asm volatile("v_mov_b32 %0, %1" : "+v"(a) : "v"(b));
#define SUBREV32(a,b) asm volatile("v_subrev_u32 %0, %0, %1" : "+v"(a) : "v"(b));
#define MULU24E(a,b) asm volatile("v_mul_u32_u24 %0, 1000, %1" : "+v"(a) : "v"(b));
#define ADDLIT(a,b) asm volatile("v_add_u32 %0, 0x12345, %0" : "+v"(a));
#define BFE(a,b) asm volatile("v_bfe_u32 %0, %0, 5, 16" : "+v"(a));
#define CVT(a,b) asm volatile("v_cvt_f32_u32 %0, %0" : "+v"(a));
DEF(k_sub32, X8(SUB32))
DEF(k_mul24, X8(MUL24))
DEF(k_mulhi24, X8(MULHI24))
DEF(k_lshl32, X8(LSHL32))
DEF(k_lshr32, X8(LSHR32))
DEF(k_and32, X8(AND32))
DEF(k_or32, X8(OR32))
DEF(k_max32, X8(MAX32))
DEF(k_addcoe64, X8(ADDCOE64))
DEF(k_add32e64, X8(ADD32E64))
DEF(k_movb32, X8(MOVB32))
DEF(k_mul24lit, X8(MULU24E))
DEF(k_addlit, X8(ADDLIT))
DEF(k_bfe, X8(BFE))
DEF(k_cvt, X8(CVT))

#define MIX1(a,b) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b)); asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(b) : "v"(a));
DEF(k_mix_add_mullo, MIX1(a0,b0) MIX1(a1,b1) MIX1(a2,b2) MIX1(a3,b3))
#define MIX2(a,b) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a) : "v"(b)); asm volatile("v_add_co_u32 %0, vcc, %0, %1" : "+v"(b) : "v"(a) : "vcc");
DEF(k_mix_add_addco, MIX2(a0,b0) MIX2(a1,b1) MIX2(a2,b2) MIX2(a3,b3))

#define CNDS(a,b) asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a) : "v"(b), "s"(msk));
#define SUBBS(a,b) asm volatile("v_subb_co_u32_e64 %0, %2, %0, %1, %2" : "+v"(a) : "v"(b), "s"(msk));
#define DEFM(NAME, BODY) \
__global__ __launch_bounds__(256) void NAME(u32 *out, u32 seed) { \
  u32 a0 = threadIdx.x ^ seed, a1 = a0 * 3, a2 = a0 * 5, a3 = a0 * 7, a4 = a0 * 11, a5 = a0 * 13, a6 = a0 * 17, a7 = a0 * 19; \
  u32 b0 = a0 + 1, b1 = a1 + 1, b2 = a2 + 1, b3 = a3 + 1, b4 = a4 + 1, b5 = a5 + 1, b6 = a6 + 1, b7 = a7 + 1; \
  u64 msk; asm volatile("s_mov_b64 %0, 0x5555" : "=s"(msk)); \
  for (int i = 0; i < ITERS; i++) { BODY } \
  out[blockIdx.x * 256 + threadIdx.x] = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7 ^ b0 ^ b1 ^ b2 ^ b3 ^ b4 ^ b5 ^ b6 ^ b7; }
DEFM(k_cnds, X8(CNDS))
DEF(k_add32, X8(ADD32))
DEF(k_xor32, X8(XOR32))
DEF(k_addco, X8(ADDCO))
DEF(k_add3, X8(ADD3))
DEF(k_alignbit, X8(ALIGNB))
DEF(k_perm, X8(PERM))
DEF(k_mullo, X8(MULLO))
DEF(k_mulhi, X8(MULHI))
DEF(k_mad24, X8(MAD24))
DEF(k_dot2, X8(DOT2))
DEF(k_dot4, X8(DOT4))
DEF(k_pkadd, X8(PKADD))
DEF(k_pkmad, X8(PKMAD))
DEF(k_lshladd32, X8(LSHLADD32))

// 64-bit ops on register pairs
#define DEF64(NAME, BODY) \
__global__ __launch_bounds__(256) void NAME(u32 *out, u32 seed) { \
  u64 a0 = threadIdx.x ^ seed, a1 = a0 * 3, a2 = a0 * 5, a3 = a0 * 7, a4 = a0 * 11, a5 = a0 * 13, a6 = a0 * 17, a7 = a0 * 19; \
  u32 c0 = (u32)a0 + 1, c1 = (u32)a1 + 1, c2 = (u32)a2 + 1, c3 = (u32)a3 + 1, c4 = (u32)a4 + 1, c5 = (u32)a5 + 1, c6 = (u32)a6 + 1, c7 = (u32)a7 + 1; \
  u64 b = seed * 0x9e37ull; \
  for (int i = 0; i < ITERS; i++) { BODY } \
  out[blockIdx.x * 256 + threadIdx.x] = (u32)(a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7) ^ c0 ^ c1 ^ c2 ^ c3 ^ c4 ^ c5 ^ c6 ^ c7; }
#define Y8(S) S(a0,c0) S(a1,c1) S(a2,c2) S(a3,c3) S(a4,c4) S(a5,c5) S(a6,c6) S(a7,c7)
#define MAD64(a,c) asm volatile("v_mad_u64_u32 %0, vcc, %1, %1, %0" : "+v"(a) : "v"(c) : "vcc");
#define LSHLADD64(a,c) asm volatile("v_lshl_add_u64 %0, %0, 3, %1" : "+v"(a) : "v"(b));
#define ADD64(a,c) asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(a) : "v"(b));
#define LSHL64(a,c) asm volatile("v_lshlrev_b64 %0, 3, %0" : "+v"(a));
#define CMP64(a,c) asm volatile("v_cmp_lt_u64 vcc, %0, %1" : : "v"(a), "v"(b) : "vcc");
#define MOV64(a,c) asm volatile("v_mov_b64 %0, %1" : "=v"(a) : "v"(b));
DEF64(k_mad64, Y8(MAD64))
DEF64(k_lshladd64, Y8(LSHLADD64))
DEF64(k_add64, Y8(ADD64))
DEF64(k_lshl64, Y8(LSHL64))
DEF64(k_cmp64, Y8(CMP64))
DEF64(k_mov64, Y8(MOV64))

typedef void (*KF)(u32 *, u32);
int main()
{
    const int blocks = 256 * 8 * 4;  // 32 waves per CU... 8 WGs of 4 waves per CU * 4
    u32 *out; hipMalloc(&out, blocks * 256 * 4);
    struct { const char *n; KF f; } ks[] = {
        {"v_add_u32", k_add32}, {"v_xor_b32", k_xor32}, {"v_add_co_u32", k_addco}, {"v_add3_u32", k_add3},
         {"v_alignbit_b32", k_alignbit}, {"v_perm_b32", k_perm},
        {"v_mul_lo_u32", k_mullo}, {"v_mul_hi_u32", k_mulhi}, {"v_mad_u32_u24", k_mad24},
        {"v_dot2_u32_u16", k_dot2}, {"v_dot4_u32_u8", k_dot4}, {"v_pk_add_u16", k_pkadd}, {"v_pk_mad_u16", k_pkmad},
        {"v_lshl_add_u32", k_lshladd32},
        {"mix add+mul_lo", k_mix_add_mullo}, {"mix add+add_co", k_mix_add_addco},
        {"v_cndmask_b32_e64 smask", k_cnds},
        {"v_sub_u32", k_sub32}, {"v_mul_u32_u24", k_mul24}, {"v_mul_hi_u32_u24", k_mulhi24}, {"v_lshlrev_b32", k_lshl32},
        {"v_lshrrev_b32", k_lshr32}, {"v_and_b32", k_and32}, {"v_or_b32", k_or32}, {"v_max_u32", k_max32},
        {"v_add_co_u32_e64(s)", k_addcoe64}, {"v_add_u32_e64", k_add32e64}, {"v_mov_b32", k_movb32},
        {"v_mul_u32_u24 lit", k_mul24lit}, {"v_add_u32 lit", k_addlit}, {"v_bfe_u32", k_bfe}, {"v_cvt_f32_u32", k_cvt},
        {"v_mad_u64_u32", k_mad64}, {"v_lshl_add_u64(sh3)", k_lshladd64}, {"v_lshl_add_u64(sh0)", k_add64},
        {"v_lshlrev_b64", k_lshl64}, {"v_cmp_lt_u64", k_cmp64}, {"v_mov_b64", k_mov64}};
    hipEvent_t e0, e1; hipEventCreate(&e0); hipEventCreate(&e1);
    for (auto &k : ks) {
        hipLaunchKernelGGL(k.f, dim3(blocks), dim3(256), 0, 0, out, 1u);  // warm
        hipEventRecord(e0);
        for (int r = 0; r < 5; r++) hipLaunchKernelGGL(k.f, dim3(blocks), dim3(256), 0, 0, out, 1u + r);
        hipEventRecord(e1); hipEventSynchronize(e1);
        float ms; hipEventElapsedTime(&ms, e0, e1);
        const double waveinstr = 5.0 * blocks * 4 * (double)ITERS * 8;  // 4 waves per block
        const double simd_cycles = ms * 1e-3 * 2.4e9 * 1024;
        printf("%-22s %8.3f ms  %5.2f cycles per wave-instruction per SIMD (2.4 GHz)\n", k.n, ms, simd_cycles / waveinstr);
    }
    hipError_t e = hipGetLastError();
    printf("last error: %s\n", hipGetErrorString(e));
    return 0;
}
