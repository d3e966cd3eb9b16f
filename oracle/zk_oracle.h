/*
 * CPU oracle for the Plonky2 commit path (Goldilocks coset LDE -> Poseidon
 * Merkle).  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU
 * baseline leg of bench.py load this library as the checker.  The product
 * (zeknox_amd/) never includes, links or calls anything under oracle/.
 *
 * Every function restates the algorithm of the reference (okx/zeknox) named
 * in its comment in plain C; the citation is the reference file:line whose
 * observable behaviour it follows.  All values are canonical (< p) on output.
 */
#ifndef ZK_ORACLE_H
#define ZK_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_P 0xFFFFFFFF00000001ull

/* field (native/ff/goldilocks.hpp:139-177, native/ff/gl64_t.cuh:93-369) */
uint64_t or_gl_add(uint64_t a, uint64_t b);
uint64_t or_gl_sub(uint64_t a, uint64_t b);
uint64_t or_gl_mul(uint64_t a, uint64_t b);
uint64_t or_gl_pow(uint64_t a, uint64_t e);
uint64_t or_gl_inv(uint64_t a);
/* reduce a 128-bit value given as (lo, hi) */
uint64_t or_gl_reduce128(uint64_t lo, uint64_t hi);
/* root of unity of order 2^lg, 7^((p-1)/2^lg) (native/ff/gl64_params.hpp:15-49) */
uint64_t or_omega(uint32_t lg);
uint64_t or_omega_inv(uint32_t lg);

/* Poseidon (native/poseidon/poseidon.cu:663-780, naive round schedule) */
void or_poseidon_permute(uint64_t st[12]);
/* sponge (native/poseidon/poseidon_permutation.hpp:84-140) */
void or_hash_one(const uint64_t *in, uint64_t n, uint64_t out[4]);
void or_hash_two(const uint64_t *a, const uint64_t *b, uint64_t out[4]);

/* Merkle linear layout (native/merkle/merkle.cpp:19-124, merkle.cu:165-244) */
int or_fill_digests_linear(uint64_t *digests, uint64_t *caps, const uint64_t *leaves,
                           uint64_t n_digests, uint64_t n_caps, uint64_t n_leaves,
                           uint64_t leaf_size, uint64_t cap_height);

/* NTT, natural order in and out, col-major batches (native/ntt/ntt.cu:253-309) */
int or_merkle_proof(const uint64_t *digests, uint64_t n_leaves, uint64_t cap_height, uint64_t index,
                    uint64_t *out);
void or_poseidon2_permute(uint64_t st[12]);
void or_p2_hash_one(const uint64_t *in, uint64_t n, uint64_t out[4]);
void or_p2_hash_two(const uint64_t *a, const uint64_t *b, uint64_t out[4]);
void or_keccak_f(uint64_t A[25]);
void or_keccak256(const uint8_t *in, uint64_t n, uint64_t out[4]);
void or_keccak_hash_one(const uint64_t *in, uint64_t n, uint64_t out[4]);
void or_keccak_hash_two(const uint64_t *a, const uint64_t *b, uint64_t out[4]);
int or_fill_digests_linear_ht(uint64_t *digests, uint64_t *caps, const uint64_t *leaves, uint64_t n_digests,
                              uint64_t n_caps, uint64_t n_leaves, uint64_t leaf_size, uint64_t cap_height, int ht);
void or_ntt(uint64_t *inout, uint32_t lg, uint64_t batches, int inverse,
            int with_coset, uint64_t coset_gen);
/* coset LDE (native/ntt/ntt.cu:477-564): out has batches x 2^(lg_n+rate_bits) */
void or_lde(uint64_t *out, const uint64_t *in, uint32_t lg_n, uint32_t rate_bits,
            uint64_t batches, int inverse, int with_coset, uint64_t coset_gen);
/* O(N^2) DFT of one column: out[i] = sum_j in[j] * (shift * w^i)^j, w = omega(lg)
 * (or omega_inv when inverse).  Independent check for small N. */
uint64_t or_poly_eval(const uint64_t *c, uint64_t n, uint64_t x);
void or_dft_naive(uint64_t *out, const uint64_t *in, uint32_t lg, int inverse, uint64_t shift);
/* out[rev(i)][j] = in[j][i] (native/ntt/kernels.cu:94-135) */
void or_transpose_rev(uint64_t *out, const uint64_t *in, uint32_t lg_n, uint64_t batches);

/* seeded inputs: xoshiro256** state from splitmix64(seed), rejection to [0,p) */
void or_fill_random(uint64_t *out, uint64_t n, uint64_t seed);
uint64_t or_splitmix64(uint64_t x);

/* thread count used by the OpenMP loops (0 = runtime default) */
void or_set_threads(int n);
int or_get_threads(void);

#ifdef __cplusplus
}
#endif
#endif
