/* The zeknox C ABI for the Goldilocks NTT / coset LDE / transpose path, as
 * exported by libzeknox.so for MI355X (gfx950).  Declarations match
 * okx/zeknox native/lib.h:21-51 (plus the exported-but-undeclared
 * init_cuda_degree_and_generator of native/lib.cu:141 and cuda_available of
 * native/utils/all_gpus.cu:68), so the Rust (wrappers/rust/src/lib.rs:12-96)
 * and Go (wrappers/go/lib/lib.go) bindings link against it unchanged.
 *
 * Semantics common to the NTT entry points: columns are stored one after the
 * other ("batches"), every call returns after the device work completed, and
 * twiddle/coset tables are built by the library on first use of a device, so
 * the init_* calls are optional (they remain valid, cheap calls). */
#ifndef ZEKNOX_LIB_H
#define ZEKNOX_LIB_H

#include <stddef.h>
#include <stdint.h>
#include <stdbool.h>
#include "utils/rusterror.h"
#include "ntt/ntt.h"
#include "merkle/merkle.h"

#ifdef __cplusplus
#define EXTERN extern "C"
#else
#define EXTERN
#endif

/* lib.cu:59-66 (all_gpus.cu:53-59): number of visible GPUs */
EXTERN RustError get_number_of_gpus(size_t *ngpus);

/* lib.cu:37-44 (all_gpus.cu:150-184): prints the device properties */
EXTERN RustError list_devices_info();

/* lib.cu:170-174: init_cuda_degree(24) */
EXTERN void init_cuda();

/* lib.cu:158-166: tables for sizes up to 2^max_degree, coset generator 7 */
EXTERN void init_cuda_degree(const uint32_t max_degree);

/* lib.cu:141-155: as init_cuda_degree with an explicit coset generator */
EXTERN void init_cuda_degree_and_generator(const uint32_t max_degree, const uint64_t group_generator);

/* lib.cu:116-123 (ntt.cu:160-173): twiddles of order 2^lg_n on device_id */
EXTERN RustError init_twiddle_factors(size_t device_id, size_t lg_n);

/* lib.cu:125-132 (ntt.cu:228-239): coset generator used by with_coset */
EXTERN RustError init_coset(size_t device_id, size_t lg_domain_size, const uint64_t coset_gen);

/* lib.cu:70-78 (ntt.cu:253-309): in-place batched NTT of size 2^lg_domain_size,
 * natural order in and out; inverse scales by 2^-lg; with_coset multiplies the
 * input by g^i (forward) or the output by g^i (inverse, as the reference). */
EXTERN RustError compute_batched_ntt(size_t device_id, void *inout, uint32_t lg_domain_size,
                                     NTT_Direction ntt_direction, NTT_Config cfg);

/* lib.cu:80-89 (ntt.cu:477-564): coset LDE.  input: batches x 2^lg_domain_size
 * coefficients; output: (batches + salt_size) x 2^(lg_domain_size + rate_bits)
 * evaluations, natural order: out[b][i] = p_b(g * w^i). */
EXTERN RustError compute_batched_lde(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                                     NTT_Direction ntt_direction, NTT_Config cfg);

/* lib.cu:91-98 (ntt.cu:315-471): single-GPU build; runs compute_batched_lde on
 * device 0 over total_num_input_elements / 2^lg_domain_size columns. */
EXTERN RustError compute_batched_lde_multi_gpu(void *output, void *input, uint32_t num_gpu,
                                               NTT_Direction ntt_direction, NTT_Config cfg,
                                               uint32_t lg_domain_size, size_t total_num_input_elements,
                                               size_t total_num_output_elements);

/* lib.cu:100-107 (ntt.cu:182-226): out[rev(i)][j] = in[j][i], in is
 * batches x 2^lg_n column-major, out is 2^lg_n x batches row-major. */
EXTERN RustError compute_transpose_rev(size_t device_id, void *output, void *input, uint32_t lg_n,
                                       NTT_TransposeConfig cfg);

/* all_gpus.cu:68: true when a GPU is usable */
EXTERN bool cuda_available();

#endif
