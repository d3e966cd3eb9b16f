/* NTT configuration types of the zeknox C ABI.
 * Same names, enumerator values and field order as okx/zeknox
 * native/ntt/ntt.h:17-62.  C callers (cgo) see the flags as char, C++ callers
 * (Rust's repr(C) bool) as bool; both are one byte, so NTT_Config is 24 bytes
 * and NTT_TransposeConfig 8 bytes in either language. */
#ifndef ZEKNOX_NTT_H
#define ZEKNOX_NTT_H

#include <stdint.h>

#ifdef __cplusplus
#define ZK_BOOL bool
#else
#define ZK_BOOL char
#endif

typedef enum { NN, NR, RN, RR } NTT_InputOutputOrder; /* only NN is implemented, as in the reference */

typedef enum { forward, inverse } NTT_Direction;

typedef enum { standard, coset } NTT_Type;

typedef enum { GS, CT } NTT_Algorithm;

typedef struct {
    uint32_t batches;              /* number of columns (polynomials) */
    NTT_InputOutputOrder order;    /* ignored (always natural -> natural) */
    NTT_Type ntt_type;             /* ignored; with_coset selects the coset */
    uint32_t extension_rate_bits;  /* LDE blowup = 2^extension_rate_bits */
    ZK_BOOL are_inputs_on_device;  /* input pointer is device memory */
    ZK_BOOL are_outputs_on_device; /* output pointer is device memory */
    ZK_BOOL with_coset;            /* evaluate on the coset g * <w> */
    ZK_BOOL is_multi_gpu;          /* ignored (single GPU) */
    uint32_t salt_size;            /* extra random columns appended to the LDE output */
} NTT_Config;

typedef struct {
    uint32_t batches;
    ZK_BOOL are_inputs_on_device;
    ZK_BOOL are_outputs_on_device;
} NTT_TransposeConfig;

#undef ZK_BOOL

#endif
