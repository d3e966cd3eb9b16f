/* Merkle-tree entry points of the zeknox C ABI (okx/zeknox native/merkle/merkle.h:21-62).
 *
 * Buffers: leaves are n_leaves x leaf_size u64 row-major; digests hold
 * 2 * (n_leaves - 2^cap_height) digests of 4 u64 in the Plonky2 linear layout;
 * cap holds 2^cap_height digests.  Only HashPoseidon is implemented in this
 * build; other hash types print an error and leave the buffers untouched.
 * Unlike the reference (which returns with kernels still queued on the legacy
 * stream), the GPU entry points return with the results complete. */
#ifndef ZEKNOX_MERKLE_H
#define ZEKNOX_MERKLE_H

#include <stdint.h>
#include "types/int_types.h"

#define HASH_SIZE 32
#define HASH_SIZE_U64 4

#ifdef __cplusplus
#define ZK_EXTERNC extern "C"
#else
#define ZK_EXTERNC
#endif

enum HashType {
    HashPoseidon = 0,
    HashKeccak = 1,
    HashPoseidonBN128 = 2,
    HashPoseidon2 = 3,
    HashMonolith = 4
};

/* merkle.cu:246-280: device pointers; runs on device gpu_id */
ZK_EXTERNC void fill_digests_buf_linear_gpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                         void *leaves_buf_gpu_ptr, u64 digests_buf_size,
                                                         u64 cap_buf_size, u64 leaves_buf_size, u64 leaf_size,
                                                         u64 cap_height, u64 hash_type, u64 gpu_id);

/* merkle.cu:283-499: the reference splits subtrees across GPUs; this build
 * runs the whole forest on device 0 (pointers must be device-0 memory). */
ZK_EXTERNC void fill_digests_buf_linear_multigpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                              void *leaves_buf_gpu_ptr, u64 digests_buf_size,
                                                              u64 cap_buf_size, u64 leaves_buf_size,
                                                              u64 leaf_size, u64 cap_height, u64 hash_type);

/* merkle.cpp:126-161: host pointers, computed on the CPU (OpenMP) */
ZK_EXTERNC void fill_digests_buf_linear_cpu(void *digests_buf_ptr, void *cap_buf_ptr, const void *leaves_buf_ptr,
                                            u64 digests_buf_size, u64 cap_buf_size, u64 leaves_buf_size,
                                            u64 leaf_size, u64 cap_height, u64 hash_type);

#undef ZK_EXTERNC
#endif
