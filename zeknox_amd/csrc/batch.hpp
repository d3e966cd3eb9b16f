// Batched commit path (batch.hip).
#pragma once
#include "runtime.hpp"

namespace zk {

// Same field layout as zk_commit_job of include/zeknox_ext.h.
struct BatchJob {
    const u64 *coeffs;
    u64 *lde;
    u64 *leaves;  // optional row-major transposed leaves (NULL: skip)
    u64 *digests;
    u64 *cap;
    u32 lg_n, rate_bits, ncols, cap_height;
};

// Runs every job's LDE and Merkle forest; returns when all are complete.
// use_graph: capture the batch into a hipGraph on the first repeat of the
// same job list and replay it afterwards.
void batch_commit(DeviceCtx &c, const BatchJob *jobs, u32 n, bool use_graph);
void batch_clear(DeviceCtx &c);

}  // namespace zk
