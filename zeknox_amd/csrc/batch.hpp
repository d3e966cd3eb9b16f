// Batched commit path (batch.hip).
#pragma once
#include "commit.hpp"

namespace zk {

// Runs every job's commit; returns when all are complete.  use_graph: the
// second call with the same job list captures a hipGraph, later calls
// replay it.
void batch_commit(DeviceCtx &c, const CommitJob *jobs, u32 n, bool use_graph);
void batch_clear(DeviceCtx &c);

}  // namespace zk
