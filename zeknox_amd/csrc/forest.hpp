// Multi-job Poseidon Merkle forests in one persistent launch (forest.hip).
#pragma once
#include <vector>

#include "merkle.hpp"

namespace zk {

// One tree job: leaves are row-major (col_stride == 0: leaf l at
// leaves[l * leaf_size]) or the rows of a column-major LDE matrix
// (leaf l = row rev(l): leaves[k * col_stride + rev(l)]).
struct ForestSpec {
    const u64 *leaves = nullptr;
    u64 col_stride = 0;
    u64 *digests = nullptr;
    u64 *caps = nullptr;
    TreeShape shape;
};

// A prepared launch: item descriptors and per-job bookkeeping in device
// memory (owned), reusable across launches (e.g. from a captured hipGraph).
struct ForestPlan {
    u32 njobs = 0, nitems = 0, nheights = 0;
    void *dev_jobs = nullptr;   // ForestJob[njobs]
    void *dev_items = nullptr;  // u32x2[nitems]: (job << 8 | height, chunk)
    u32 *dev_sync = nullptr;    // [0] queue head, [1] error flag, [2..] done counters njobs x kMaxHeights
    u32 grid = 0;
    std::vector<u32> height_first;  // items of height h: [height_first[h], height_first[h+1])
    ~ForestPlan();
};

ForestPlan *forest_plan(const std::vector<ForestSpec> &jobs, int device);
// One launch per tree height over the items of every job, on `stream` (the
// stream order provides the level dependencies; kernels only, graph-safe).
void forest_launch(const ForestPlan &plan, hipStream_t stream);
// Persistent one-launch form: disabled (throws).
void forest_launch_persistent(const ForestPlan &plan, hipStream_t stream);
// Throws if a bounded dependency wait of a persistent launch timed out.
void forest_check(const ForestPlan &plan, hipStream_t stream);

}  // namespace zk
