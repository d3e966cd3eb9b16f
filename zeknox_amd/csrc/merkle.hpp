// Host-side interface of the Poseidon Merkle kernels (merkle.hip).
#pragma once
#include <memory>
#include <vector>

#include "forest_map.hpp"
#include "runtime.hpp"

namespace zk {

// Shape of a Plonky2 linear-layout forest of 2^cap_height subtrees.
struct TreeShape {
    u64 n_leaves = 0, leaf_size = 0, cap_height = 0;
    u64 n_caps() const { return 1ull << cap_height; }
    u64 sub_leaves() const { return n_leaves >> cap_height; }  // L
    u64 n_digests() const { return n_caps() == n_leaves ? 0 : 2 * (n_leaves - n_caps()); }
    u64 sub_digests() const { return n_digests() >> cap_height; }  // D = 2L - 2
};

// Validates the reference's (digests, caps, leaves) size triple; throws on a
// shape the linear layout cannot describe.
TreeShape tree_shape(u64 n_digests, u64 n_caps, u64 n_leaves, u64 leaf_size, u64 cap_height);

// Leaf layout in memory:
//  row-major  leaf l at leaves[l * leaf_size + k]  (the ABI layout)
//  col-major  leaf l = row rev(l) of an LDE matrix: leaves[k * col_stride + rev(l)]
enum class LeafLayout { RowMajor, ColMajorRev };

// Fills digests (linear layout) and caps from leaves on `stream`.  With the
// column-major layout, a non-null rows_out also receives the row-major leaf
// matrix (leaf l at rows_out[l * leaf_size], as compute_transpose_rev writes).
void merkle_launch(const TreeShape &t, u64 *digests, u64 *caps, const u64 *leaves, LeafLayout layout,
                   u64 col_stride, hipStream_t stream, u64 *rows_out = nullptr);

// The same forest with the hasher of a Merkle HashType: 0 Poseidon, 1 Keccak,
// 3 Poseidon2 (throws for the others).
void merkle_launch_hashed(const TreeShape &t, u64 *digests, u64 *caps, const u64 *leaves, LeafLayout layout,
                          u64 col_stride, hipStream_t stream, u32 hash_type);

// Merkle openings of leaves dev_indices[0..n_idx) (device array, each <
// n_leaves): siblings_out gets lg(n_leaves) - cap_height digests per query
// (leaf level first); rows_out (optional) the leaf rows, read from row-major
// leaves (col_stride 0) or a column-major LDE (leaf l = row rev(l)).
void proofs_launch(const TreeShape &t, const u64 *digests, const u64 *leaves, u64 col_stride, const u64 *dev_indices,
                   u64 n_idx, u64 *siblings_out, u64 *rows_out, hipStream_t stream);

// The Poseidon forests of many commits at once (batched commits): leaf l of
// a job is row rev(l) of its column-major LDE (as merkle_launch with
// LeafLayout::ColMajorRev and col_stride 2^lg_leaves); rows (optional)
// receives the row-major leaf matrix; digests and caps as merkle_launch.
// (ForestJob: forest_map.hpp.)
struct ForestPlan;
struct ForestPlanDeleter {
    void operator()(ForestPlan *p) const;
};
using ForestPlanPtr = std::unique_ptr<ForestPlan, ForestPlanDeleter>;

// Device tables of the launches (allocates and copies synchronously).
ForestPlanPtr forest_plan(const std::vector<ForestJob> &jobs);

// Enqueues the forests on `stream`: one launch for all leaves (two when some
// jobs write leaf rows and some do not), then one per tree height, each
// covering every job; kernels only, so it can be captured into a graph.
void forest_launch(const ForestPlan &plan, hipStream_t stream);

// The same from tables the caller placed in device memory: `launches` and
// the bytes of forest_tables (forest_map.hpp), copied to `tables`.
void forest_launch_tables(const std::vector<ForestLaunch> &launches, const void *tables, hipStream_t stream);

}  // namespace zk
