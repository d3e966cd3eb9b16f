// Index maps of the batched (multi-job) Poseidon forest launches in
// merkle.hip: which job, and which leaf or tree node of it, a thread of a
// launch that covers many commits works on.
//
// Every job of a launch owns a run of whole workgroups starting at its
// `block0` (the table is sorted by block0, block0 of the first entry is 0), so
// the job of a workgroup is found by a binary search that is uniform across
// the workgroup, and lanes past the job's last leaf or node stay idle (or,
// where lanes cooperate on one permutation, hash node 0 again and store
// nothing).  Shared with the host test of the plans (tests/native).
#pragma once
#include <string.h>

#include <algorithm>
#include <vector>

#include "gl64.hpp"

namespace zk {

constexpr int kLeafThreads = 128;  // threads per workgroup of the leaf kernels
constexpr int kNodeThreads = 128;  // threads per workgroup of the level kernels
// Levels with at most this many nodes (all subtrees) use a quad of lanes per
// node.
constexpr u64 kQuadMaxNodes = 1u << 16;
// Levels with at most this many nodes use 16 lanes per node: 14-17 us per
// level against 20 us with a quad (2^11-2^12 nodes; each lane does one S-box
// and one MDS row per round, against three of each, plus a lane exchange per
// element; 4/3 the work of a quad, so wider levels keep the quad).
constexpr u64 kHexMaxNodes = 1u << 12;
// Workgroups of one forest launch at most (HIP counts a grid in threads, as a
// 32-bit number); a longer launch is split, each part with its own table.
constexpr u64 kForestMaxBlocks = 1ull << 23;

// Arguments of one leaf-hash launch (merkle.hip).
struct LeafArgs {
    const u64 *leaves;
    u64 n_leaves;
    u32 leaf_size;
    u32 lg_leaves;
    u64 col_stride;
    u64 *out;        // digests (or caps for the direct case)
    u64 sub_leaves;  // L (0: direct mapping, out[l])
    u64 sub_digests; // D
    u64 *rows;       // column-major layout only: optional row-major copy of the leaves
};

// Index of the last table entry with block0 <= b.
template <typename T>
ZK_HD u32 forest_find(const T *jobs, u32 njobs, u64 b)
{
    u32 lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const u32 mid = (lo + hi) >> 1;
        if (jobs[mid].block0 <= b)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// One job's part of a tree-height launch: the level with 2^lg_nodes nodes per
// subtree (lg_nodes == 0: the cap entry of every subtree).
struct ForestLevelJob {
    u64 *d;        // the job's digests, linear layout
    u64 *caps;     // the job's caps
    u64 D;         // digests per subtree
    u64 total;     // nodes of this level over all subtrees
    u64 block0;    // first workgroup of the job in the launch
    u32 lg_nodes;  // log2 nodes per subtree
    u32 pad;
};

// Node x (< total) of a level: subtree s, first child c0 (the second is
// c0 + 1) and output index within the subtree's digests; out < 0 is the
// subtree's cap entry.  A level with m = 2^lg_nodes nodes per subtree starts
// at digest m - 2 and node i hashes digests 2i + 2 and 2i + 3; the cap hashes
// digests 0 and 1.
struct ForestNode {
    u64 s;
    u64 c0;
    long long out;
};

ZK_HD ForestNode forest_node(const ForestLevelJob &J, u64 x)
{
    const u64 s = x >> J.lg_nodes;
    if (J.lg_nodes == 0) return {s, 0, -1};
    const u64 node = (1ull << J.lg_nodes) - 2 + (x & ((1ull << J.lg_nodes) - 1));
    return {s, 2 * node + 2, (long long)node};
}

// One job's part of a multi-job leaf launch.
struct ForestLeafJob {
    LeafArgs a;
    u64 block0;  // first workgroup of the job in the launch
};

// One commit of a batched forest: leaf l is row rev(l) of the column-major
// LDE (2^lg_leaves rows); rows (optional) receives the row-major leaf matrix;
// digests (linear layout, unused when caps == leaves) and caps.
struct ForestJob {
    const u64 *lde;
    u64 *rows;
    u64 *digests;
    u64 *caps;
    u32 lg_leaves, leaf_size, cap_height;
};

// One launch of a batched forest: its kind, lanes per node, workgroups, and
// its table (njobs entries at byte offset off of the plan's buffer).
struct ForestLaunch {
    enum Kind { LeavesRows, Leaves, Level };
    Kind kind;
    int g;
    u32 blocks;
    u32 njobs;
    size_t off;
};

template <typename T>
inline void forest_append(std::vector<unsigned char> &bytes, std::vector<ForestLaunch> &launches,
                          ForestLaunch::Kind kind, int g, std::vector<T> &table, u64 blocks)
{
    if (table.empty()) return;
    const size_t off = (bytes.size() + 255) & ~(size_t)255;
    bytes.resize(off + table.size() * sizeof(T));
    memcpy(bytes.data() + off, table.data(), table.size() * sizeof(T));
    launches.push_back({kind, g, (u32)blocks, (u32)table.size(), off});
    table.clear();
}

// The launches of a batched forest, in order, and their tables: the leaves
// of the jobs that write leaf rows, the leaves of the others, then one launch
// per tree height k = 1, 2, ... (the level k above the leaves of every job
// whose subtrees are at least k tall; a job's last height is its caps), each
// level's children written by an earlier launch.  Returns false for an
// invalid job (more than 2^31 leaves, cap_height above lg_leaves, empty
// leaves).
inline bool forest_tables(const std::vector<ForestJob> &jobs, std::vector<unsigned char> &bytes,
                          std::vector<ForestLaunch> &launches)
{
    auto blocks_for = [](u64 n, int tpb) { return (n + tpb - 1) / tpb; };
    for (const ForestJob &j : jobs)
        if (j.lg_leaves > 31 || j.cap_height > j.lg_leaves || j.leaf_size == 0 || !j.lde || !j.caps ||
            (j.cap_height < j.lg_leaves && !j.digests))
            return false;
    for (int pass = 0; pass < 2; pass++) {
        const bool rows = pass == 0;
        const ForestLaunch::Kind kind = rows ? ForestLaunch::LeavesRows : ForestLaunch::Leaves;
        std::vector<ForestLeafJob> table;
        u64 blocks = 0;
        for (const ForestJob &j : jobs) {
            if ((j.rows != nullptr) != rows) continue;
            const u64 N = 1ull << j.lg_leaves, nc = 1ull << j.cap_height;
            const u64 L = N >> j.cap_height;
            ForestLeafJob e{};
            e.a.leaves = j.lde;
            e.a.rows = j.rows;
            e.a.n_leaves = N;
            e.a.leaf_size = j.leaf_size;
            e.a.lg_leaves = j.lg_leaves;
            e.a.col_stride = N;
            e.a.out = nc == N ? j.caps : j.digests;
            e.a.sub_leaves = (nc == N || L == 2) ? 0 : L;
            e.a.sub_digests = 2 * L - 2;
            const u64 nb = blocks_for(N, kLeafThreads);
            if (blocks + nb > kForestMaxBlocks) {
                forest_append(bytes, launches, kind, 1, table, blocks);
                blocks = 0;
            }
            e.block0 = blocks;
            table.push_back(e);
            blocks += nb;
        }
        forest_append(bytes, launches, kind, 1, table, blocks);
    }
    u32 maxh = 0;
    for (const ForestJob &j : jobs)
        if (j.cap_height < j.lg_leaves) maxh = std::max(maxh, j.lg_leaves - j.cap_height);
    for (u32 k = 1; k <= maxh; k++) {
        u64 total = 0;
        for (const ForestJob &j : jobs) {
            const u32 lg_sub = j.lg_leaves - j.cap_height;
            if (j.cap_height < j.lg_leaves && lg_sub >= k) total += 1ull << (lg_sub - k + j.cap_height);
        }
        const int g = total > kQuadMaxNodes ? 1 : (total > kHexMaxNodes ? 4 : 16);
        std::vector<ForestLevelJob> table;
        u64 blocks = 0;
        for (const ForestJob &j : jobs) {
            const u32 lg_sub = j.lg_leaves - j.cap_height;
            if (j.cap_height >= j.lg_leaves || lg_sub < k) continue;
            ForestLevelJob e{};
            e.d = j.digests;
            e.caps = j.caps;
            e.D = (2ull << lg_sub) - 2;
            e.lg_nodes = lg_sub - k;
            e.total = 1ull << (e.lg_nodes + j.cap_height);
            const u64 nb = blocks_for(e.total * (u64)g, kNodeThreads);
            if (blocks + nb > kForestMaxBlocks) {
                forest_append(bytes, launches, ForestLaunch::Level, g, table, blocks);
                blocks = 0;
            }
            e.block0 = blocks;
            table.push_back(e);
            blocks += nb;
        }
        forest_append(bytes, launches, ForestLaunch::Level, g, table, blocks);
    }
    return true;
}

}  // namespace zk
