// Poseidon Merkle forest in the Plonky2 "linear" digest layout on gfx950.
//
// Same outputs as okx/zeknox fill_digests_buf_linear_gpu_with_gpu_ptr
// (native/merkle/merkle.cu:165-280) and the CPU path
// (native/merkle/merkle.cpp:19-124): per subtree of L leaves the D = 2L-2
// digests hold the leaf hashes at [D-L, D), internal node i = H(d[2i+2],
// d[2i+3]) level by level, and cap = H(d[0], d[1]); caps == leaves hashes the
// leaves straight into the cap; L == 2 keeps the two leaf hashes at d[0..2).
//
// One Poseidon sponge per lane.  Leaves are read either row-major (the ABI
// layout) or straight from a column-major LDE matrix at row rev(l), where
// consecutive lanes read consecutive rows of every column (coalesced).
#include <algorithm>
#include <string>
#include <vector>

#include "hashers.hpp"
#include "forest_map.hpp"
#include "leaf_rows.hpp"
#include "merkle.hpp"
#include "poseidon.hpp"
#include "poseidon_hex.hpp"
#include "poseidon_quad.hpp"

namespace zk {

// Trees with at most this many leaves use a quad of lanes per leaf.
constexpr u64 kQuadMaxLeaves = 1u << 15;

static u32 ilog2(u64 x)
{
    u32 r = 0;
    while (x > 1) {
        x >>= 1;
        r++;
    }
    return r;
}

TreeShape tree_shape(u64 n_digests, u64 n_caps, u64 n_leaves, u64 leaf_size, u64 cap_height)
{
    auto bad = [](const char *m) { return Error((int)hipErrorInvalidValue, m); };
    if (n_leaves == 0 || (n_leaves & (n_leaves - 1))) throw bad("leaves count must be a power of two");
    if (cap_height >= 63 || n_caps != (1ull << cap_height)) throw bad("cap size must be 2^cap_height");
    if (n_caps > n_leaves) throw bad("more caps than leaves");
    if (leaf_size == 0 || leaf_size > 0xFFFFFFFFull) throw bad("invalid leaf size");
    TreeShape t;
    t.n_leaves = n_leaves;
    t.leaf_size = leaf_size;
    t.cap_height = cap_height;
    if (n_caps != n_leaves && n_digests != t.n_digests()) throw bad("digests size must be 2 * (leaves - caps)");
    return t;
}

// The work of leaf thread t (one sponge per lane).  ring: the workgroup's
// LDS ring of the row writer (ROWS only).
template <LeafLayout LAYOUT, bool ROWS>
__device__ __forceinline__ void leaf_hash_thread(const LeafArgs &a, u64 t, u64 *ring)
{
    if (t >= a.n_leaves) return;
    u64 leaf;
    u64 h[4];
    if constexpr (LAYOUT == LeafLayout::RowMajor) {
        leaf = t;
        const u64 *p = a.leaves + leaf * a.leaf_size;
        hash_one([p](u32 i) { return p[i]; }, a.leaf_size, h);
    } else {
        // lane t owns LDE row t, i.e. leaf rev(t); with a.rows it also writes
        // that row to the row-major leaf matrix (what compute_transpose_rev
        // produces), so the leaves are not read a second time, in whole
        // 64-byte sectors through a per-lane LDS ring (leaf_rows.hpp)
        leaf = zk_bitrev((u32)t, a.lg_leaves);
        const u64 *p = a.leaves + t;
        const u64 cs = a.col_stride;
        if constexpr (ROWS) {
            RowSectorWriter w(a.rows, leaf, a.leaf_size, ring + threadIdx.x * kRowRingPitch);
            hash_one_chunks([p, cs](u32 i) { return p[(u64)i * cs]; }, a.leaf_size, h,
                            [&w](u32, u32 m, const u64 *v) { w.chunk(m, v); });
            w.finish();
        } else {
            hash_one([p, cs](u32 i) { return p[(u64)i * cs]; }, a.leaf_size, h);
        }
    }
    u64 slot = leaf;
    if (a.sub_leaves) {
        const u64 s = leaf / a.sub_leaves, i = leaf % a.sub_leaves;
        slot = s * a.sub_digests + (a.sub_digests - a.sub_leaves) + i;
    }
    u64 *o = a.out + slot * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = h[k];
}

// At least four waves per SIMD: the leaf-row variant fits in 111 VGPRs.
template <LeafLayout LAYOUT, bool ROWS>
__global__ __launch_bounds__(kLeafThreads, 4) void leaf_hash_kernel(LeafArgs a)
{
    __shared__ u64 ring[ROWS ? kLeafThreads * kRowRingPitch : 1];
    leaf_hash_thread<LAYOUT, ROWS>(a, (u64)blockIdx.x * kLeafThreads + threadIdx.x, ring);
}

// The leaves of many commits in one launch (batched commits): workgroup b
// works for the job whose run of workgroups holds b (forest_map.hpp), with
// that job's arguments, exactly as that job's own leaf_hash_kernel would.
template <bool ROWS>
__global__ __launch_bounds__(kLeafThreads, 4) void leaf_hash_forest_kernel(const ForestLeafJob *__restrict__ jobs,
                                                                           u32 njobs)
{
    __shared__ u64 ring[ROWS ? kLeafThreads * kRowRingPitch : 1];
    const ForestLeafJob &J = jobs[forest_find(jobs, njobs, blockIdx.x)];
    leaf_hash_thread<LeafLayout::ColMajorRev, ROWS>(J.a, ((u64)blockIdx.x - J.block0) * kLeafThreads + threadIdx.x,
                                                    ring);
}

// Leaf hashing with a quad of lanes per leaf, for trees too small to fill
// the GPU with one lane per leaf (a third of the latency).  Lanes past the
// last leaf hash leaf 0 again and store nothing.
template <LeafLayout LAYOUT, bool ROWS>
__global__ __launch_bounds__(kLeafThreads) void leaf_hash_quad_kernel(LeafArgs a)
{
    const u64 t = ((u64)blockIdx.x * kLeafThreads + threadIdx.x) >> 2;
    const u32 j = threadIdx.x & 3;
    const bool ok = t < a.n_leaves;
    const u64 tc = ok ? t : 0;
    u64 leaf, h[3];
    if constexpr (LAYOUT == LeafLayout::RowMajor) {
        leaf = tc;
        const u64 *p = a.leaves + leaf * a.leaf_size;
        hash_one_quad([p](u32 i) { return p[i]; }, a.leaf_size, j, h);
    } else {
        leaf = zk_bitrev((u32)tc, a.lg_leaves);
        const u64 *p = a.leaves + tc;
        const u64 cs = a.col_stride;
        if constexpr (ROWS) {
            u64 *q = a.rows + leaf * a.leaf_size;
            const bool w = ok;
            hash_one_quad([p, cs, q, w](u32 i) {
                const u64 v = p[(u64)i * cs];
                if (w) q[i] = v;
                return v;
            }, a.leaf_size, j, h);
        } else {
            hash_one_quad([p, cs](u32 i) { return p[(u64)i * cs]; }, a.leaf_size, j, h);
        }
    }
    if (!ok) return;
    u64 slot = leaf;
    if (a.sub_leaves) {
        const u64 s = leaf / a.sub_leaves, i = leaf % a.sub_leaves;
        slot = s * a.sub_digests + (a.sub_digests - a.sub_leaves) + i;
    }
    u64 *o = a.out + slot * 4;
    if (j == 0) {
        o[0] = h[0];
        o[1] = h[1];
        o[2] = h[2];
    } else if (j == 1) {
        o[3] = h[0];
    }
}

// One level: 2^r nodes per subtree starting at `base`.
__global__ __launch_bounds__(kNodeThreads) void level_hash_kernel(u64 *d, u64 nodes, u64 base, u64 n_sub, u64 D)
{
    const u64 t = (u64)blockIdx.x * kNodeThreads + threadIdx.x;
    if (t >= nodes * n_sub) return;
    const u64 s = t / nodes, node = base + t % nodes;
    u64 *sd = d + s * D * 4;
    u64 l[4], r[4], o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        l[k] = sd[(2 * node + 2) * 4 + k];
        r[k] = sd[(2 * node + 3) * 4 + k];
    }
    hash_two(l, r, o);
#pragma unroll
    for (int k = 0; k < 4; k++) sd[node * 4 + k] = o[k];
}

// One level with a quad of lanes per node (levels too small to fill the GPU
// with one lane per node: a third of the latency).  Lanes past the last node
// hash node 0 again and store nothing, so every quad stays complete.
__global__ __launch_bounds__(kNodeThreads) void level_hash_quad_kernel(u64 *d, u64 nodes, u64 base, u64 n_sub,
                                                                       u64 D)
{
    const u64 t = (u64)blockIdx.x * kNodeThreads + threadIdx.x;
    const u32 j = threadIdx.x & 3;
    const u64 total = nodes * n_sub, x = t >> 2;
    const u64 xc = x < total ? x : 0;
    const u64 s = xc / nodes, node = base + xc % nodes;
    u64 *sd = d + s * D * 4;
    u64 out[3];
    hash_two_quad(sd + (2 * node + 2) * 4, sd + (2 * node + 3) * 4, j, out);
    if (x < total) {
        u64 *o = sd + node * 4;
        if (j == 0) {
            o[0] = out[0];
            o[1] = out[1];
            o[2] = out[2];
        } else if (j == 1) {
            o[3] = out[0];
        }
    }
}

// One level with 16 lanes per node (hash_two_hex), for levels whose latency
// rather than their work bounds them.  Groups past the last node hash node 0
// again and store nothing, so every group stays complete.
__global__ __launch_bounds__(kNodeThreads) void level_hash_hex_kernel(u64 *d, u64 nodes, u64 base, u64 n_sub,
                                                                      u64 D)
{
    const u64 t = (u64)blockIdx.x * kNodeThreads + threadIdx.x;
    const u32 j = threadIdx.x & 15, lb = threadIdx.x & 48;  // lane in the group, group's first wave lane
    const u64 total = nodes * n_sub, x = t >> 4;
    const u64 xc = x < total ? x : 0;
    const u64 s = xc / nodes, node = base + xc % nodes;
    u64 *sd = d + s * D * 4;
    const u64 out = hash_two_hex(sd + (2 * node + 2) * 4, sd + (2 * node + 3) * 4, j, lb);
    if (x < total && j < 4) sd[node * 4 + j] = out;
}

// One tree height of many commits in one launch (batched commits): G lanes
// per node (1, or 4 / 16 cooperating on one permutation for heights whose
// latency rather than their work bounds them).  Every node's children were
// written by an earlier launch on the same stream; there are no device-side
// waits.  Lanes past the job's last node hash node 0 again and store nothing.
template <int G>
__global__ __launch_bounds__(kNodeThreads) void level_hash_forest_kernel(const ForestLevelJob *__restrict__ jobs,
                                                                         u32 njobs)
{
    const ForestLevelJob &J = jobs[forest_find(jobs, njobs, blockIdx.x)];
    const u64 x = (((u64)blockIdx.x - J.block0) * kNodeThreads + threadIdx.x) / G;
    const bool ok = x < J.total;
    const ForestNode n = forest_node(J, ok ? x : 0);
    u64 *sd = J.d + n.s * J.D * 4;
    const u64 *ca = sd + n.c0 * 4, *cb = ca + 4;
    u64 *o = n.out < 0 ? J.caps + n.s * 4 : sd + (u64)n.out * 4;
    if constexpr (G == 1) {
        if (!ok) return;
        u64 l[4], r[4], h[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            l[k] = ca[k];
            r[k] = cb[k];
        }
        hash_two(l, r, h);
#pragma unroll
        for (int k = 0; k < 4; k++) o[k] = h[k];
    } else if constexpr (G == 4) {
        const u32 j = threadIdx.x & 3;
        u64 out[3];
        hash_two_quad(ca, cb, j, out);
        if (ok) {
            if (j == 0) {
                o[0] = out[0];
                o[1] = out[1];
                o[2] = out[2];
            } else if (j == 1) {
                o[3] = out[0];
            }
        }
    } else {
        static_assert(G == 16, "1, 4 or 16 lanes per node");
        const u32 j = threadIdx.x & 15, lb = threadIdx.x & 48;
        const u64 h = hash_two_hex(ca, cb, j, lb);
        if (ok && j < 4) o[j] = h;
    }
}

// The top of every subtree in one launch: workgroup s loads the 2m digests
// below its m-node level (m <= kTopNodes) into LDS, then computes the levels
// of m, m/2, ..., 2 nodes (16 lanes per node, levels kept in LDS and written
// to the digests) and the cap entry.  m = 1 (two leaves per subtree) is the
// cap alone.  A wave whose four groups are all past the level's last node
// skips it (the groups of a wave hash together), so the narrow levels run one
// wave on one SIMD.  The level of 64 nodes per subtree stays with the level
// kernels: 64 nodes x 16 lanes would put four waves on each SIMD of the CU.
constexpr u32 kTopNodes = 32;
__global__ __launch_bounds__(16 * kTopNodes) void tree_top_kernel(u64 *d, u64 *caps, u32 m, u64 D)
{
    __shared__ u64 bufA[2 * kTopNodes * 4];
    __shared__ u64 bufB[kTopNodes * 4];
    u64 *sd = d + (u64)blockIdx.x * D * 4;
    for (u32 i = threadIdx.x; i < 8 * m; i += blockDim.x) bufA[i] = sd[(u64)(2 * m - 2) * 4 + i];
    __syncthreads();
    const u32 j = threadIdx.x & 15, q = threadIdx.x >> 4, lb = threadIdx.x & 48;
    const u32 wave_q0 = q & ~3u;  // the wave's first group
    u64 *cur = bufA, *nxt = bufB;
    for (u32 nodes = m; nodes >= 2; nodes >>= 1) {
        if (wave_q0 < nodes) {  // uniform across the wave
            const u32 qc = q < nodes ? q : 0;
            const u64 h = hash_two_hex(cur + 8 * qc, cur + 8 * qc + 4, j, lb);
            if (q < nodes && j < 4) {
                sd[(u64)(nodes - 2 + q) * 4 + j] = h;
                nxt[4 * q + j] = h;
            }
        }
        __syncthreads();
        u64 *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    if (wave_q0 == 0) {
        const u64 h = hash_two_hex(cur, cur + 4, j, lb);
        if (q == 0 && j < 4) caps[(u64)blockIdx.x * 4 + j] = h;
    }
}

// ----------------------------------------------------------------- other hashers
// Keccak and Poseidon2 trees: one lane per leaf / node / cap entry.
template <LeafLayout LAYOUT, typename H>
__global__ __launch_bounds__(kLeafThreads) void leaf_hash_h_kernel(LeafArgs a)
{
    const u64 t = (u64)blockIdx.x * kLeafThreads + threadIdx.x;
    if (t >= a.n_leaves) return;
    u64 leaf, h[4];
    if constexpr (LAYOUT == LeafLayout::RowMajor) {
        leaf = t;
        const u64 *p = a.leaves + leaf * a.leaf_size;
        H::one([p](u32 i) { return p[i]; }, a.leaf_size, h);
    } else {
        leaf = zk_bitrev((u32)t, a.lg_leaves);
        const u64 *p = a.leaves + t;
        const u64 cs = a.col_stride;
        H::one([p, cs](u32 i) { return p[(u64)i * cs]; }, a.leaf_size, h);
    }
    u64 slot = leaf;
    if (a.sub_leaves) slot = (leaf / a.sub_leaves) * a.sub_digests + (a.sub_digests - a.sub_leaves) + leaf % a.sub_leaves;
    u64 *o = a.out + slot * 4;
#pragma unroll
    for (int k = 0; k < 4; k++) o[k] = h[k];
}

template <typename H>
__global__ __launch_bounds__(kNodeThreads) void level_hash_h_kernel(u64 *d, u64 nodes, u64 base, u64 n_sub, u64 D)
{
    const u64 t = (u64)blockIdx.x * kNodeThreads + threadIdx.x;
    if (t >= nodes * n_sub) return;
    const u64 s = t / nodes, node = base + t % nodes;
    u64 *sd = d + s * D * 4;
    u64 l[4], r[4], o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        l[k] = sd[(2 * node + 2) * 4 + k];
        r[k] = sd[(2 * node + 3) * 4 + k];
    }
    H::two(l, r, o);
#pragma unroll
    for (int k = 0; k < 4; k++) sd[node * 4 + k] = o[k];
}

template <typename H>
__global__ __launch_bounds__(kNodeThreads) void cap_hash_h_kernel(u64 *caps, const u64 *d, u64 n_caps, u64 D)
{
    const u64 s = (u64)blockIdx.x * kNodeThreads + threadIdx.x;
    if (s >= n_caps) return;
    const u64 *sd = d + s * D * 4;
    u64 l[4], r[4], o[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        l[k] = sd[k];
        r[k] = sd[4 + k];
    }
    H::two(l, r, o);
#pragma unroll
    for (int k = 0; k < 4; k++) caps[s * 4 + k] = o[k];
}

// Merkle openings: thread t = (query, layer k) copies the sibling of the
// query's leaf-to-root path at layer k (linear layout: position q has
// sibling q ^ 1 and parent q / 2 - 1).
__global__ __launch_bounds__(256) void proof_siblings_kernel(const u64 *digests, const u64 *indices, u64 n_idx,
                                                             u32 layers, u64 D, u64 *out)
{
    const u64 t = (u64)blockIdx.x * 256 + threadIdx.x;
    if (t >= n_idx * layers) return;
    const u64 qi = t / layers;
    const u32 k = (u32)(t % layers);
    const u64 l = indices[qi], L = 1ull << layers;
    const u64 *sd = digests + (l >> layers) * D * 4;
    u64 q = L - 2 + (l & (L - 1));
    for (u32 j = 0; j < k; j++) q = q / 2 - 1;
    const u64 *src = sd + (q ^ 1) * 4;
    u64 *dst = out + t * 4;
#pragma unroll
    for (int e = 0; e < 4; e++) dst[e] = src[e];
}

// Leaf rows of the queries: row-major leaves (col_stride 0) or a
// column-major LDE (leaf l = row rev(l)).
__global__ __launch_bounds__(256) void proof_rows_kernel(const u64 *leaves, u64 col_stride, u32 lg_leaves,
                                                         u64 leaf_size, const u64 *indices, u64 n_idx, u64 *out)
{
    const u64 t = (u64)blockIdx.x * 256 + threadIdx.x;
    if (t >= n_idx * leaf_size) return;
    const u64 qi = t / leaf_size, c = t % leaf_size, l = indices[qi];
    out[t] = col_stride ? leaves[c * col_stride + zk_bitrev((u32)l, lg_leaves)] : leaves[l * leaf_size + c];
}

static inline u32 blocks_for(u64 n, int tpb) { return (u32)((n + tpb - 1) / tpb); }

void merkle_launch(const TreeShape &t, u64 *digests, u64 *caps, const u64 *leaves, LeafLayout layout,
                   u64 col_stride, hipStream_t st, u64 *rows_out)
{
    LeafArgs a{};
    a.leaves = leaves;
    a.rows = layout == LeafLayout::ColMajorRev ? rows_out : nullptr;
    a.n_leaves = t.n_leaves;
    a.leaf_size = (u32)t.leaf_size;
    a.lg_leaves = ilog2(t.n_leaves);
    a.col_stride = col_stride;
    const u64 L = t.sub_leaves(), D = t.sub_digests(), nc = t.n_caps();
    if (nc == t.n_leaves) {
        a.out = caps;
        a.sub_leaves = 0;
    } else if (L == 2) {
        a.out = digests;
        a.sub_leaves = 0;
    } else {
        a.out = digests;
        a.sub_leaves = L;
        a.sub_digests = D;
    }
    const u32 nb = blocks_for(t.n_leaves, kLeafThreads);
    if (t.n_leaves <= kQuadMaxLeaves && t.leaf_size > 4) {
        ProfScope ps(layout == LeafLayout::RowMajor ? "merkle.leaves" : "merkle.leaves_colmajor", st);
        const u32 nq = blocks_for(4 * t.n_leaves, kLeafThreads);
        if (layout == LeafLayout::RowMajor)
            hipLaunchKernelGGL((leaf_hash_quad_kernel<LeafLayout::RowMajor, false>), dim3(nq), dim3(kLeafThreads), 0,
                               st, a);
        else if (a.rows)
            hipLaunchKernelGGL((leaf_hash_quad_kernel<LeafLayout::ColMajorRev, true>), dim3(nq), dim3(kLeafThreads),
                               0, st, a);
        else
            hipLaunchKernelGGL((leaf_hash_quad_kernel<LeafLayout::ColMajorRev, false>), dim3(nq),
                               dim3(kLeafThreads), 0, st, a);
        ZK_HIP(hipGetLastError());
    } else {
    ProfScope ps(layout == LeafLayout::RowMajor ? "merkle.leaves" : "merkle.leaves_colmajor", st);
    if (layout == LeafLayout::RowMajor)
        hipLaunchKernelGGL((leaf_hash_kernel<LeafLayout::RowMajor, false>), dim3(nb), dim3(kLeafThreads), 0, st, a);
    else if (a.rows)
        hipLaunchKernelGGL((leaf_hash_kernel<LeafLayout::ColMajorRev, true>), dim3(nb), dim3(kLeafThreads), 0, st, a);
    else
        hipLaunchKernelGGL((leaf_hash_kernel<LeafLayout::ColMajorRev, false>), dim3(nb), dim3(kLeafThreads), 0, st,
                           a);
    ZK_HIP(hipGetLastError());
    }
    if (nc == t.n_leaves) return;
    // Levels wider than the top kernel's: one lane per node while the level
    // fills the GPU, a quad per node below that.  Level with r nodes per
    // subtree starts at digest r - 2.
    const u64 m = std::min<u64>(L / 2, kTopNodes);
    for (u64 nodes = L / 2; nodes > m; nodes /= 2) {
        const u64 total = nodes * nc;
        ProfScope ps("merkle.levels", st);
        if (total > kQuadMaxNodes)
            hipLaunchKernelGGL(level_hash_kernel, dim3(blocks_for(total, kNodeThreads)), dim3(kNodeThreads), 0, st,
                               digests, nodes, nodes - 2, nc, D);
        else if (total > kHexMaxNodes)
            hipLaunchKernelGGL(level_hash_quad_kernel, dim3(blocks_for(4 * total, kNodeThreads)), dim3(kNodeThreads),
                               0, st, digests, nodes, nodes - 2, nc, D);
        else
            hipLaunchKernelGGL(level_hash_hex_kernel, dim3(blocks_for(16 * total, kNodeThreads)), dim3(kNodeThreads),
                               0, st, digests, nodes, nodes - 2, nc, D);
        ZK_HIP(hipGetLastError());
    }
    ProfScope ps("merkle.top", st);
    const u32 threads = 16 * (u32)std::max<u64>(m, 4);
    hipLaunchKernelGGL(tree_top_kernel, dim3((u32)nc), dim3(threads), 0, st, digests, caps, (u32)m, D);
    ZK_HIP(hipGetLastError());
}

// ------------------------------------------------------------ batched forest
struct ForestPlan {
    void *dev = nullptr;  // every launch's table
    std::vector<ForestLaunch> launches;
    ~ForestPlan()
    {
        if (dev) (void)hipFree(dev);
    }
};

void ForestPlanDeleter::operator()(ForestPlan *p) const { delete p; }

ForestPlanPtr forest_plan(const std::vector<ForestJob> &jobs)
{
    std::vector<unsigned char> bytes;
    ForestPlanPtr p(new ForestPlan());
    if (!forest_tables(jobs, bytes, p->launches)) throw Error((int)hipErrorInvalidValue, "invalid forest job");
    if (!bytes.empty()) {
        ZK_HIP(hipMalloc(&p->dev, bytes.size()));
        ZK_HIP(hipMemcpy(p->dev, bytes.data(), bytes.size(), hipMemcpyHostToDevice));
    }
    return p;
}

void forest_launch(const ForestPlan &p, hipStream_t st) { forest_launch_tables(p.launches, p.dev, st); }

void forest_launch_tables(const std::vector<ForestLaunch> &launches, const void *tables, hipStream_t st)
{
    for (const ForestLaunch &l : launches) {
        const void *tab = (const char *)tables + l.off;
        if (l.kind != ForestLaunch::Level) {
            ProfScope ps("batch.leaves", st);
            if (l.kind == ForestLaunch::LeavesRows)
                hipLaunchKernelGGL(leaf_hash_forest_kernel<true>, dim3(l.blocks), dim3(kLeafThreads), 0, st,
                                   (const ForestLeafJob *)tab, l.njobs);
            else
                hipLaunchKernelGGL(leaf_hash_forest_kernel<false>, dim3(l.blocks), dim3(kLeafThreads), 0, st,
                                   (const ForestLeafJob *)tab, l.njobs);
        } else {
            ProfScope ps("batch.levels", st);
            const ForestLevelJob *t = (const ForestLevelJob *)tab;
            if (l.g == 1)
                hipLaunchKernelGGL(level_hash_forest_kernel<1>, dim3(l.blocks), dim3(kNodeThreads), 0, st, t,
                                   l.njobs);
            else if (l.g == 4)
                hipLaunchKernelGGL(level_hash_forest_kernel<4>, dim3(l.blocks), dim3(kNodeThreads), 0, st, t,
                                   l.njobs);
            else
                hipLaunchKernelGGL(level_hash_forest_kernel<16>, dim3(l.blocks), dim3(kNodeThreads), 0, st, t,
                                   l.njobs);
        }
        ZK_HIP(hipGetLastError());
    }
}

template <typename H>
static void merkle_launch_h(const TreeShape &t, u64 *digests, u64 *caps, const u64 *leaves, LeafLayout layout,
                            u64 col_stride, hipStream_t st)
{
    LeafArgs a{};
    a.leaves = leaves;
    a.n_leaves = t.n_leaves;
    a.leaf_size = (u32)t.leaf_size;
    a.lg_leaves = ilog2(t.n_leaves);
    a.col_stride = col_stride;
    const u64 L = t.sub_leaves(), D = t.sub_digests(), nc = t.n_caps();
    a.out = nc == t.n_leaves ? caps : digests;
    a.sub_leaves = (nc == t.n_leaves || L == 2) ? 0 : L;
    a.sub_digests = D;
    const u32 nb = blocks_for(t.n_leaves, kLeafThreads);
    if (layout == LeafLayout::RowMajor)
        hipLaunchKernelGGL((leaf_hash_h_kernel<LeafLayout::RowMajor, H>), dim3(nb), dim3(kLeafThreads), 0, st, a);
    else
        hipLaunchKernelGGL((leaf_hash_h_kernel<LeafLayout::ColMajorRev, H>), dim3(nb), dim3(kLeafThreads), 0, st, a);
    ZK_HIP(hipGetLastError());
    if (nc == t.n_leaves) return;
    for (u64 nodes = L / 2; nodes >= 2; nodes /= 2) {
        hipLaunchKernelGGL(level_hash_h_kernel<H>, dim3(blocks_for(nodes * nc, kNodeThreads)), dim3(kNodeThreads), 0,
                           st, digests, nodes, nodes - 2, nc, D);
        ZK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(cap_hash_h_kernel<H>, dim3(blocks_for(nc, kNodeThreads)), dim3(kNodeThreads), 0, st, caps,
                       (const u64 *)digests, nc, D);
    ZK_HIP(hipGetLastError());
}

void merkle_launch_hashed(const TreeShape &t, u64 *digests, u64 *caps, const u64 *leaves, LeafLayout layout,
                          u64 col_stride, hipStream_t st, u32 hash_type)
{
    switch (hash_type) {
    case 0: merkle_launch(t, digests, caps, leaves, layout, col_stride, st); break;
    case 1: merkle_launch_h<KeccakHasher>(t, digests, caps, leaves, layout, col_stride, st); break;
    case 3: merkle_launch_h<Poseidon2Hasher>(t, digests, caps, leaves, layout, col_stride, st); break;
    default: throw Error((int)hipErrorInvalidValue, "hash type " + std::to_string(hash_type) + " is not implemented");
    }
}

void proofs_launch(const TreeShape &t, const u64 *digests, const u64 *leaves, u64 col_stride, const u64 *dev_indices,
                   u64 n_idx, u64 *siblings_out, u64 *rows_out, hipStream_t st)
{
    const u32 layers = ilog2(t.n_leaves) - (u32)t.cap_height;
    if (layers && n_idx) {
        hipLaunchKernelGGL(proof_siblings_kernel, dim3(blocks_for(n_idx * layers, 256)), dim3(256), 0, st, digests,
                           dev_indices, n_idx, layers, t.sub_digests(), siblings_out);
        ZK_HIP(hipGetLastError());
    }
    if (rows_out && n_idx) {
        hipLaunchKernelGGL(proof_rows_kernel, dim3(blocks_for(n_idx * t.leaf_size, 256)), dim3(256), 0, st, leaves,
                           col_stride, ilog2(t.n_leaves), t.leaf_size, dev_indices, n_idx, rows_out);
        ZK_HIP(hipGetLastError());
    }
}

}  // namespace zk
