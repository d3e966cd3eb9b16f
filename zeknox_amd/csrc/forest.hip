// Multi-job Poseidon Merkle forests in one persistent launch.
//
// The work of every tree of every job (leaf hashes, each internal level, the
// cap) is cut into items of up to 256 nodes, ordered by height and then job,
// and handed out by one atomic queue head to a persistent grid.  An item at
// height t > 0 waits until every item of the same job at height t-1 is done.
// Because items are claimed in that order, everything an item waits for was
// claimed earlier by a running block that never waits on later items, so the
// queue always drains; the waits are bounded anyway (a timeout sets an error
// flag and the block moves on, so the grid always exits).
//
// Per subtree of L leaves (linear layout, D = 2L - 2 digests): the 2^r nodes
// of one level start at digest index 2^r - 2 and node i = H(d[2i+2], d[2i+3])
// (okx/zeknox native/merkle/merkle.cu:116-163); leaves occupy [L-2, 2L-2).
#include <stdio.h>

#include "forest.hpp"
#include "poseidon.hpp"

namespace zk {

constexpr int kForestThreads = 256;
constexpr int kMaxHeights = 40;
constexpr u32 kSpinLimit = 1u << 24;

struct ForestJob {
    const u64 *leaves;
    u64 col_stride;
    u64 *digests;
    u64 *caps;
    u64 n_leaves;
    u32 leaf_size, lg_leaves, cap_height, lg_sub;
    u32 mode;  // 0: caps == leaves, 1: two leaves per subtree, 2: general
    u32 top;   // height of the cap items (0 for mode 0)
    u32 item_count[kMaxHeights];
};

ForestPlan::~ForestPlan()
{
    if (dev_jobs) (void)hipFree(dev_jobs);
    if (dev_items) (void)hipFree(dev_items);
    if (dev_sync) (void)hipFree(dev_sync);
}

__global__ void forest_reset_kernel(u32 *sync, u32 n)
{
    for (u32 i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) sync[i] = 0;
}

// Children written by other workgroups (possibly on another XCD) of the same
// launch: device-coherent loads.
ZK_HD void load_digest_coherent(const u64 *p, u64 out[4])
{
#pragma unroll
    for (int k = 0; k < 4; k++) {
#ifdef __HIP_DEVICE_COMPILE__
        out[k] = __hip_atomic_load(p + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#else
        out[k] = p[k];
#endif
    }
}

// The work of thread `tid` on item (job J, height t, chunk): one node each.
ZK_HD void forest_thread(const ForestJob &J, u32 t, u32 chunk, u32 tid)
{
    const u64 x = (u64)chunk * kForestThreads + tid;
    const u64 ncaps = 1ull << J.cap_height;
    if (t == 0) {
        if (x < J.n_leaves) {
            u64 h[4];
            u64 leaf;
            if (J.col_stride) {
                leaf = zk_bitrev((u32)x, J.lg_leaves);  // lane x owns LDE row x
                const u64 *p = J.leaves + x;
                const u64 cs = J.col_stride;
                hash_one([p, cs](u32 i) { return p[(u64)i * cs]; }, J.leaf_size, h);
            } else {
                leaf = x;
                const u64 *p = J.leaves + x * J.leaf_size;
                hash_one([p](u32 i) { return p[i]; }, J.leaf_size, h);
            }
            u64 *o;
            if (J.mode == 0) {
                o = J.caps + leaf * 4;
            } else if (J.mode == 1) {
                o = J.digests + leaf * 4;
            } else {
                const u64 L = 1ull << J.lg_sub, D = 2 * L - 2;
                o = J.digests + ((leaf >> J.lg_sub) * D + (L - 2) + (leaf & (L - 1))) * 4;
            }
#pragma unroll
            for (int k = 0; k < 4; k++) o[k] = h[k];
        }
    } else if (t < J.top) {
        // internal level: 2^r nodes per subtree, r = lg_sub - t
        const u32 r = J.lg_sub - t;
        if (x < (ncaps << r)) {
            const u64 D = (2ull << J.lg_sub) - 2;
            const u64 s = x >> r, node = ((1ull << r) - 2) + (x & ((1ull << r) - 1));
            u64 *sd = J.digests + s * D * 4;
            u64 l[4], rr[4], o[4];
            load_digest_coherent(sd + (2 * node + 2) * 4, l);
            load_digest_coherent(sd + (2 * node + 3) * 4, rr);
            hash_two(l, rr, o);
#pragma unroll
            for (int k = 0; k < 4; k++) sd[node * 4 + k] = o[k];
        }
    } else if (x < ncaps) {
        const u64 D = (2ull << J.lg_sub) - 2;
        const u64 *sd = J.digests + x * D * 4;
        u64 l[4], rr[4], o[4];
        load_digest_coherent(sd, l);
        load_digest_coherent(sd + 4, rr);
        hash_two(l, rr, o);
#pragma unroll
        for (int k = 0; k < 4; k++) J.caps[x * 4 + k] = o[k];
    }
}

__global__ __launch_bounds__(kForestThreads) void forest_kernel(const ForestJob *__restrict__ jobs,
                                                                const uint2 *__restrict__ items, u32 nitems,
                                                                u32 *sync)
{
    __shared__ u32 s_item;
    u32 *head = sync, *err = sync + 1, *done = sync + 2;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(head, 1u);
        __syncthreads();
        const u32 it = s_item;
        __syncthreads();
        if (it >= nitems) return;
        const uint2 d = items[it];
        const u32 ji = d.x >> 8, t = d.x & 0xFF, chunk = d.y;
        const ForestJob &J = jobs[ji];
        if (t > 0) {
            if (threadIdx.x == 0) {
                const u32 need = J.item_count[t - 1];
                u32 *cnt = done + ji * kMaxHeights + (t - 1);
                u32 spins = 0;
                while (__hip_atomic_load(cnt, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < need) {
                    __builtin_amdgcn_s_sleep(4);
                    if (++spins > kSpinLimit) {
                        atomicOr(err, 1u);
                        break;
                    }
                }
            }
            __syncthreads();
        }
        forest_thread(J, t, chunk, threadIdx.x);
        __syncthreads();
        if (threadIdx.x == 0)
            __hip_atomic_fetch_add(done + ji * kMaxHeights + t, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Level-synchronous form: one launch per height, block b runs item first + b.
// Every item of the launch has its inputs written by earlier launches of the
// same stream, so there are no device-side waits or atomics.
__global__ __launch_bounds__(kForestThreads) void forest_height_kernel(const ForestJob *__restrict__ jobs,
                                                                       const uint2 *__restrict__ items, u32 first)
{
    const uint2 d = items[first + blockIdx.x];
    forest_thread(jobs[d.x >> 8], d.x & 0xFF, d.y, threadIdx.x);
}

static u32 ceil_div(u64 a, u64 b) { return (u32)((a + b - 1) / b); }

// Host side of a plan: job descriptors and the ordered item list.
static void forest_host_plan(const std::vector<ForestSpec> &specs, std::vector<ForestJob> &jobs,
                             std::vector<uint2> &items, u32 &maxT_out)
{
    jobs.assign(specs.size(), ForestJob{});
    u32 maxT = 0;
    for (size_t j = 0; j < specs.size(); j++) {
        const TreeShape &t = specs[j].shape;
        ForestJob &J = jobs[j];
        J = ForestJob{};
        J.leaves = specs[j].leaves;
        J.col_stride = specs[j].col_stride;
        J.digests = specs[j].digests;
        J.caps = specs[j].caps;
        J.n_leaves = t.n_leaves;
        J.leaf_size = (u32)t.leaf_size;
        u32 lg = 0;
        while ((1ull << lg) < t.n_leaves) lg++;
        J.lg_leaves = lg;
        J.cap_height = (u32)t.cap_height;
        J.lg_sub = lg - J.cap_height;
        if (J.col_stride && lg > 32) throw Error((int)hipErrorInvalidValue, "too many leaves");
        J.item_count[0] = ceil_div(t.n_leaves, kForestThreads);
        if (t.n_caps() == t.n_leaves) {
            J.mode = 0;
            J.top = 0;
        } else {
            J.mode = J.lg_sub == 1 ? 1 : 2;
            J.top = J.lg_sub;  // heights 1 .. lg_sub-1 internal, lg_sub caps
            for (u32 h = 1; h < J.top; h++) J.item_count[h] = ceil_div(t.n_caps() << (J.lg_sub - h), kForestThreads);
            J.item_count[J.top] = ceil_div(t.n_caps(), kForestThreads);
        }
        if (J.top + 1 > (u32)kMaxHeights) throw Error((int)hipErrorInvalidValue, "tree too deep");
        maxT = std::max(maxT, J.top);
    }
    items.clear();
    for (u32 h = 0; h <= maxT; h++)
        for (size_t j = 0; j < jobs.size(); j++) {
            if (h > jobs[j].top) continue;
            for (u32 c = 0; c < jobs[j].item_count[h]; c++) items.push_back(make_uint2(((u32)j << 8) | h, c));
        }
    maxT_out = maxT;
}

ForestPlan *forest_plan(const std::vector<ForestSpec> &specs, int device)
{
    std::vector<ForestJob> jobs;
    std::vector<uint2> items;
    u32 maxT = 0;
    forest_host_plan(specs, jobs, items, maxT);
    auto *p = new ForestPlan();
    p->height_first.assign(maxT + 2, 0);
    for (size_t i = 0; i < items.size(); i++) p->height_first[(items[i].x & 0xFF) + 1] = (u32)(i + 1);
    for (u32 h = 1; h <= maxT + 1; h++) p->height_first[h] = std::max(p->height_first[h], p->height_first[h - 1]);
    p->njobs = (u32)jobs.size();
    p->nitems = (u32)items.size();
    p->nheights = maxT + 1;
    ZK_HIP(hipMalloc(&p->dev_jobs, jobs.size() * sizeof(ForestJob)));
    ZK_HIP(hipMemcpy(p->dev_jobs, jobs.data(), jobs.size() * sizeof(ForestJob), hipMemcpyHostToDevice));
    ZK_HIP(hipMalloc(&p->dev_items, std::max<size_t>(items.size(), 1) * sizeof(uint2)));
    if (!items.empty())
        ZK_HIP(hipMemcpy(p->dev_items, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
    ZK_HIP(hipMalloc(&p->dev_sync, (2 + jobs.size() * kMaxHeights) * sizeof(u32)));
    (void)device;
    p->grid = 0;  // persistent form only
    return p;
}

void forest_launch(const ForestPlan &p, hipStream_t st)
{
    if (!p.nitems) return;
    ProfScope ps("merkle.forest", st);
    for (u32 h = 0; h + 1 < (u32)p.height_first.size(); h++) {
        const u32 first = p.height_first[h], count = p.height_first[h + 1] - first;
        if (!count) continue;
        hipLaunchKernelGGL(forest_height_kernel, dim3(count), dim3(kForestThreads), 0, st,
                           (const ForestJob *)p.dev_jobs, (const uint2 *)p.dev_items, first);
        ZK_HIP(hipGetLastError());
    }
}

// The persistent single-launch form (forest_kernel with the atomic work queue)
// is kept in the source but not launched: its first GPU runs ended in a
// memory fault whose cause was not found (see the design notes).
void forest_launch_persistent(const ForestPlan &p, hipStream_t st)
{
    (void)p;
    (void)st;
    throw Error((int)hipErrorNotSupported, "persistent Merkle forest launch is disabled");
}

void forest_check(const ForestPlan &p, hipStream_t st)
{
    u32 err = 0;
    ZK_HIP(hipMemcpyAsync(&err, p.dev_sync + 1, sizeof(u32), hipMemcpyDeviceToHost, st));
    ZK_HIP(hipStreamSynchronize(st));
    if (err) throw Error(-(int)hipErrorLaunchTimeOut, "Merkle forest: a dependency wait timed out");
}

}  // namespace zk
