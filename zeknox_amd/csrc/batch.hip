// Batched commits: many LDE + Merkle instances of uneven size in one call.
//
// The coset LDEs of the jobs run on kLanes side streams (longest first, each
// onto the least-loaded lane) forked from and joined back into one origin
// stream with events, so the small transforms overlap.  Then the Merkle
// forests of all jobs run on the origin stream as fused launches
// (forest_launch, merkle.hip): one leaf-hash launch over the leaves of every
// job and one launch per tree height over every job's trees, so the leaf
// hashing of small commits, which cannot fill the GPU one commit at a time,
// is part of one launch that does, and the latency of a narrow tree level is
// paid once per height instead of once per job.  There are no device-side
// waits: every launch reads only what earlier launches on its stream wrote.
// With ZK_BATCH_GRAPH, the second call with the same job list captures the
// whole DAG into a hipGraph (kernels and event edges only) and later calls
// replay it.  At most kMaxGraphs job lists are remembered per device (least
// recently used first out), so a caller that passes fresh buffers on every
// call does not grow the cache.
#include <string.h>

#include <algorithm>
#include <list>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "commit.hpp"
#include "merkle.hpp"

namespace zk {

constexpr int kLanes = 4;
constexpr size_t kMaxGraphs = 8;

struct BatchExec {
    ForestPlanPtr forest;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchExec()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct BatchState {
    hipStream_t origin = nullptr;
    hipStream_t lanes[kLanes] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[kLanes] = {};
    // job list bytes -> plan, most recently used first
    std::list<std::pair<std::string, std::unique_ptr<BatchExec>>> cache;
};

static std::mutex g_batch_mu;
static std::map<int, std::unique_ptr<BatchState>> g_batch;

static BatchState &batch_state(DeviceCtx &c)
{
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto &b = g_batch[c.dev];
    if (!b) {
        auto n = std::make_unique<BatchState>();
        ZK_HIP(hipStreamCreateWithFlags(&n->origin, hipStreamNonBlocking));
        for (int l = 0; l < kLanes; l++) {
            ZK_HIP(hipStreamCreateWithFlags(&n->lanes[l], hipStreamNonBlocking));
            ZK_HIP(hipEventCreateWithFlags(&n->join[l], hipEventDisableTiming));
        }
        ZK_HIP(hipEventCreateWithFlags(&n->fork, hipEventDisableTiming));
        b = std::move(n);
    }
    return *b;
}

// Lane of every job: longest processing time first onto the least-loaded lane.
static std::vector<int> assign_lanes(const CommitJob *jobs, u32 n)
{
    std::vector<u32> order(n);
    for (u32 i = 0; i < n; i++) order[i] = i;
    std::vector<double> cost(n);
    for (u32 i = 0; i < n; i++) cost[i] = lde_cost_us(jobs[i]);
    std::stable_sort(order.begin(), order.end(), [&](u32 a, u32 b) { return cost[a] > cost[b]; });
    double load[kLanes] = {};
    std::vector<int> lane(n);
    for (u32 i : order) {
        int best = 0;
        for (int l = 1; l < kLanes; l++)
            if (load[l] < load[best]) best = l;
        lane[i] = best;
        load[best] += cost[i];
    }
    return lane;
}

static ForestPlanPtr plan_forests(const CommitJob *jobs, u32 n)
{
    std::vector<ForestJob> f(n);
    for (u32 i = 0; i < n; i++) {
        const CommitJob &j = jobs[i];
        f[i] = ForestJob{j.lde, j.leaves, j.digests, j.cap, j.lg_n + j.rate_bits, j.ncols, j.cap_height};
    }
    return forest_plan(f);
}

// Fork, per-lane LDEs, join, then the fused forests on the origin stream:
// everything ends up ordered before the origin stream's next operation.
static void issue(DeviceCtx &c, BatchState &b, const CommitJob *jobs, u32 n, const ForestPlan &forest)
{
    const std::vector<int> lane = assign_lanes(jobs, n);
    ZK_HIP(hipEventRecord(b.fork, b.origin));
    for (int l = 0; l < kLanes; l++) ZK_HIP(hipStreamWaitEvent(b.lanes[l], b.fork, 0));
    for (u32 i = 0; i < n; i++) lde_launch(c, jobs[i], b.lanes[lane[i]]);
    for (int l = 0; l < kLanes; l++) {
        ZK_HIP(hipEventRecord(b.join[l], b.lanes[l]));
        ZK_HIP(hipStreamWaitEvent(b.origin, b.join[l], 0));
    }
    forest_launch(forest, b.origin);
}

void batch_commit(DeviceCtx &c, const CommitJob *jobs, u32 n, bool use_graph)
{
    if (n == 0) return;
    for (u32 i = 0; i < n; i++) commit_validate(jobs[i]);
    BatchState &b = batch_state(c);
    // inputs written by the caller's earlier work (library or null stream)
    ZK_HIP(hipStreamSynchronize(c.stream));
    ZK_HIP(hipStreamSynchronize(nullptr));
    if (!use_graph) {
        const ForestPlanPtr forest = plan_forests(jobs, n);
        issue(c, b, jobs, n, *forest);
        ZK_HIP(hipStreamSynchronize(b.origin));
        return;
    }
    const std::string key(reinterpret_cast<const char *>(jobs), sizeof(CommitJob) * n);
    auto it = std::find_if(b.cache.begin(), b.cache.end(), [&](const auto &kv) { return kv.first == key; });
    if (it == b.cache.end()) {
        // first call with this job list: eager (sets kernel attributes); the
        // origin stream is idle here, so an evicted graph is not running
        auto e = std::make_unique<BatchExec>();
        e->forest = plan_forests(jobs, n);
        if (b.cache.size() >= kMaxGraphs) b.cache.pop_back();
        b.cache.emplace_front(key, std::move(e));
        issue(c, b, jobs, n, *b.cache.front().second->forest);
        ZK_HIP(hipStreamSynchronize(b.origin));
        return;
    }
    b.cache.splice(b.cache.begin(), b.cache, it);
    BatchExec &e = *b.cache.front().second;
    if (!e.exec) {
        const bool prof = prof_enabled();
        prof_set(false);  // no timing events inside a captured graph
        try {
            ZK_HIP(hipStreamBeginCapture(b.origin, hipStreamCaptureModeThreadLocal));
            issue(c, b, jobs, n, *e.forest);
            ZK_HIP(hipStreamEndCapture(b.origin, &e.graph));
            ZK_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
        } catch (...) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(b.origin, &g);
            if (g) (void)hipGraphDestroy(g);
            if (e.graph) (void)hipGraphDestroy(e.graph);
            e.graph = nullptr;
            prof_set(prof);
            throw;
        }
        prof_set(prof);
    }
    ProfScope ps("batch.graph", b.origin);
    ZK_HIP(hipGraphLaunch(e.exec, b.origin));
    ZK_HIP(hipStreamSynchronize(b.origin));
}

void batch_clear(DeviceCtx &c)
{
    BatchState &b = batch_state(c);
    ZK_HIP(hipStreamSynchronize(b.origin));
    b.cache.clear();
}

}  // namespace zk
