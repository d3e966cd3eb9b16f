// Batched commits: many LDE + Merkle instances of uneven size in one go.
//
// Every job's coset LDE is issued back to back on one stream, then the Merkle
// trees of all jobs are built level-synchronously: one launch hashes the
// leaves of every job (read straight from the column-major LDEs), then one
// launch per tree height processes that height of every job, so the latency
// of the top levels is paid once per batch instead of once per job.  With
// ZK_BATCH_GRAPH the whole sequence is captured once into a hipGraph per
// distinct job list and replayed; the graph contains kernels only.
#include <string.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "batch.hpp"
#include "forest.hpp"
#include "ntt.hpp"

namespace zk {

struct BatchExec {
    std::unique_ptr<ForestPlan> forest;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    ~BatchExec()
    {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
    }
};

struct BatchState {
    hipStream_t main = nullptr;
    std::map<std::string, std::unique_ptr<BatchExec>> cache;
};

static std::mutex g_batch_mu;
static std::map<int, std::unique_ptr<BatchState>> g_batch;

static BatchState &batch_state(DeviceCtx &c)
{
    std::lock_guard<std::mutex> lk(g_batch_mu);
    auto &b = g_batch[c.dev];
    if (!b) {
        b = std::make_unique<BatchState>();
        ZK_HIP(hipStreamCreateWithFlags(&b->main, hipStreamNonBlocking));
    }
    return *b;
}

static void validate(const BatchJob &j)
{
    const u32 lg = j.lg_n + j.rate_bits;
    if (j.lg_n == 0 || lg > 31 || j.ncols == 0 || !j.coeffs || !j.lde || !j.cap || j.cap_height > lg)
        throw Error((int)hipErrorInvalidValue, "invalid batch job");
    if (j.cap_height < lg && !j.digests) throw Error((int)hipErrorInvalidValue, "invalid batch job");
}

static std::vector<ForestSpec> forest_specs(const BatchJob *jobs, u32 n)
{
    std::vector<ForestSpec> v(n);
    for (u32 i = 0; i < n; i++) {
        const BatchJob &j = jobs[i];
        const u64 N = 1ull << (j.lg_n + j.rate_bits);
        v[i].leaves = j.lde;
        v[i].col_stride = N;
        v[i].digests = j.digests;
        v[i].caps = j.cap;
        v[i].shape = tree_shape(2 * (N - (1ull << j.cap_height)), 1ull << j.cap_height, N, j.ncols, j.cap_height);
    }
    return v;
}

// Issues the whole batch on b.main.
static void issue(DeviceCtx &c, BatchState &b, const BatchJob *jobs, u32 n, const ForestPlan &fp)
{
    for (u32 i = 0; i < n; i++) {
        const BatchJob &j = jobs[i];
        hipStream_t st = b.main;
        NttJob nj;
        nj.src = j.coeffs;
        nj.dst = j.lde;
        nj.src_stride = 1ull << j.lg_n;
        nj.dst_stride = 1ull << (j.lg_n + j.rate_bits);
        nj.ncols = j.ncols;
        nj.lg = j.lg_n + j.rate_bits;
        nj.lg_in = j.lg_n;
        nj.coset_in = true;
        nj.prof_tag = "lde";
        ntt_launch(c, nj, st);
        if (j.leaves) transpose_rev_launch(j.lde, j.leaves, nj.lg, j.ncols, st);
    }
    forest_launch(fp, b.main);
}

void batch_commit(DeviceCtx &c, const BatchJob *jobs, u32 n, bool use_graph)
{
    if (n == 0) return;
    for (u32 i = 0; i < n; i++) validate(jobs[i]);
    BatchState &b = batch_state(c);
    // caller's prior work on the library stream / null stream completes first
    ZK_HIP(hipStreamSynchronize(c.stream));
    const std::string key(reinterpret_cast<const char *>(jobs), sizeof(BatchJob) * n);
    auto it = b.cache.find(key);
    if (it == b.cache.end()) {
        auto e = std::make_unique<BatchExec>();
        e->forest.reset(forest_plan(forest_specs(jobs, n), c.dev));
        it = b.cache.emplace(key, std::move(e)).first;
        // first use runs eagerly (sets kernel attributes, validates the plan)
        issue(c, b, jobs, n, *it->second->forest);
        ZK_HIP(hipStreamSynchronize(b.main));
        return;
    }
    BatchExec &e = *it->second;
    if (use_graph) {
        if (!e.exec) {
            const bool prof = prof_enabled();
            prof_set(false);  // no timing events inside a captured graph
            try {
                ZK_HIP(hipStreamBeginCapture(b.main, hipStreamCaptureModeThreadLocal));
                issue(c, b, jobs, n, *e.forest);
                ZK_HIP(hipStreamEndCapture(b.main, &e.graph));
                ZK_HIP(hipGraphInstantiate(&e.exec, e.graph, nullptr, nullptr, 0));
            } catch (...) {
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(b.main, &g);
                if (g) (void)hipGraphDestroy(g);
                prof_set(prof);
                throw;
            }
            prof_set(prof);
        }
        ProfScope ps("batch.graph", b.main);
        ZK_HIP(hipGraphLaunch(e.exec, b.main));
    } else {
        issue(c, b, jobs, n, *e.forest);
    }
    ZK_HIP(hipStreamSynchronize(b.main));
}

void batch_clear(DeviceCtx &c)
{
    BatchState &b = batch_state(c);
    ZK_HIP(hipStreamSynchronize(b.main));
    b.cache.clear();
}

}  // namespace zk
