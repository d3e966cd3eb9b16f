// Host model of the persistent multi-job Merkle forest that an early batched
// commit ran in one launch (its work-queue protocol, restated here from that
// design; the product no longer has it): items of up to 256 nodes, ordered by
// tree height and then job, are claimed through one atomic queue head by W
// "workgroups" (host threads); an item at height t > 0 waits until every item
// of its job at height t - 1 is done (per-job, per-height done counters,
// acquire / release); a wait that exceeds its spin bound either proceeds
// (the design as it ran) or stops the workgroup with an error (the design a
// replacement must have).  Every digest, leaf and cap address is bounds-checked
// against its buffer, every child read is checked to have been written first,
// and the caps and digests are compared with a one-lane reference forest.
//
// usage: forest_queue_model workgroups spin_limit proceed_on_timeout
// Prints one line per job list and "ok" when every check passed.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>

#include "../../zeknox_amd/csrc/poseidon.hpp"

using namespace zk;

constexpr int kItem = 256;  // nodes per item (threads of a workgroup)

struct Job {
    u32 lg_leaves, leaf_size, cap_height;
    std::vector<u64> lde;        // column-major, leaf l = row rev(l)
    std::vector<u64> digests;    // 2 (N - 2^h) x 4
    std::vector<u64> caps;       // 2^h x 4
    std::vector<std::atomic<int>> written;  // per digest slot, then per cap slot
    u32 top = 0;                 // height of the cap items (0: caps == leaves)
    std::vector<u32> items_at;   // items per height
    std::vector<std::atomic<u32>> done;
    Job(u32 lg, u32 c, u32 h) : lg_leaves(lg), leaf_size(c), cap_height(h) {}
};

// out-of-bounds addresses, and reads of a child digest no item has written yet
static std::atomic<int> violations{0}, oob{0}, unwritten{0};
static void violation(const char *what, u64 a, u64 b, bool bounds = true)
{
    (bounds ? oob : unwritten)++;
    if (violations++ < 10) printf("violation: %s (%llu, %llu)\n", what, (unsigned long long)a, (unsigned long long)b);
}

static u64 *at(std::vector<u64> &v, u64 i, const char *what)
{
    if (i + 4 > v.size()) {
        violation(what, i, v.size());
        static u64 sink[4];
        return sink;
    }
    return &v[i];
}

// one node of an item (the per-thread work of the kernel)
static void item_thread(Job &J, u32 t, u64 x)
{
    const u64 N = 1ull << J.lg_leaves, ncaps = 1ull << J.cap_height;
    const u32 lg_sub = J.lg_leaves - J.cap_height;
    const u64 L = 1ull << lg_sub, D = 2 * L - 2;
    const u64 nd = J.digests.size() / 4;
    if (t == 0) {
        if (x >= N) return;
        const u64 leaf = zk_bitrev((u32)x, J.lg_leaves);
        u64 h[4];
        const u64 *p = J.lde.data() + x;
        const u64 cs = N;
        if ((u64)(J.leaf_size - 1) * cs + x >= J.lde.size()) violation("leaf read", x, J.lde.size());
        hash_one([p, cs](u32 i) { return p[(u64)i * cs]; }, J.leaf_size, h);
        u64 slot;
        u64 *o;
        if (ncaps == N) {
            slot = nd + leaf;
            o = at(J.caps, leaf * 4, "cap write");
        } else if (lg_sub == 1) {
            slot = leaf;
            o = at(J.digests, leaf * 4, "leaf digest write");
        } else {
            slot = (leaf >> lg_sub) * D + (L - 2) + (leaf & (L - 1));
            o = at(J.digests, slot * 4, "leaf digest write");
        }
        for (int k = 0; k < 4; k++) o[k] = h[k];
        if (slot < J.written.size()) J.written[slot].store(1, std::memory_order_relaxed);
    } else if (t < J.top) {
        const u32 r = lg_sub - t;
        if (x >= (ncaps << r)) return;
        const u64 s = x >> r, node = ((1ull << r) - 2) + (x & ((1ull << r) - 1));
        const u64 c0 = s * D + 2 * node + 2, c1 = c0 + 1, me = s * D + node;
        for (u64 c : {c0, c1})
            if (c >= nd || !J.written[c].load(std::memory_order_relaxed)) violation("unwritten child read", c, nd, c >= nd);
        u64 a[4], b[4], o[4];
        const u64 *pa = at(J.digests, c0 * 4, "child read"), *pb = at(J.digests, c1 * 4, "child read");
        for (int k = 0; k < 4; k++) a[k] = pa[k], b[k] = pb[k];
        hash_two(a, b, o);
        u64 *w = at(J.digests, me * 4, "node write");
        for (int k = 0; k < 4; k++) w[k] = o[k];
        J.written[me].store(1, std::memory_order_relaxed);
    } else {
        if (x >= ncaps) return;
        const u64 c0 = x * D, c1 = c0 + 1;
        for (u64 c : {c0, c1})
            if (c >= nd || !J.written[c].load(std::memory_order_relaxed)) violation("unwritten cap child read", c, nd, c >= nd);
        u64 a[4], b[4], o[4];
        const u64 *pa = at(J.digests, c0 * 4, "cap child"), *pb = at(J.digests, c1 * 4, "cap child");
        for (int k = 0; k < 4; k++) a[k] = pa[k], b[k] = pb[k];
        hash_two(a, b, o);
        u64 *w = at(J.caps, x * 4, "cap write");
        for (int k = 0; k < 4; k++) w[k] = o[k];
    }
}

struct Item {
    u32 job, t, chunk;
};

static bool run_list(std::vector<std::unique_ptr<Job>> &jobs, int W, long spin_limit, bool proceed)
{
    std::vector<Item> items;
    u32 maxT = 0;
    for (auto &jp : jobs) {
        Job &J = *jp;
        const u64 N = 1ull << J.lg_leaves, ncaps = 1ull << J.cap_height;
        const u32 lg_sub = J.lg_leaves - J.cap_height;
        J.items_at.assign(lg_sub + 1, 0);
        J.items_at[0] = (u32)((N + kItem - 1) / kItem);
        if (ncaps == N) {
            J.top = 0;
        } else {
            J.top = lg_sub;
            for (u32 h = 1; h < J.top; h++) J.items_at[h] = (u32)(((ncaps << (lg_sub - h)) + kItem - 1) / kItem);
            J.items_at[J.top] = (u32)((ncaps + kItem - 1) / kItem);
        }
        J.done = std::vector<std::atomic<u32>>(J.top + 1);
        maxT = std::max(maxT, J.top);
    }
    for (u32 h = 0; h <= maxT; h++)
        for (u32 j = 0; j < jobs.size(); j++)
            if (h <= jobs[j]->top)
                for (u32 c = 0; c < jobs[j]->items_at[h]; c++) items.push_back({j, h, c});
    std::atomic<u32> head{0};
    std::atomic<int> err{0};
    auto workgroup = [&]() {
        for (;;) {
            const u32 it = head.fetch_add(1);
            if (it >= items.size()) return;
            const Item d = items[it];
            Job &J = *jobs[d.job];
            if (d.t > 0) {
                long spins = 0;
                bool timed_out = false;
                while (J.done[d.t - 1].load(std::memory_order_acquire) < J.items_at[d.t - 1]) {
                    std::this_thread::yield();
                    if (++spins > spin_limit) {
                        timed_out = true;
                        break;
                    }
                }
                if (timed_out) {
                    err.store(1);
                    if (!proceed) return;  // the replacement's rule: never run on missing inputs
                }
            }
            for (u32 tid = 0; tid < kItem; tid++) item_thread(J, d.t, (u64)d.chunk * kItem + tid);
            J.done[d.t].fetch_add(1, std::memory_order_release);
        }
    };
    std::vector<std::thread> ws;
    for (int w = 0; w < W; w++) ws.emplace_back(workgroup);
    for (auto &w : ws) w.join();
    return err.load() == 0;
}

// one-lane reference forest of a job
static void reference(Job &J, std::vector<u64> &dg, std::vector<u64> &cp)
{
    const u64 N = 1ull << J.lg_leaves, ncaps = 1ull << J.cap_height;
    const u32 lg_sub = J.lg_leaves - J.cap_height;
    const u64 L = 1ull << lg_sub, D = 2 * L - 2;
    dg.assign(J.digests.size(), 0);
    cp.assign(J.caps.size(), 0);
    std::vector<u64> lh(N * 4);
    for (u64 x = 0; x < N; x++) {
        const u64 leaf = zk_bitrev((u32)x, J.lg_leaves);
        const u64 *p = J.lde.data() + x;
        hash_one([p, N](u32 i) { return p[(u64)i * N]; }, J.leaf_size, &lh[leaf * 4]);
    }
    if (ncaps == N) {
        cp = lh;
        return;
    }
    for (u64 s = 0; s < ncaps; s++) {
        u64 *sd = dg.data() + s * D * 4;
        for (u64 i = 0; i < L; i++)
            for (int k = 0; k < 4; k++) sd[(D - L + i) * 4 + k] = lh[(s * L + i) * 4 + k];
        for (u64 node = D - L; node-- > 0;) hash_two(&sd[(2 * node + 2) * 4], &sd[(2 * node + 3) * 4], &sd[node * 4]);
        hash_two(&sd[0], &sd[4], &cp[s * 4]);
    }
}

int main(int argc, char **argv)
{
    const int W = argc > 1 ? atoi(argv[1]) : 8;
    const long spin = argc > 2 ? atol(argv[2]) : 1L << 30;
    const bool proceed = argc > 3 && atoi(argv[3]) != 0;
    // (lg_n, rate_bits, columns, cap_height): the small job list the persistent
    // launch faulted on, and a mixed list
    const std::vector<std::vector<int>> lists[] = {
        {{3, 1, 5, 2}, {4, 2, 7, 3}, {6, 3, 16, 4}, {5, 1, 1, 0}, {7, 2, 9, 6}, {2, 1, 3, 3}, {8, 1, 12, 1},
         {9, 3, 4, 0}, {6, 2, 135, 4}},
        {{10, 1, 33, 4}, {8, 3, 9, 2}, {11, 2, 70, 4}, {12, 1, 4, 3}, {9, 3, 135, 4}, {13, 1, 8, 4}}};
    std::mt19937_64 g(7);
    bool all_ok = true;
    int li = 0;
    for (const auto &list : lists) {
        std::vector<std::unique_ptr<Job>> jobs;
        for (const auto &s : list) {
            const u32 lg = (u32)(s[0] + s[1]);
            auto J = std::make_unique<Job>(lg, (u32)s[2], (u32)s[3]);
            const u64 N = 1ull << lg, ncaps = 1ull << s[3];
            J->lde.resize(N * s[2]);
            for (auto &v : J->lde) v = g() % ZK_P;
            const u64 nd = ncaps < N ? 2 * (N - ncaps) : 0;
            J->digests.assign(nd * 4, 0);
            J->caps.assign(ncaps * 4, 0);
            J->written = std::vector<std::atomic<int>>(nd + ncaps);
            jobs.push_back(std::move(J));
        }
        const int v0 = violations.load(), o0 = oob.load(), u0 = unwritten.load();
        const bool ok = run_list(jobs, W, spin, proceed);
        int wrong = 0;
        for (auto &jp : jobs) {
            std::vector<u64> dg, cp;
            reference(*jp, dg, cp);
            if (dg != jp->digests || cp != jp->caps) wrong++;
        }
        printf("list %d: %zu jobs, %s, %d out-of-bounds, %d unwritten reads, %d jobs with wrong results\n", li++,
               jobs.size(), ok ? "no wait timed out" : "a wait timed out", oob.load() - o0, unwritten.load() - u0,
               wrong);
        all_ok = all_ok && ok && wrong == 0 && violations.load() == v0;
    }
    printf("%s\n", all_ok ? "ok" : "FAIL");
    return all_ok ? 0 : 1;
}
