// Host test of the batched Merkle forest's launch plans (forest_tables and the
// index maps forest_find / forest_node of zeknox_amd/csrc/forest_map.hpp):
// every launch is emulated workgroup by workgroup and thread by thread, in
// launch order, with the kernels' index arithmetic, on job lists that include
// the shapes of the batched workload (lg 11..21 leaves, 4..135 columns,
// cap_height min(4, lg - 2)) and the edge cases (caps == leaves, two leaves per
// subtree, one leaf per cap, leaf rows on some jobs only).  Checks: every leaf
// of every job is hashed exactly once, into its linear-layout slot; every
// level node reads two children that an earlier launch wrote and writes its
// own slot exactly once; every digest and cap ends up written exactly once;
// no access leaves its buffer; tables are sorted with block0 starting at 0; no
// launch exceeds 2^32 threads.  Plan sizes only (no emulation) for a 2^31-leaf
// job.  Prints "ok".
#include <cstdio>
#include <random>
#include <vector>

#include "../../zeknox_amd/csrc/forest_map.hpp"

using namespace zk;

static int fails = 0;
#define CHECK(c, ...)                         \
    do {                                      \
        if (!(c)) {                           \
            if (fails++ < 20) {               \
                printf("FAIL: " __VA_ARGS__); \
                printf("\n");                 \
            }                                 \
        }                                     \
    } while (0)

struct Spec {
    u32 lg, C, h;
    bool rows;
};

struct Buf {
    std::vector<u64> dig, cap;        // markers written (slot id + 1)
    std::vector<int> dig_w, cap_w;    // write counts
    std::vector<int> leaf_hashed;
    u64 lde_dummy[1], rows_dummy[1];  // distinct pointers, never dereferenced
};

static bool in(const std::vector<u64> &v, const u64 *p) { return p >= v.data() && p + 4 <= v.data() + v.size(); }

static void run(const std::vector<Spec> &specs, const char *name)
{
    std::vector<Buf> bufs(specs.size());
    std::vector<ForestJob> jobs;
    for (size_t i = 0; i < specs.size(); i++) {
        const Spec &s = specs[i];
        const u64 N = 1ull << s.lg, nc = 1ull << s.h, nd = nc < N ? 2 * (N - nc) : 0;
        Buf &b = bufs[i];
        b.dig.assign(nd * 4, 0);
        b.cap.assign(nc * 4, 0);
        b.dig_w.assign(nd, 0);
        b.cap_w.assign(nc, 0);
        b.leaf_hashed.assign(N, 0);
        jobs.push_back({b.lde_dummy, s.rows ? b.rows_dummy : nullptr, nd ? b.dig.data() : nullptr, b.cap.data(), s.lg,
                        s.C, s.h});
    }
    std::vector<unsigned char> bytes;
    std::vector<ForestLaunch> launches;
    CHECK(forest_tables(jobs, bytes, launches), "%s: plan refused", name);
    auto job_of = [&](const u64 *lde) {
        for (size_t i = 0; i < bufs.size(); i++)
            if (lde == bufs[i].lde_dummy) return (int)i;
        return -1;
    };
    auto job_of_digests = [&](const u64 *d, const u64 *caps) {
        for (size_t i = 0; i < bufs.size(); i++)
            if (caps == bufs[i].cap.data() && (d == bufs[i].dig.data() || (!d && bufs[i].dig.empty()))) return (int)i;
        return -1;
    };
    int phase = 0;  // leaves, then levels
    for (const ForestLaunch &l : launches) {
        CHECK((u64)l.blocks * 128 < (1ull << 32), "%s: launch of %u workgroups", name, l.blocks);
        if (l.kind != ForestLaunch::Level) {
            CHECK(phase == 0, "%s: leaves after levels", name);
            const ForestLeafJob *t = (const ForestLeafJob *)(bytes.data() + l.off);
            CHECK(t[0].block0 == 0, "%s: first block0", name);
            for (u32 e = 1; e < l.njobs; e++) CHECK(t[e].block0 > t[e - 1].block0, "%s: unsorted leaf table", name);
            for (u64 b = 0; b < l.blocks; b++) {
                const ForestLeafJob &J = t[forest_find(t, l.njobs, b)];
                const int ji = job_of(J.a.leaves);
                CHECK(ji >= 0, "%s: leaf table points nowhere", name);
                if (ji < 0) continue;
                Buf &B = bufs[ji];
                CHECK((J.a.rows != nullptr) == (l.kind == ForestLaunch::LeavesRows), "%s: rows variant", name);
                CHECK(J.a.col_stride == J.a.n_leaves && J.a.n_leaves == (1ull << specs[ji].lg), "%s: leaf shape", name);
                for (u32 tid = 0; tid < 128; tid++) {
                    const u64 x = (b - J.block0) * 128 + tid;
                    if (x >= J.a.n_leaves) continue;
                    B.leaf_hashed[x]++;
                    // the slot arithmetic of leaf_hash_thread
                    const u64 leaf = zk_bitrev((u32)x, J.a.lg_leaves);
                    u64 slot = leaf;
                    if (J.a.sub_leaves)
                        slot = (leaf / J.a.sub_leaves) * J.a.sub_digests + (J.a.sub_digests - J.a.sub_leaves) +
                               leaf % J.a.sub_leaves;
                    u64 *o = J.a.out + slot * 4;
                    if (J.a.out == B.cap.data()) {
                        CHECK(in(B.cap, o), "%s: leaf writes past the caps", name);
                        if (in(B.cap, o)) B.cap_w[slot]++, o[0] = slot + 1;
                    } else {
                        CHECK(J.a.out == B.dig.data() && in(B.dig, o), "%s: leaf writes past the digests", name);
                        if (in(B.dig, o)) B.dig_w[slot]++, o[0] = slot + 1;
                    }
                }
            }
        } else {
            phase = 1;
            CHECK(l.g == 1 || l.g == 4 || l.g == 16, "%s: lanes per node %d", name, l.g);
            const ForestLevelJob *t = (const ForestLevelJob *)(bytes.data() + l.off);
            CHECK(t[0].block0 == 0, "%s: first block0", name);
            for (u32 e = 1; e < l.njobs; e++) CHECK(t[e].block0 > t[e - 1].block0, "%s: unsorted level table", name);
            // writes of this launch become visible to the next launch only
            std::vector<std::pair<u64 *, u64>> pending;
            for (u64 b = 0; b < l.blocks; b++) {
                const ForestLevelJob &J = t[forest_find(t, l.njobs, b)];
                const int ji = job_of_digests(J.d, J.caps);
                CHECK(ji >= 0, "%s: level table points nowhere", name);
                if (ji < 0) continue;
                Buf &B = bufs[ji];
                for (u32 tid = 0; tid < 128; tid++) {
                    const u64 x = ((b - J.block0) * 128 + tid) / l.g;
                    const bool ok = x < J.total;
                    const ForestNode n = forest_node(J, ok ? x : 0);
                    u64 *sd = J.d + n.s * J.D * 4;
                    const u64 *ca = sd + n.c0 * 4, *cb = ca + 4;
                    // every lane reads the children (also lanes that store nothing)
                    CHECK(in(B.dig, ca) && in(B.dig, cb), "%s: child outside the digests", name);
                    if (!in(B.dig, ca) || !in(B.dig, cb)) continue;
                    CHECK(ca[0] != 0 && cb[0] != 0, "%s: job %d child %llu read before it is written", name, ji,
                          (unsigned long long)((ca - B.dig.data()) / 4));
                    if (!ok || tid % l.g) continue;  // one writer per node
                    if (n.out < 0) {
                        u64 *o = J.caps + n.s * 4;
                        CHECK(in(B.cap, o), "%s: cap outside", name);
                        if (in(B.cap, o)) B.cap_w[n.s]++, pending.push_back({o, n.s + 1});
                    } else {
                        u64 *o = sd + (u64)n.out * 4;
                        CHECK(in(B.dig, o) && (u64)n.out < J.D, "%s: node outside", name);
                        if (in(B.dig, o)) B.dig_w[(o - B.dig.data()) / 4]++, pending.push_back({o, 1});
                    }
                }
            }
            for (auto &w : pending) w.first[0] = w.second;
        }
    }
    for (size_t i = 0; i < bufs.size(); i++) {
        const Buf &B = bufs[i];
        for (u64 x = 0; x < B.leaf_hashed.size(); x++)
            CHECK(B.leaf_hashed[x] == 1, "%s: job %zu leaf %llu hashed %d times", name, i, (unsigned long long)x,
                  B.leaf_hashed[x]);
        for (u64 k = 0; k < B.dig_w.size(); k++)
            CHECK(B.dig_w[k] == 1, "%s: job %zu digest %llu written %d times", name, i, (unsigned long long)k,
                  B.dig_w[k]);
        for (u64 k = 0; k < B.cap_w.size(); k++)
            CHECK(B.cap_w[k] == 1, "%s: job %zu cap %llu written %d times", name, i, (unsigned long long)k, B.cap_w[k]);
    }
    printf("%s: %zu jobs, %zu launches\n", name, specs.size(), launches.size());
}

int main()
{
    // the edge cases: (lg leaves, columns, cap_height, rows)
    run({{4, 5, 2, true}, {6, 7, 3, false}, {9, 16, 4, true}, {6, 1, 0, false}, {9, 9, 6, true}, {3, 3, 3, false},
         {9, 12, 1, true}, {12, 4, 0, false}, {8, 135, 4, true}, {1, 8, 0, false}, {1, 8, 1, true}, {2, 3, 1, false}},
        "edge cases");
    // batched-workload shapes, all with rows and mixed
    std::mt19937_64 g(3);
    for (int list = 0; list < 4; list++) {
        std::vector<Spec> v;
        for (int i = 0; i < 48; i++) {
            const u32 k = 10 + g() % 9, r = 1 + g() % 3, lg = std::min<u32>(k + r, list < 2 ? 21 : 16);
            v.push_back({lg, (u32)(4 + g() % 132), std::min<u32>(4, lg - 2), list % 2 == 0 || (g() & 1)});
        }
        if (list >= 2)
            for (auto &s : v) s.lg = std::min<u32>(s.lg, 14);  // smaller lists: more, shorter trees
        char name[32];
        snprintf(name, sizeof name, "mixed list %d", list);
        run(v, name);
    }
    // plan sizes only: a 2^31-leaf job and many small ones
    {
        std::vector<ForestJob> jobs;
        static u64 dummy[4];
        jobs.push_back({dummy, nullptr, dummy, dummy, 31, 8, 4});
        for (int i = 0; i < 100; i++) jobs.push_back({dummy, nullptr, dummy, dummy, 20, 8, 4});
        std::vector<unsigned char> bytes;
        std::vector<ForestLaunch> launches;
        CHECK(forest_tables(jobs, bytes, launches), "large plan refused");
        u64 leaf_blocks = 0;
        for (const ForestLaunch &l : launches) {
            CHECK((u64)l.blocks * 128 < (1ull << 32), "large plan: launch of %u workgroups", l.blocks);
            if (l.kind != ForestLaunch::Level) leaf_blocks += l.blocks;
        }
        CHECK(leaf_blocks == (1ull << 24) + 100 * (1ull << 13), "large plan: %llu leaf workgroups",
              (unsigned long long)leaf_blocks);
        jobs[0].cap_height = 32;
        CHECK(!forest_tables(jobs, bytes, launches), "cap_height above lg accepted");
    }
    printf("%s\n", fails ? "FAIL" : "ok");
    return fails ? 1 : 0;
}
