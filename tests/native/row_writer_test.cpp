// Host test of the leaf kernel's row writer (zeknox_amd/csrc/leaf_rows.hpp)
// driven by the sponge's chunk hook (hash_one_chunks), as the fused commit's
// leaf kernel drives it: for leaf sizes 1..20, 135 and 400 and every 8-byte
// misalignment of the row matrix's base, the leaves of a 64-leaf matrix are
// written in a shuffled order (lanes finish in any order) into a buffer with
// guard zones.  Checks: the matrix equals the leaves row by row, nothing
// outside it is written, the digest equals hash_one's, and every sector that
// lies inside one row is stored in a single call (whole), at most two per row
// partially.  Prints "ok".
#include <algorithm>
#include <cstdio>
#include <random>
#include <vector>

#include "../../zeknox_amd/csrc/leaf_rows.hpp"
#include "../../zeknox_amd/csrc/poseidon.hpp"

using namespace zk;

static int fails = 0;
#define CHECK(c, ...)                        \
    do {                                     \
        if (!(c)) {                          \
            if (fails++ < 10) {              \
                printf("FAIL: " __VA_ARGS__); \
                printf("\n");                \
            }                                \
        }                                    \
    } while (0)

static u64 value(u64 leaf, u32 i) { return (leaf * 0x9E3779B97F4A7C15ull + i * 0xD1B54A32D192ED03ull) % ZK_P; }

int main()
{
    constexpr u64 kSentinel = 0xA5A5A5A5A5A5A5A5ull;
    constexpr u32 kLeaves = 64;
    std::mt19937_64 g(3);
    std::vector<u32> sizes;
    for (u32 c = 1; c <= 20; c++) sizes.push_back(c);
    sizes.push_back(135);
    sizes.push_back(400);
    long full = 0, partial = 0;
    for (u32 C : sizes) {
        for (u32 mis = 0; mis < 8; mis++) {
            // 64-byte aligned allocation, guard of 16 elements on both sides
            const size_t n = (size_t)kLeaves * C;
            std::vector<u64> mem(n + 48, kSentinel);
            u64 *al = (u64 *)(((uintptr_t)mem.data() + 63) & ~(uintptr_t)63);
            u64 *rows = al + 16 + mis;
            CHECK(rows + n + 16 <= mem.data() + mem.size(), "buffer");
            std::vector<u32> order(kLeaves);
            for (u32 i = 0; i < kLeaves; i++) order[i] = i;
            std::shuffle(order.begin(), order.end(), g);
            u64 ring[kRowRingPitch];
            for (u32 leaf : order) {
                for (auto &r : ring) r = kSentinel;
                RowSectorWriter w(rows, leaf, C, ring);
                auto get = [leaf](u32 i) { return value(leaf, i); };
                u64 h[4], h0[4];
                u32 last_q = w.q, chunks = 0;
                hash_one_chunks(get, C, h, [&](u32 idx, u32 m, const u64 *v) {
                    CHECK(idx == 8 * chunks && m == std::min<u32>(8, C - idx), "chunk %u of %u: idx %u m %u", chunks,
                          C, idx, m);
                    for (u32 i = 0; i < m; i++) CHECK(v[i] == value(leaf, idx + i), "chunk value");
                    const u32 before = w.next;
                    w.chunk(m, v);
                    // sectors stored by this chunk are whole unless it is the row's first
                    for (u32 s = before; s < w.next; s++) (s == 0 && w.o ? partial : full)++;
                    CHECK(w.q == last_q + m, "q");
                    last_q = w.q;
                    CHECK(w.q - 8 * w.next < 8, "pending positions %u", w.q - 8 * w.next);
                    chunks++;
                });
                if (8 * w.next < w.q) partial++;
                w.finish();
                hash_one(get, C, h0);
                for (int k = 0; k < 4; k++) CHECK(h[k] == h0[k], "digest C=%u", C);
                CHECK(ring[kRowRingPitch - 1] == kSentinel, "ring overrun");
            }
            for (u32 leaf = 0; leaf < kLeaves; leaf++)
                for (u32 i = 0; i < C; i++)
                    CHECK(rows[(u64)leaf * C + i] == value(leaf, i), "C=%u mis=%u leaf %u element %u", C, mis, leaf,
                          i);
            for (u64 *p = mem.data(); p < rows; p++) CHECK(*p == kSentinel, "write before the matrix C=%u mis=%u", C, mis);
            for (u64 *p = rows + n; p < mem.data() + mem.size(); p++)
                CHECK(*p == kSentinel, "write after the matrix C=%u mis=%u", C, mis);
        }
    }
    // at most two partial sectors per row
    long rows_total = (long)sizes.size() * 8 * kLeaves;
    CHECK(partial <= 2 * rows_total, "partial sectors %ld for %ld rows", partial, rows_total);
    printf("%s: %ld whole sectors, %ld partial, %ld rows\n", fails ? "FAIL" : "ok", full, partial, rows_total);
    return fails ? 1 : 0;
}
