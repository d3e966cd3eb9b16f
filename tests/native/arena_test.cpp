// Randomised check of the commit server's first-fit arena (zeknox_amd/csrc/arena.hpp):
// live ranges never overlap and stay inside the arena, accounting matches, and
// freeing everything coalesces back to one range.  Prints "ok" on success.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../zeknox_amd/csrc/arena.hpp"

struct Live {
    uint64_t off, bytes, len;
};

int main()
{
    const uint64_t size = 1ull << 30, align = 256;
    zk::Arena a(size, align);
    std::mt19937_64 g(7);
    std::vector<Live> live;
    uint64_t used = 0;
    for (int step = 0; step < 200000; step++) {
        const bool do_alloc = live.empty() || (g() % 100) < 55;
        if (do_alloc) {
            const uint64_t bytes = (g() % 4 == 0) ? (g() % (64ull << 20)) : (g() % 65536);
            uint64_t off = 0;
            if (a.alloc(bytes, off)) {
                const uint64_t len = a.round(bytes ? bytes : 1);
                if (off % align || off + len > size) { printf("bad range\n"); return 1; }
                for (auto &l : live)
                    if (off < l.off + l.len && l.off < off + len) { printf("overlap\n"); return 1; }
                live.push_back({off, bytes, len});
                used += len;
            } else if (a.largest_free() >= a.round(bytes ? bytes : 1)) {
                printf("fit refused\n");
                return 1;
            }
        } else {
            const size_t i = g() % live.size();
            a.release(live[i].off, live[i].bytes);
            used -= live[i].len;
            live[i] = live.back();
            live.pop_back();
        }
        if (a.used() != used) { printf("accounting\n"); return 1; }
    }
    for (auto &l : live) a.release(l.off, l.bytes);
    if (a.used() != 0 || a.free_ranges() != 1 || a.largest_free() != size) { printf("no coalesce\n"); return 1; }
    uint64_t off;
    if (a.alloc(size + 1, off)) { printf("oversize\n"); return 1; }
    printf("ok\n");
    return 0;
}
