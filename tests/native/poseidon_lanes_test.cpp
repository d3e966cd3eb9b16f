// Host emulation of the lane-parallel Poseidon permutations
// (zeknox_amd/csrc/poseidon_hex.hpp: 16 lanes, one state element each;
// poseidon_quad.hpp: 4 lanes, three elements each): the lanes of a group run
// their S-box and MDS phases in lock step, the gather between them done here
// the way the device exchange does it (lane j reads element (j + k) mod 12 of
// its group).  Every output must equal the exact one-lane permutation
// (poseidon_permute_safe, canonical) on random states, on edge values and on
// states whose first S-box square hits the relaxed arithmetic's rare case,
// both on the relaxed path (when it reports no rare case) and on the exact one.
// Prints "ok".
#include <cstdio>
#include <random>
#include <vector>

#include "../../zeknox_amd/csrc/poseidon_hex.hpp"
#include "../../zeknox_amd/csrc/poseidon_quad.hpp"

using namespace zk;

static int fails = 0;

// 16-lane group: element e of lane j is j < 12 ? j : j - 12.
template <bool SAFE>
static void permute_hex(u64 st[12], bool &rare)
{
    const u64 *rc = round_constants();
    u64 s[16];
    for (u32 j = 0; j < 16; j++) {
        const u32 e = j < 12 ? j : j - 12;
        s[j] = SAFE ? gl_add(st[e], rc[e]) : gl_add_nc(st[e], rc[e], rare);
    }
    for (int r = 0; r < 30; r++) {
        const bool full = r < 4 || r >= 26;
        for (u32 j = 0; j < 16; j++) s[j] = hex_sbox<SAFE>(s[j], j < 12 ? j : j - 12, full, rare);
        u64 n[16];
        for (u32 j = 0; j < 16; j++) {
            const u32 e = j < 12 ? j : j - 12;
            u64 t[12];
            for (u32 k = 0; k < 12; k++) t[k] = s[(e + k) % 12];  // lane (e + k) mod 12 of the group
            n[j] = hex_mds<SAFE>(t, e, rc[12 * (r + 1) + e]);
        }
        for (u32 j = 0; j < 16; j++) s[j] = n[j];
    }
    for (u32 j = 12; j < 16; j++)
        if (s[j] != s[j - 12] && fails++ < 10) printf("hex: mirror lane %u differs\n", j);
    for (u32 e = 0; e < 12; e++) st[e] = s[e];
}

// 4-lane quad: lane j holds elements 3j..3j+2.
template <bool SAFE>
static void permute_quad(u64 st[12], bool &rare)
{
    const u64 *rc = round_constants();
    u64 s[4][3];
    for (u32 j = 0; j < 4; j++)
        for (int k = 0; k < 3; k++)
            s[j][k] = SAFE ? gl_add(st[3 * j + k], rc[3 * j + k]) : gl_add_nc(st[3 * j + k], rc[3 * j + k], rare);
    for (int r = 0; r < 30; r++) {
        const bool full = r < 4 || r >= 26;
        for (u32 j = 0; j < 4; j++) quad_sbox<SAFE>(s[j], j, full, rare);
        u64 n[4][3];
        for (u32 j = 0; j < 4; j++) {
            u64 t[12];
            for (u32 e = 0; e < 12; e++) {
                const u32 g = (3 * j + e) % 12;  // element g is lane g / 3, slot g % 3
                t[e] = s[g / 3][g % 3];
            }
            const u64 rk[3] = {rc[12 * (r + 1) + 3 * j], rc[12 * (r + 1) + 3 * j + 1], rc[12 * (r + 1) + 3 * j + 2]};
            quad_mds<SAFE>(t, j, rk, n[j]);
        }
        for (u32 j = 0; j < 4; j++)
            for (int k = 0; k < 3; k++) s[j][k] = n[j][k];
    }
    for (u32 e = 0; e < 12; e++) st[e] = s[e / 3][e % 3];
}

static long relaxed_rare = 0, checked = 0;

static void check(const u64 in[12], const char *what)
{
    u64 ref[12];
    for (int i = 0; i < 12; i++) ref[i] = gl_canon(in[i]);
    poseidon_permute_safe(ref);
    for (int v = 0; v < 2; v++) {
        u64 a[12], b[12];
        for (int i = 0; i < 12; i++) a[i] = b[i] = in[i];
        bool rare = false;
        if (v == 0) permute_hex<false>(a, rare); else permute_quad<false>(a, rare);
        checked++;
        if (!rare) {
            for (int i = 0; i < 12; i++)
                if (gl_canon(a[i]) != ref[i] && fails++ < 10)
                    printf("%s %s relaxed: element %d %016llx want %016llx\n", v ? "quad" : "hex", what, i,
                           (unsigned long long)gl_canon(a[i]), (unsigned long long)ref[i]);
        } else {
            relaxed_rare++;
        }
        for (int i = 0; i < 12; i++) b[i] = gl_canon(in[i]);
        bool unused = false;
        if (v == 0) permute_hex<true>(b, unused); else permute_quad<true>(b, unused);
        for (int i = 0; i < 12; i++)
            if (b[i] != ref[i] && fails++ < 10)
                printf("%s %s exact: element %d %016llx want %016llx\n", v ? "quad" : "hex", what, i,
                       (unsigned long long)b[i], (unsigned long long)ref[i]);
    }
}

int main(int argc, char **argv)
{
    std::mt19937_64 g(99);
    const u64 edge[] = {0, 1, ZK_P - 1, ZK_P, ZK_P + 1, 0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFull, 0x100000000ull};
    for (int n = 0; n < 2000; n++) {
        u64 s[12];
        for (int i = 0; i < 12; i++) s[i] = g();
        check(s, "random");
    }
    for (u64 x : edge) {
        u64 s[12];
        for (int i = 0; i < 12; i++) s[i] = (i & 1) ? x : g();
        check(s, "edge");
        for (int i = 0; i < 12; i++) s[i] = x;
        check(s, "edge");
    }
    // rare-case states (element 0 / 1 of the first S-box) from argv: hex pairs
    for (int a = 1; a + 1 < argc; a += 2) {
        u64 s[12] = {0};
        s[strtoull(argv[a], nullptr, 10)] = strtoull(argv[a + 1], nullptr, 16);
        for (int i = 0; i < 12; i++)
            if (i != (int)strtoull(argv[a], nullptr, 10)) s[i] = g() % ZK_P;
        check(s, "rare");
    }
    printf("%s: %ld permutations, %ld relaxed runs reported a rare case\n", fails ? "FAIL" : "ok", checked,
           relaxed_rare);
    return fails ? 1 : 0;
}
