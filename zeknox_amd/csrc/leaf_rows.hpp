// Row-major leaf rows written in whole 64-byte sectors by the lane that hashes
// the leaf.
//
// A lane of the fused commit's leaf kernel absorbs its leaf 8 elements at a
// time and also writes the leaf to the row-major leaf matrix.  A row of C
// elements starts at element (leaf * C) of that matrix, so for C = 135 (1080-
// byte rows) a 64-byte chunk of a row straddles two 64-byte sectors in seven
// rows of eight; the two halves of such a sector are written one absorb step
// (one permutation, ~0.1 ms) apart, by which time the first half has left L2 as
// a partial write.  The writer instead keeps the row's pending elements in a
// 16-element per-lane LDS ring and stores a sector only once all of its
// elements in this row are known: every sector inside the row is written whole
// in one burst, and only the row's first and last sectors (shared with the
// neighbouring rows) are partial.
#pragma once
#include "gl64.hpp"

namespace zk {

constexpr u32 kRowRing = 16;       // ring elements per lane
constexpr u32 kRowRingPitch = 17;  // LDS pitch per lane (odd: spreads the lanes over the banks)

struct RowSectorWriter {
    u64 *base;  // 64-byte aligned: sector k of the matrix is base[8k .. 8k + 8)
    u64 s0;     // first sector of the row
    u32 o;      // position of the row's element 0 in sector s0
    u32 next;   // next sector to store, relative to s0
    u32 q;      // positions of the row known so far, from the start of sector s0
    u64 *ring;  // kRowRing elements; position q lives at ring[q % kRowRing]

    ZK_HD RowSectorWriter(u64 *rows, u64 leaf, u32 leaf_size, u64 *lane_ring)
    {
        const uintptr_t addr = (uintptr_t)rows;  // rows is 8-byte aligned
        base = (u64 *)(addr & ~(uintptr_t)63);
        const u64 g0 = (u64)((addr & 63) >> 3) + leaf * leaf_size;
        s0 = g0 >> 3;
        o = (u32)(g0 & 7);
        next = 0;
        q = o;
        ring = lane_ring;
    }

    // positions [lo, hi) of sector `next`
    ZK_HD void store(u32 lo, u32 hi)
    {
        u64 *dst = base + (s0 + next) * 8;
#pragma unroll
        for (u32 e = 0; e < 8; e++)
            if (e >= lo && e < hi) dst[e] = ring[(8 * next + e) % kRowRing];
    }

    // the row's next m <= 8 elements; at most 7 positions are pending before
    // it, so the m new ones never overwrite a pending one
    ZK_HD void chunk(u32 m, const u64 *v)
    {
#pragma unroll
        for (u32 i = 0; i < 8; i++)
            if (i < m) ring[(q + i) % kRowRing] = v[i];
        q += m;
        while (8 * next + 8 <= q) {
            store(next == 0 ? o : 0, 8);
            next++;
        }
    }

    // the row's last, partial sector
    ZK_HD void finish()
    {
        if (8 * next < q) store(next == 0 ? o : 0, q - 8 * next);
    }
};

}  // namespace zk
