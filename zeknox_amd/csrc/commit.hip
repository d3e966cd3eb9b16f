// One commit on a stream (commit.hpp).
#include "commit.hpp"
#include "merkle.hpp"
#include "ntt.hpp"

namespace zk {

void commit_validate(const CommitJob &j)
{
    const u32 lg = j.lg_n + j.rate_bits;
    if (j.lg_n == 0 || lg > 31 || j.ncols == 0 || !j.coeffs || !j.lde || !j.cap || j.cap_height > lg)
        throw Error((int)hipErrorInvalidValue, "invalid args");
    if (j.cap_height < lg && !j.digests) throw Error((int)hipErrorInvalidValue, "invalid args");
    // the LDE reads the coefficients while it writes the evaluations (an
    // in-place LDE would also need the shared scratch buffer), and the leaf
    // hash reads the LDE while it writes leaves and digests: no two ranges
    // may overlap
    const u64 N = (u64)1 << lg, ncaps = (u64)1 << j.cap_height;
    const u64 nd = ncaps < N ? 2 * (N - ncaps) : 0;
    struct Range {
        const void *p;
        u64 bytes;
    } r[5] = {{j.coeffs, 8 * (u64)j.ncols << j.lg_n},
              {j.lde, 8 * (u64)j.ncols * N},
              {j.leaves, j.leaves ? 8 * (u64)j.ncols * N : 0},
              {j.digests, nd ? 32 * nd : 0},
              {j.cap, 32 * ncaps}};
    for (int a = 0; a < 5; a++)
        for (int b = a + 1; b < 5; b++) {
            const uintptr_t pa = (uintptr_t)r[a].p, pb = (uintptr_t)r[b].p;
            if (r[a].bytes && r[b].bytes && pa < pb + r[b].bytes && pb < pa + r[a].bytes)
                throw Error((int)hipErrorInvalidValue, "invalid args: overlapping buffers");
        }
}

void lde_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st, bool use_scratch, const u64 *gpw)
{
    const u32 lg = j.lg_n + j.rate_bits;
    const u64 N = (u64)1 << lg;
    NttJob nj;
    nj.src = j.coeffs;
    nj.dst = j.lde;
    nj.src_stride = (u64)1 << j.lg_n;
    nj.dst_stride = N;
    nj.ncols = j.ncols;
    nj.lg = lg;
    nj.lg_in = j.lg_n;
    nj.coset_in = true;
    nj.prof_tag = "lde";
    nj.use_scratch = use_scratch;
    nj.gpw = gpw;
    ntt_launch(c, nj, st);
}

void commit_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st, bool use_scratch, const u64 *gpw)
{
    const u32 lg = j.lg_n + j.rate_bits;
    const u64 N = (u64)1 << lg, ncaps = (u64)1 << j.cap_height;
    const TreeShape t = tree_shape(ncaps < N ? 2 * (N - ncaps) : 0, ncaps, N, j.ncols, j.cap_height);
    lde_launch(c, j, st, use_scratch, gpw);
    merkle_launch(t, j.digests, j.cap, j.lde, LeafLayout::ColMajorRev, N, st, j.leaves);
}

double commit_cost_us(const CommitJob &j)
{
    const u32 lg = j.lg_n + j.rate_bits;
    const double N = (double)((u64)1 << lg);
    const double perms_per_leaf = j.ncols > 4 ? (double)((j.ncols + 7) / 8) : 0.0;
    const double passes = lg <= 10 ? 1 : (lg <= 20 ? 2 : 3);
    return N * perms_per_leaf / 1800.0 + 16.0 * N * j.ncols / 2.0e6 * passes + 50.0 * (lg - j.cap_height + 1 + passes);
}

double lde_cost_us(const CommitJob &j)
{
    const u32 lg = j.lg_n + j.rate_bits;
    const double N = (double)((u64)1 << lg);
    const double passes = lg <= 10 ? 1 : (lg <= 20 ? 2 : 3);
    return 16.0 * N * j.ncols / 2.0e6 * passes + 10.0 * passes;
}

u64 commit_output_bytes(u32 lg_n, u32 rate_bits, u32 ncols, u32 cap_height, bool leaves)
{
    const u64 N = (u64)1 << (lg_n + rate_bits), ncaps = (u64)1 << cap_height;
    const u64 nd = ncaps < N ? 2 * (N - ncaps) : 0;
    return 8 * (u64)ncols * N * (leaves ? 2 : 1) + 32 * nd + 32 * ncaps;
}

}  // namespace zk
