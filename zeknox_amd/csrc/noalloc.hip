// Allocation-free commit (zk_commit_noalloc): the commit of commit.hip with
// every device buffer it needs beyond the caller's inputs and outputs taken
// from one fixed array per device, so that callers that cannot allocate
// device memory (code being captured into a graph, tight serving loops,
// harnesses that forbid hipMalloc) can still commit.
//
// The only such buffer of the commit is the blocked intermediate of a
// two-pass 2^20-point LDE (ntt.hip, ntt_launch): pass 1 writes the 2^10
// elements of each 8 consecutive pass-2 fast indices as one contiguous block,
// pass 2 reads it.  Columns are independent, so the LDE runs in chunks of as
// many columns as the fixed array holds; other transform sizes run their
// passes in the output buffer and use no scratch.  No new twiddle table is
// built: an existing one is used, else the passes compute their twiddles.
//
// The array lives in this code object on every device the library runs on,
// until the process ends.  Calls on one device share it and are serialised by
// the device context's mutex; it keeps the last call's contents, which
// nothing reads before pass 1 has rewritten them.
#include <algorithm>

#include "commit.hpp"
#include "merkle.hpp"
#include "ntt.hpp"

namespace zk {

constexpr size_t kFixedScratchBytes = (size_t)256 << 20;  // 256 MiB: 32 columns of 2^20 points

static __device__ __align__(256) u64 scratch[kFixedScratchBytes / sizeof(u64)];

void fixed_scratch_setup(DeviceCtx &c)
{
    // the array's name is only a host-side handle: kernels and copies get
    // this device's address of it
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        c.fixed_scratch = nullptr;
        c.fixed_scratch_bytes = 0;
        c.fixed_scratch_err = std::string("hipGetSymbolAddress of the fixed scratch failed: ") +
                              hipGetErrorString(e);
        return;
    }
    c.fixed_scratch = (u64 *)p;
    c.fixed_scratch_bytes = kFixedScratchBytes;
}

void commit_noalloc_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st)
{
    if (!c.fixed_scratch) throw Error(-(int)hipErrorInvalidSymbol, c.fixed_scratch_err);
    const u32 lg = j.lg_n + j.rate_bits;
    const u64 N = (u64)1 << lg, n = (u64)1 << j.lg_n, ncaps = (u64)1 << j.cap_height;
    const bool blocked = ntt_blocked_intermediate(lg);
    const u64 col_bytes = N * sizeof(u64);
    if (blocked && col_bytes > c.fixed_scratch_bytes)
        throw Error((int)hipErrorInvalidValue, "invalid args: one column of the LDE exceeds the fixed scratch of " +
                                                   std::to_string(c.fixed_scratch_bytes >> 20) + " MiB");
    const u32 per = blocked ? (u32)std::min<u64>(j.ncols, c.fixed_scratch_bytes / col_bytes) : j.ncols;
    for (u32 c0 = 0; c0 < j.ncols; c0 += per) {
        NttJob nj;
        nj.src = j.coeffs + (u64)c0 * n;
        nj.dst = j.lde + (u64)c0 * N;
        nj.src_stride = n;
        nj.dst_stride = N;
        nj.ncols = std::min<u32>(per, j.ncols - c0);
        nj.lg = lg;
        nj.lg_in = j.lg_n;
        nj.coset_in = true;
        nj.prof_tag = "lde";
        nj.no_alloc = true;
        nj.inter = blocked ? c.fixed_scratch : nullptr;
        nj.inter_bytes = blocked ? c.fixed_scratch_bytes : 0;
        ntt_launch(c, nj, st);
    }
    const TreeShape t = tree_shape(ncaps < N ? 2 * (N - ncaps) : 0, ncaps, N, j.ncols, j.cap_height);
    merkle_launch(t, j.digests, j.cap, j.lde, LeafLayout::ColMajorRev, N, st, j.leaves);
}

}  // namespace zk
