// One Plonky2 commit on a stream: coset LDE -> (optional row-major leaves) ->
// Poseidon Merkle digests -> cap.  Shared by zk_commit, the batched entry
// point and the commit server.
#pragma once
#include "runtime.hpp"

namespace zk {

// Same field layout as zk_commit_job of include/zeknox_ext.h.
struct CommitJob {
    const u64 *coeffs;  // ncols x 2^lg_n, column-major, device
    u64 *lde;           // ncols x 2^(lg_n + rate_bits), column-major, device
    u64 *leaves;        // optional row-major leaves (nullptr: not written)
    u64 *digests;       // 2 (2^(lg_n+rate_bits) - 2^cap_height) x 4 (unused when that is 0)
    u64 *cap;           // 2^cap_height x 4
    u32 lg_n, rate_bits, ncols, cap_height;
};

// Throws Error(invalid args) for a job the path cannot run, including one
// whose coefficients, LDE, leaves, digests or cap ranges overlap.
void commit_validate(const CommitJob &j);

// Enqueues the whole commit on `st` (no host synchronisation) with the coset
// power rows `gpw` (3 x 2^11 u64 on the device; nullptr: the device
// context's, i.e. its current generator).  use_scratch: the LDE may use the
// library scratch buffer (callers serialised on the device context only).
void commit_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st, bool use_scratch = false,
                   const u64 *gpw = nullptr);

// The coset LDE of the commit alone (the first part of commit_launch).
void lde_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st, bool use_scratch = false, const u64 *gpw = nullptr);

// The same commit with no device allocation (noalloc.hip): the blocked LDE
// intermediate lives in the library's fixed per-device scratch array (the LDE
// runs in column chunks that fit it) and no twiddle table is built.  The
// caller holds c.mu; the coset is the context's current generator.
void commit_noalloc_launch(DeviceCtx &c, const CommitJob &j, hipStream_t st);

// Rough device time of a commit in microseconds (scheduling weight only):
// leaf permutations at ~1.8 G/s plus ~50 us per tree level and LDE pass.
double commit_cost_us(const CommitJob &j);

// Rough device time of the LDE alone in microseconds (scheduling weight).
double lde_cost_us(const CommitJob &j);

// Device bytes a commit writes beside its coefficients.
u64 commit_output_bytes(u32 lg_n, u32 rate_bits, u32 ncols, u32 cap_height, bool leaves);

}  // namespace zk
