/* Extensions of the zeknox C ABI in libzeknox.so for MI355X.  None of these
 * has a counterpart in the reference (okx/zeknox); they sit beside the
 * unchanged lib.h / merkle.h entry points, take device pointers and sizes
 * only, and report errors as RustError like the rest of the ABI.
 *
 * The reference's commit path is three calls (compute_batched_lde,
 * compute_transpose_rev, fill_digests_buf_linear_gpu_with_gpu_ptr, see
 * wrappers/rust/tests/merkle_tree.rs and native/lib.cu:80-113); zk_commit
 * computes the same LDE, digests and cap in one call and skips the transpose
 * round trip by hashing leaves straight from the column-major LDE. */
#ifndef ZEKNOX_EXT_H
#define ZEKNOX_EXT_H

#include <stddef.h>
#include <stdint.h>
#include "utils/rusterror.h"

#ifdef __cplusplus
extern "C" {
#endif

/* One commit: coset LDE of ncols polynomials of 2^lg_n coefficients
 * (column-major, device) with blowup 2^rate_bits and shift `shift`, then the
 * Poseidon Merkle forest of its 2^(lg_n+rate_bits) leaves (leaf l = LDE row
 * rev(l)) with 2^cap_height subtrees.
 *   lde     : ncols x 2^(lg_n+rate_bits) u64 (column-major), required
 *   leaves  : optional (NULL to skip) row-major transposed leaves, as
 *             compute_transpose_rev would write them
 *   digests : 2 * (2^(lg_n+rate_bits) - 2^cap_height) x 4 u64
 *   cap     : 2^cap_height x 4 u64
 * Results equal compute_batched_lde -> compute_transpose_rev ->
 * fill_digests_buf_linear_gpu_with_gpu_ptr bit for bit. */
RustError zk_commit(size_t device_id, const uint64_t *coeffs, uint64_t *lde, uint64_t *leaves,
                    uint64_t *digests, uint64_t *cap, uint32_t lg_n, uint32_t rate_bits, uint32_t ncols,
                    uint32_t cap_height, uint64_t shift);

/* zk_commit with no device allocation, for callers that cannot allocate
 * (code under graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The device must already be
 * initialised (init_cuda, init_twiddle_factors or any other call on it) and
 * `shift` must equal its coset generator (7 unless init_coset changed it);
 * code 1 otherwise.  The LDE's blocked intermediate (2^20-row commits) lives
 * in a fixed 256 MiB array the library holds on every device it runs on until
 * the process ends; the LDE runs in column chunks that fit it (32 columns of
 * 2^20 rows), so no size is refused for it.  Every call on a device shares
 * that array: calls are serialised on the device, and calls that overlapped
 * on one device (e.g. from a caller's own copy of the kernels) would race.
 * Results equal zk_commit bit for bit. */
RustError zk_commit_noalloc(size_t device_id, const uint64_t *coeffs, uint64_t *lde, uint64_t *leaves,
                            uint64_t *digests, uint64_t *cap, uint32_t lg_n, uint32_t rate_bits, uint32_t ncols,
                            uint32_t cap_height, uint64_t shift);

/* One instance of a batched commit (field meaning as zk_commit; shift is the
 * device's current coset generator, 7 unless init_coset changed it). */
typedef struct {
    const uint64_t *coeffs; /* ncols x 2^lg_n, column-major, device */
    uint64_t *lde;          /* ncols x 2^(lg_n+rate_bits), column-major, device */
    uint64_t *leaves;       /* optional row-major leaves (NULL: not written) */
    uint64_t *digests;      /* 2 * (2^(lg_n+rate_bits) - 2^cap_height) x 4 */
    uint64_t *cap;          /* 2^cap_height x 4 */
    uint32_t lg_n, rate_bits, ncols, cap_height;
} zk_commit_job;

#define ZK_BATCH_GRAPH 1u /* capture the batch in a hipGraph, replay on repeats */

/* Commits every job.  The jobs' coset LDEs (the kernels zk_commit launches)
 * are spread over four streams, longest first onto the least-loaded one,
 * forked from and joined back into one origin stream with events; then the
 * Merkle forests of all jobs run as fused launches on the origin stream: one
 * leaf-hash launch over the leaves of every job and one launch per tree
 * height over every job's trees (no device-side waits).  Results equal n
 * separate zk_commit calls.  With ZK_BATCH_GRAPH the first call with a given
 * job list runs eagerly, the second captures the whole DAG into a hipGraph
 * and later calls with the same list replay it; at most 8 job lists (with
 * their launch tables in device memory) are kept per device (least recently
 * used evicted first; zk_batch_clear drops them all).  Buffers of one job
 * must not overlap (code 1 otherwise).  Returns with all results complete. */
RustError zk_batch_commit(size_t device_id, const zk_commit_job *jobs, uint32_t njobs, uint32_t flags);

/* Frees cached batch plans and graphs of a device. */
RustError zk_batch_clear(size_t device_id);

/* Merkle openings for FRI queries over a linear-layout forest (the digests of
 * fill_digests_buf_linear_gpu_with_gpu_ptr / zk_commit; device buffers).
 * For each leaf index (host array, every index < n_leaves) it writes the
 * lg(n_leaves) - cap_height sibling digests of the leaf-to-cap path, leaf
 * level first (siblings_out: n_indices x layers x 4 u64), and, when rows_out
 * is not NULL, the leaf row (n_indices x leaf_size u64) read from row-major
 * leaves (col_stride 0) or from a column-major LDE with col_stride rows (leaf
 * l = LDE row rev(l)).  Hashing the row, then combining with each sibling
 * (the path node is the left child when its position bit is 0), gives
 * cap[index >> layers]. */
RustError zk_merkle_proofs(size_t device_id, const uint64_t *digests, const uint64_t *leaves, uint64_t col_stride,
                           uint64_t n_leaves, uint64_t leaf_size, uint64_t cap_height, const uint64_t *indices,
                           uint64_t n_indices, uint64_t *siblings_out, uint64_t *rows_out);

/* Commit server: continuous batching of commit requests on one device.
 * Requests are admitted in arrival order while fewer than max_inflight run and
 * the server's device-memory budget (allocated once at creation) has room for
 * the request's coefficients, LDE, digests and cap.  The requests admitted at
 * one moment form a group (at most max_inflight / 2): their LDEs run on
 * several streams, then their Merkle forests as fused launches (one leaf-hash
 * launch over the group's leaves, one launch per tree height); groups admitted
 * while others run overlap with them on the device.  Results equal zk_commit of the
 * same request with the coset generator the device had when the server was
 * created (7 unless init_coset changed it before; the server keeps its own
 * copy of the coset table, so later init_coset or zk_commit calls with
 * another shift do not affect it).  Thread-safe. */
typedef struct zk_server zk_server;

typedef struct {
    uint32_t lg_n, rate_bits, ncols, cap_height; /* cap_height <= 12 */
    const uint64_t *coeffs; /* device, ncols x 2^lg_n column-major, valid until done, not
                               inside the server's own memory (code 1 otherwise);
                               NULL: synthetic coefficients (zk_fill_random with seed) */
    uint64_t seed;
    uint32_t keep;          /* 1: LDE and digests stay on the device until zk_server_release */
    uint32_t reserved;
} zk_serve_request;

typedef struct {
    uint64_t submit_ns, admit_ns, done_ns; /* host steady clock */
    uint64_t *lde, *digests;               /* device, when kept (else NULL) */
    uint64_t device_bytes;                 /* server memory the request held */
} zk_serve_result;

RustError zk_server_create(size_t device_id, uint64_t budget_bytes, uint32_t max_inflight, zk_server **server);
/* Queues a request; cap_out (host, 4 << cap_height u64) is filled by the time
 * zk_server_wait returns for the ticket. */
RustError zk_server_submit(zk_server *server, const zk_serve_request *req, uint64_t *cap_out, uint64_t *ticket);
RustError zk_server_wait(zk_server *server, uint64_t ticket, zk_serve_result *result);
/* Frees a kept request's device memory. */
RustError zk_server_release(zk_server *server, uint64_t ticket);
/* Waits for every admitted and queued request, then frees the server and all
 * its memory (kept results included: their lde / digests pointers are invalid
 * afterwards). */
RustError zk_server_destroy(zk_server *server);

/* NTT plan tuning.  No counterpart in the reference, whose NTT has one fixed
 * stage structure.  The plan of a transform of 2^lg points (lg <= 30) is its
 * split into 1..3 HBM passes of at most 10 bits (bits[0..npasses)), each
 * pass's groups per workgroup (gbits, log2; -1: the most the pass allows)
 * and, for a 10-bit pass, whether the 16-element kernel may run it
 * (kernel16 = 1) or the 32-element kernel does (0).  The autotuner
 * (zeknox_amd/tune.py) measures candidate plans per size and writes the best
 * to ntt_tuning_gfx950.txt beside the library, which is loaded when a device
 * is first used (ZK_NTT_TUNING names another file).  Every plan computes the
 * same transform bit for bit.  zk_ntt_tune_set with npasses = 0 restores the
 * default plan; code 1 for a plan the passes cannot run.  Setup calls: not to
 * be made while transforms run on other threads. */
RustError zk_ntt_tune_set(uint32_t lg, uint32_t npasses, const uint32_t *bits, const int32_t *gbits,
                          const uint32_t *kernel16);
/* The tuned plan of 2^lg (npasses = 0: none, the default plan applies);
 * arrays of 3. */
RustError zk_ntt_tune_get(uint32_t lg, uint32_t *npasses, uint32_t *bits, int32_t *gbits, uint32_t *kernel16);

/* n field elements uniform on [0, p) from a counter-based generator:
 * element i = splitmix64 finaliser of seed + (i+1) * 0x9e3779b97f4a7c15,
 * re-mixed while >= p.  Device pointer. */
RustError zk_fill_random(size_t device_id, uint64_t *out, uint64_t n, uint64_t seed);

/* Per-kernel timing on the library's own stream (HIP events around every
 * launch).  Disabled by default; zk_profile_get reports kernel i's name,
 * accumulated milliseconds and launch count; returns code 1 past the end. */
void zk_profile_enable(int on);
void zk_profile_reset(void);
RustError zk_profile_get(int idx, char *name, size_t name_len, double *total_ms, uint64_t *launches);

#ifdef __cplusplus
}
#endif
#endif
