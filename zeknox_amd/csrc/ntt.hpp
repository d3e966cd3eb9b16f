// Host-side interface of the NTT / LDE / transpose kernels (ntt.hip).
#pragma once
#include "runtime.hpp"

namespace zk {

struct NttJob {
    const u64 *src = nullptr;  // column c at src + c * src_stride
    u64 *dst = nullptr;        // column c at dst + c * dst_stride
    u64 src_stride = 0, dst_stride = 0;
    u32 ncols = 0;
    u32 lg = 0;         // transform size N = 2^lg
    u32 lg_in = 0;      // input columns hold 2^lg_in values, the rest is zero (LDE)
    bool inverse = false;
    bool coset_in = false;   // forward coset: x_n *= g^n before the transform
    bool coset_out = false;  // reference inverse-coset quirk: X_k *= g^k after
    bool scale_out = false;  // X_k *= 2^-lg (inverse)
    const char *prof_tag = "ntt";  // "ntt" / "intt" / "lde": kernel timing names
    // the transform may keep its intermediate in the library scratch buffer
    // (only for callers serialised on the device context: the scratch is
    // shared)
    bool use_scratch = false;
    // coset power rows (3 x 2^11 u64, device) to use instead of the device
    // context's, e.g. a commit server's snapshot (nullptr: the context's)
    const u64 *gpw = nullptr;
    // allocation-free launch: no library scratch, no new twiddle table (an
    // existing one is used, else the passes compute their twiddles); a
    // two-pass transform that keeps a blocked intermediate keeps it in
    // `inter` (inter_bytes long, at least ncols * 2^lg * 8 bytes) when given
    bool no_alloc = false;
    u64 *inter = nullptr;
    size_t inter_bytes = 0;
};

// Runs the transform on ctx.stream.  When src == dst (in place) and more than
// one pass is needed, the library scratch buffer is used for the first pass.
void ntt_launch(DeviceCtx &c, const NttJob &job, hipStream_t stream);

// out[rev(i)][j] = in[j][i] for an in: ncols x 2^lg column-major matrix.
void transpose_rev_launch(const u64 *in, u64 *out, u32 lg, u64 ncols, hipStream_t stream);

// Seeded synthetic field elements (uniform on [0, p)), counter based.
void fill_random_launch(u64 *out, u64 n, u64 seed, hipStream_t stream);

// Number of HBM passes the transform of size 2^lg takes and each pass's bits.
int ntt_plan(u32 lg, u32 bits[4]);

// Loads the NTT plan table chosen by the autotuner (zeknox_amd/tune.py), once
// per process: the file named by ZK_NTT_TUNING, else ntt_tuning_gfx950.txt in
// the directory of libzeknox.so; without one the default plans apply.  Lines
// "lg np b1 b2 b3 g1 g2 g3 r1 r2 r3" (g < 0: auto), '#' starts a comment;
// entries that are not a runnable plan are skipped with a message on stderr.
void ntt_tune_load_once();

// Sets (np > 0) or clears (np == 0) the plan of size 2^lg; false if the entry
// is not a runnable plan.
bool ntt_tune_set(u32 lg, u32 np, const u32 bits[3], const int32_t gbits[3], const u32 r10[3]);

// The tuned plan of size 2^lg (np = 0: none); gbits -1 for auto.
void ntt_tune_read(u32 lg, u32 &np, u32 bits[3], int32_t gbits[3], u32 r10[3]);

// Whether a transform of size 2^lg (not in place) keeps a blocked
// intermediate when one is available: two 10-bit passes.
bool ntt_blocked_intermediate(u32 lg);

}  // namespace zk
