// Per-device runtime for the zeknox C ABI on MI355X: device selection, the
// library stream, device-resident root-of-unity and coset power tables, a
// grow-only scratch buffer, and the error type carried back to the ABI.
//
// Replaces the reference's gpu_t / all_gpus / static twiddle arrays
// (okx/zeknox native/utils/gpu_t.cuh:237-342, native/utils/all_gpus.cu:11-66,
// native/ntt/ntt.cu:25-27).  Differences: every table is built once per device
// at first use for all sizes up to 2^32 (so init_twiddle_factors / init_coset
// are cheap and never required), and all state is guarded by a mutex.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <stdexcept>
#include <string>

#include "gl64.hpp"

namespace zk {

// Number of entries per power table: PW[L][i] = omega_{2^L}^i for i < 2^11.
constexpr int kPowBits = 11;
constexpr int kPowLen = 1 << kPowBits;
constexpr int kMaxLg = 32;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ZK_HIP(expr)                                                                     \
    do {                                                                                 \
        hipError_t _e = (expr);                                                          \
        if (_e != hipSuccess)                                                            \
            throw ::zk::Error(-(int)_e, std::string(#expr "@" __FILE__ ":") +            \
                                            std::to_string(__LINE__) + " failed: " +     \
                                            hipGetErrorString(_e));                      \
    } while (0)

struct DeviceCtx {
    int dev = -1;
    hipStream_t stream = nullptr;
    std::mutex mu;  // serialises ABI calls on this device
    // pw[dir]: (kMaxLg+1) x kPowLen table, row L = powers of omega_{2^L} (dir 0)
    // or omega_{2^L}^-1 (dir 1).
    u64 *pw[2] = {nullptr, nullptr};
    // r10tb[dir][64 v + j] = omega_1024^(rev4(v) j) (v < 16, j < 64), dir 1
    // with the inverse root: the middle twiddles of the 10-bit NTT pass
    u64 *r10tb[2] = {nullptr, nullptr};
    // coset powers: row k (k = 0,1,2) holds g^(i * 2^(11k)), i < 2^11
    u64 *gpw = nullptr;
    u64 coset_gen = 7;
    u64 coset_gen_inv = 0;
    // dynamic LDS bytes one NTT workgroup may use (two workgroups per CU)
    size_t lds_budget = 65536;
    // grow-only scratch (NTT ping-pong buffer)
    void *scratch = nullptr;
    size_t scratch_bytes = 0;
    // allocation-free commits: this device's address of the library's fixed
    // scratch array (noalloc.hip), looked up once when the context is made;
    // null with fixed_scratch_err set when the lookup failed
    u64 *fixed_scratch = nullptr;
    size_t fixed_scratch_bytes = 0;
    std::string fixed_scratch_err;

    // grow-only staging for small host-provided arrays (e.g. query indices)
    void *aux = nullptr;
    size_t aux_bytes = 0;

    // inter-pass twiddle tables of the NTT passes after the first, built on
    // first use and kept for the process (see ntt.hip); guarded by tw_mu
    std::mutex tw_mu;
    std::map<uint64_t, u64 *> tw_tables;

    void *get_scratch(size_t bytes);
    void *get_aux(size_t bytes);
    void set_coset(u64 g);
};

int device_count();

// Optional per-kernel timing: HIP events recorded on the launch stream around
// each kernel while enabled; resolved after the stream has synchronised..
struct ProfScope {
    hipEvent_t a = nullptr, b = nullptr;
    const char *name;
    hipStream_t st;
    ProfScope(const char *name, hipStream_t st);
    ~ProfScope();
};
bool prof_enabled();
void prof_set(bool on);
void prof_reset();
void prof_flush();  // after a stream sync: accumulate completed event pairs
bool prof_get(int idx, std::string &name, double &ms, uint64_t &launches);
// Selects `dev` (hipSetDevice) and returns its context, created on first use.
DeviceCtx &ctx(size_t dev);
// The same for a context that already exists; throws Error(invalid args)
// instead of creating one (creating a context allocates its tables).
DeviceCtx &ctx_existing(size_t dev);
// Looks up this device's address of the fixed scratch array (noalloc.hip);
// called by ctx() with the device current, once per device.
void fixed_scratch_setup(DeviceCtx &c);

}  // namespace zk
