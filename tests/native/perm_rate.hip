// Issue-limited rate of the library's Poseidon permutation on gfx950: every
// lane runs the leaf kernel's sponge (hash_one) over a 135-element leaf whose
// elements it computes from its index (no memory traffic but the digest
// store), launched in the leaf-hash kernel's shape
// (workgroups of 128 lanes, at least four waves per SIMD, 2^20 lanes), so the
// rate is what the permutation's instruction stream reaches when nothing but
// VALU issue limits it, with as many waves per SIMD as the registers allow
// and, for comparison, with four or five (capped through LDS).  bench.py's
// leaf-hash roofline divides the leaf kernel's permutation rate by the first.  Prints permutations per second and
// SIMD-cycles per wave-permutation (256 CUs x 4 SIMDs at 2.4 GHz).
//
//   hipcc --offload-arch=gfx950 -O3 -std=c++17 -Xclang -target-feature -Xclang -packed-fp32-ops \
//         -o perm_rate tests/native/perm_rate.hip && ./perm_rate
#include <hip/hip_runtime.h>
#include <stdio.h>

#include "../../zeknox_amd/csrc/poseidon.hpp"

using namespace zk;

constexpr int kThreads = 128;
constexpr int kPerms = 17;  // permutations per lane: a 135-element leaf

// LDS bytes per workgroup caps the waves per SIMD (160 KiB per CU, two waves
// per workgroup): 0 leaves the register count as the only limit.
template <u32 LDS_BYTES>
__global__ __launch_bounds__(kThreads, 4) void perm_rate_kernel(u64 *out, u64 seed)
{
    const u64 t = (u64)blockIdx.x * kThreads + threadIdx.x;
    // the leaf kernel's sponge over a 135-element leaf, elements computed
    // from the lane index instead of loaded
    const u64 base = (t + seed) << 8;
    u64 h[4];
    hash_one([base](u32 i) { return base + i; }, 8 * kPerms - 1, h);
    u64 x = h[0] ^ h[1] ^ h[2] ^ h[3];
    if constexpr (LDS_BYTES > 0) {
        __shared__ u64 pad[LDS_BYTES / 8];
        pad[threadIdx.x] = x;
        __syncthreads();
        x ^= pad[(threadIdx.x + 1) % kThreads];
    }
    out[t] = x;
}

template <u32 LDS_BYTES>
static int run(u64 *out, u32 lanes, const char *label)
{
    const u32 blocks = lanes / kThreads;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    for (int w = 0; w < 3; w++)
        hipLaunchKernelGGL(perm_rate_kernel<LDS_BYTES>, dim3(blocks), dim3(kThreads), 0, 0, out, (u64)w);
    const int reps = 10;
    (void)hipEventRecord(e0);
    for (int r = 0; r < reps; r++)
        hipLaunchKernelGGL(perm_rate_kernel<LDS_BYTES>, dim3(blocks), dim3(kThreads), 0, 0, out, 7u + r);
    (void)hipEventRecord(e1);
    (void)hipEventSynchronize(e1);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        printf("error: %s\n", hipGetErrorString(e));
        return 1;
    }
    const double perms = (double)reps * lanes * kPerms, sec = ms * 1e-3;
    const double cycles_per_wave_perm = sec * 2.4e9 * 1024 / (perms / 64);
    printf("{\"occupancy\": \"%s\", \"lds_bytes_per_workgroup\": %u, \"permutations_per_s\": %.6g, "
           "\"simd_cycles_per_wave_permutation\": %.1f, \"ms_per_launch\": %.4f, \"lanes\": %u, "
           "\"permutations_per_lane\": %d}\n",
           label, LDS_BYTES, perms / sec, cycles_per_wave_perm, ms / reps, lanes, kPerms);
    return 0;
}

int main()
{
    const u32 lanes = 1u << 20;
    u64 *out = nullptr;
    if (hipMalloc(&out, (size_t)lanes * sizeof(u64)) != hipSuccess) return 1;
    // registers only (80 VGPRs: six waves per SIMD); LDS-capped to five and
    // four waves per SIMD (10 and 8 workgroups per CU)
    int rc = run<0>(out, lanes, "registers only");
    rc |= run<16384>(out, lanes, "5 waves per SIMD");
    rc |= run<20480>(out, lanes, "4 waves per SIMD");
    (void)hipFree(out);
    return rc;
}
