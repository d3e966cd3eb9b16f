// extern "C" entry points of libzeknox.so: the okx/zeknox lib.h / merkle.h C
// ABI (declared in include/lib.h, include/merkle/merkle.h) and the extensions
// of include/zeknox_ext.h.  No C++ exception crosses this boundary: failures
// become RustError{code, strdup(message)} (NTT family, native/ntt/ntt.cu:218-224)
// or a message on stderr (void Merkle family, native/utils/cuda_utils.cuh:18-29).
#include <stddef.h>
#include <stdio.h>
#include <string.h>
#include <time.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "batch.hpp"
#include "hashers.hpp"
#include "serve.hpp"
#include "merkle.hpp"
#include "ntt.hpp"
#include "poseidon.hpp"

#include "../../include/lib.h"
#include "../../include/zeknox_ext.h"

static_assert(sizeof(NTT_Config) == 24, "NTT_Config layout (native/ntt/ntt.h:43-55)");
static_assert(offsetof(NTT_Config, extension_rate_bits) == 12, "NTT_Config layout");
static_assert(offsetof(NTT_Config, are_inputs_on_device) == 16, "NTT_Config layout");
static_assert(offsetof(NTT_Config, salt_size) == 20, "NTT_Config layout");
static_assert(sizeof(NTT_TransposeConfig) == 8, "NTT_TransposeConfig layout (native/ntt/ntt.h:57-62)");
static_assert(sizeof(RustError) == 16, "RustError layout (native/utils/rusterror.h:15-41)");
static_assert(sizeof(zk_commit_job) == sizeof(zk::CommitJob), "zk_commit_job layout");
static_assert(offsetof(zk_commit_job, lg_n) == offsetof(zk::CommitJob, lg_n), "zk_commit_job layout");

using namespace zk;

namespace {

RustError ok() { return RustError{0, nullptr}; }

RustError err(int code, const std::string &msg)
{
    return RustError{code, msg.empty() ? nullptr : strdup(msg.c_str())};
}

template <typename F>
RustError guarded(F &&f)
{
    try {
        f();
        return ok();
    } catch (const Error &e) {
        return err(e.code, e.what());
    } catch (const std::exception &e) {
        return err(-(int)hipErrorUnknown, e.what());
    } catch (...) {
        return err(-(int)hipErrorUnknown, "unknown error");
    }
}

// Device ids cross the ABI as size_t, but the Rust wrapper passes an i32 for
// compute_transpose_rev (wrappers/rust/src/lib.rs:49-55); only the low 32 bits
// are meaningful.
inline size_t dev_id(size_t d) { return (size_t)(uint32_t)d; }

// Device buffer that either aliases a caller pointer or is owned (stream-ordered).
struct DevBuf {
    u64 *p = nullptr;
    bool owned = false;
    hipStream_t st = nullptr;
    DevBuf(void *caller, bool on_device, size_t elems, hipStream_t s) : st(s)
    {
        if (on_device) {
            p = (u64 *)caller;
        } else if (elems) {
            ZK_HIP(hipMallocAsync((void **)&p, elems * sizeof(u64), st));
            owned = true;
        }
    }
    ~DevBuf()
    {
        if (owned && p) (void)hipFreeAsync(p, st);
    }
};

void sync(DeviceCtx &c)
{
    ZK_HIP(hipStreamSynchronize(c.stream));
    prof_flush();
}

// LDE of `cols` columns (shared by compute_batched_lde and the multi-GPU entry)
void run_lde(size_t device_id, void *output, void *input, uint32_t lg_n, NTT_Direction dir, const NTT_Config &cfg)
{
    if (lg_n == 0 || cfg.extension_rate_bits < 1) throw Error((int)hipErrorInvalidValue, "invalid args");
    const u32 lg = lg_n + cfg.extension_rate_bits;
    if (lg > (u32)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
    DeviceCtx &c = ctx(dev_id(device_id));
    std::lock_guard<std::mutex> lk(c.mu);
    const size_t n = (size_t)1 << lg_n, N = (size_t)1 << lg;
    const size_t in_elems = n * cfg.batches, out_elems = N * ((size_t)cfg.batches + cfg.salt_size);
    DevBuf din(input, cfg.are_inputs_on_device, in_elems, c.stream);
    DevBuf dout(output, cfg.are_outputs_on_device, out_elems, c.stream);
    if (!cfg.are_inputs_on_device && in_elems)
        ZK_HIP(hipMemcpyAsync(din.p, input, in_elems * sizeof(u64), hipMemcpyHostToDevice, c.stream));
    NttJob j;
    j.src = din.p;
    j.dst = dout.p;
    j.src_stride = n;
    j.dst_stride = N;
    j.ncols = cfg.batches;
    j.lg = lg;
    j.lg_in = lg_n;
    j.inverse = dir == inverse;
    j.coset_in = cfg.with_coset && dir == forward;
    j.coset_out = cfg.with_coset && dir == inverse;
    j.scale_out = dir == inverse;
    j.prof_tag = "lde";
    j.use_scratch = true;
    ntt_launch(c, j, c.stream);
    if (cfg.salt_size > 0)  // time-seeded like the reference (ntt.cu:541-546)
        fill_random_launch(dout.p + N * cfg.batches, N * cfg.salt_size, (u64)time(nullptr), c.stream);
    if (!cfg.are_outputs_on_device && out_elems)
        ZK_HIP(hipMemcpyAsync(output, dout.p, out_elems * sizeof(u64), hipMemcpyDeviceToHost, c.stream));
    sync(c);
}

static bool hash_type_supported(u64 ht) { return ht == HashPoseidon || ht == HashKeccak || ht == HashPoseidon2; }

// Merkle forest on one device; errors are printed (the ABI returns void).
void run_merkle(const char *fn, size_t device_id, void *digests, void *cap, void *leaves, u64 n_digests,
                u64 n_caps, u64 n_leaves, u64 leaf_size, u64 cap_height, u64 hash_type)
{
    try {
        if (!hash_type_supported(hash_type))
            throw Error((int)hipErrorInvalidValue, "hash type " + std::to_string(hash_type) +
                                                       " is not implemented (Poseidon, Keccak, Poseidon2 are)");
        TreeShape t = tree_shape(n_digests, n_caps, n_leaves, leaf_size, cap_height);
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        merkle_launch_hashed(t, (u64 *)digests, (u64 *)cap, (const u64 *)leaves, LeafLayout::RowMajor, 0, c.stream,
                             (u32)hash_type);
        sync(c);
    } catch (const std::exception &e) {
        fprintf(stderr, "%s: %s\n", fn, e.what());
    }
}

template <typename F>
void parallel_for(u64 n, F &&f)
{
    unsigned nt = std::max(1u, std::thread::hardware_concurrency());
    if (n < 2 * (u64)nt) nt = 1;
    if (nt == 1) {
        for (u64 i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++)
        th.emplace_back([&, t] {
            for (u64 i = t; i < n; i += nt) f(i);
        });
    for (auto &x : th) x.join();
}

}  // namespace

extern "C" {

RustError get_number_of_gpus(size_t *ngpus)
{
    if (!ngpus) return err((int)hipErrorInvalidValue, "null pointer");
    *ngpus = (size_t)device_count();
    return ok();
}

RustError list_devices_info()
{
    return guarded([] {
        int n = device_count();
        for (int d = 0; d < n; d++) {
            hipDeviceProp_t p;
            ZK_HIP(hipGetDeviceProperties(&p, d));
            printf("Device %d - %s (%s)\n", d, p.name, p.gcnArchName);
            printf("compute units: %d, clock %d kHz, wavefront %d\n", p.multiProcessorCount, p.clockRate,
                   p.warpSize);
            printf("global memory: %zu MiB, LDS per workgroup: %zu B, L2: %d B\n", p.totalGlobalMem >> 20,
                   p.sharedMemPerBlock, p.l2CacheSize);
            printf("maxThreadsPerMultiProcessor: %d, managedMemory: %d, isMultiGpuBoard: %d\n\n",
                   p.maxThreadsPerMultiProcessor, p.managedMemory, p.isMultiGpuBoard);
        }
    });
}

bool cuda_available() { return device_count() > 0; }

void init_cuda_degree_and_generator(const uint32_t max_degree, const uint64_t group_generator)
{
    (void)max_degree;  // every size up to 2^32 is served from the same tables
    int n = device_count();
    for (int d = 0; d < n; d++) {
        try {
            DeviceCtx &c = ctx((size_t)d);
            std::lock_guard<std::mutex> lk(c.mu);
            c.set_coset(group_generator);
        } catch (const std::exception &e) {
            fprintf(stderr, "init_cuda_degree_and_generator: %s\n", e.what());
        }
    }
}

void init_cuda_degree(const uint32_t max_degree) { init_cuda_degree_and_generator(max_degree, 7); }

void init_cuda() { init_cuda_degree(24); }

RustError init_twiddle_factors(size_t device_id, size_t lg_n)
{
    return guarded([&] {
        if (lg_n > (size_t)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
        ctx(dev_id(device_id));
    });
}

RustError init_coset(size_t device_id, size_t lg_domain_size, const uint64_t coset_gen)
{
    return guarded([&] {
        if (lg_domain_size > (size_t)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        c.set_coset(coset_gen);
    });
}

RustError compute_batched_ntt(size_t device_id, void *inout, uint32_t lg_domain_size, NTT_Direction dir,
                              NTT_Config cfg)
{
    if (lg_domain_size == 0) return RustError{(int)hipErrorInvalidValue, nullptr};
    return guarded([&] {
        if (lg_domain_size > (u32)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        const size_t N = (size_t)1 << lg_domain_size, elems = N * cfg.batches;
        DevBuf d(inout, cfg.are_inputs_on_device, elems, c.stream);
        if (!cfg.are_inputs_on_device && elems)
            ZK_HIP(hipMemcpyAsync(d.p, inout, elems * sizeof(u64), hipMemcpyHostToDevice, c.stream));
        NttJob j;
        j.src = d.p;
        j.dst = d.p;
        j.src_stride = j.dst_stride = N;
        j.ncols = cfg.batches;
        j.lg = j.lg_in = lg_domain_size;
        j.inverse = dir == inverse;
        j.coset_in = cfg.with_coset && dir == forward;
        j.coset_out = cfg.with_coset && dir == inverse;
        j.scale_out = dir == inverse;
        j.prof_tag = dir == inverse ? "intt" : "ntt";
        j.use_scratch = true;
        ntt_launch(c, j, c.stream);
        // results land where the reference leaves them: the device buffer when
        // outputs stay on device, else copied back over the host input
        if (cfg.are_inputs_on_device && !cfg.are_outputs_on_device) {
            // reference semantics (ntt.cu:294-296): copy back into `inout`,
            // which is then the same device pointer; nothing to do
        } else if (!cfg.are_inputs_on_device && !cfg.are_outputs_on_device && elems) {
            ZK_HIP(hipMemcpyAsync(inout, d.p, elems * sizeof(u64), hipMemcpyDeviceToHost, c.stream));
        } else if (!cfg.are_inputs_on_device && cfg.are_outputs_on_device) {
            throw Error((int)hipErrorInvalidValue, "host input with device output is not supported in place");
        }
        sync(c);
    });
}

RustError compute_batched_lde(size_t device_id, void *output, void *input, uint32_t lg_domain_size,
                              NTT_Direction dir, NTT_Config cfg)
{
    if (lg_domain_size == 0 || cfg.extension_rate_bits < 1)
        return err((int)hipErrorInvalidValue, "invalid args");
    return guarded([&] { run_lde(device_id, output, input, lg_domain_size, dir, cfg); });
}

RustError compute_batched_lde_multi_gpu(void *output, void *input, uint32_t num_gpu, NTT_Direction dir,
                                        NTT_Config cfg, uint32_t lg_domain_size, size_t total_num_input_elements,
                                        size_t total_num_output_elements)
{
    (void)num_gpu;
    (void)total_num_output_elements;
    if (total_num_input_elements == 0 || lg_domain_size == 0 || cfg.extension_rate_bits < 1)
        return RustError{(int)hipErrorInvalidValue, nullptr};
    return guarded([&] {
        NTT_Config c1 = cfg;
        c1.batches = (uint32_t)(total_num_input_elements >> lg_domain_size);
        c1.is_multi_gpu = false;
        run_lde(0, output, input, lg_domain_size, dir, c1);
    });
}

RustError compute_transpose_rev(size_t device_id, void *output, void *input, uint32_t lg_n,
                                NTT_TransposeConfig cfg)
{
    return guarded([&] {
        if (lg_n > (u32)kMaxLg) throw Error((int)hipErrorInvalidValue, "invalid args");
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        const size_t elems = ((size_t)1 << lg_n) * cfg.batches;
        DevBuf din(input, cfg.are_inputs_on_device, elems, c.stream);
        DevBuf dout(output, cfg.are_outputs_on_device, elems, c.stream);
        if (!cfg.are_inputs_on_device && elems)
            ZK_HIP(hipMemcpyAsync(din.p, input, elems * sizeof(u64), hipMemcpyHostToDevice, c.stream));
        transpose_rev_launch(din.p, dout.p, lg_n, cfg.batches, c.stream);
        if (!cfg.are_outputs_on_device && elems)
            ZK_HIP(hipMemcpyAsync(output, dout.p, elems * sizeof(u64), hipMemcpyDeviceToHost, c.stream));
        sync(c);
    });
}

void fill_digests_buf_linear_gpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                              void *leaves_buf_gpu_ptr, u64 digests_buf_size, u64 cap_buf_size,
                                              u64 leaves_buf_size, u64 leaf_size, u64 cap_height, u64 hash_type,
                                              u64 gpu_id)
{
    run_merkle("fill_digests_buf_linear_gpu_with_gpu_ptr", (size_t)gpu_id, digests_buf_gpu_ptr, cap_buf_gpu_ptr,
               leaves_buf_gpu_ptr, digests_buf_size, cap_buf_size, leaves_buf_size, leaf_size, cap_height,
               hash_type);
}

void fill_digests_buf_linear_multigpu_with_gpu_ptr(void *digests_buf_gpu_ptr, void *cap_buf_gpu_ptr,
                                                   void *leaves_buf_gpu_ptr, u64 digests_buf_size,
                                                   u64 cap_buf_size, u64 leaves_buf_size, u64 leaf_size,
                                                   u64 cap_height, u64 hash_type)
{
    run_merkle("fill_digests_buf_linear_multigpu_with_gpu_ptr", 0, digests_buf_gpu_ptr, cap_buf_gpu_ptr,
               leaves_buf_gpu_ptr, digests_buf_size, cap_buf_size, leaves_buf_size, leaf_size, cap_height,
               hash_type);
}

// Host Merkle (merkle.cpp:126-161) with the same Poseidon code as the GPU
// kernels, parallel over leaves and per level with std::thread.
extern "C++" {
template <typename H>
static void host_forest(const TreeShape &t, u64 *dg, u64 *cp, const u64 *lv)
{
    const u32 ls = (u32)t.leaf_size;
    auto leaf_hash = [&](u64 l, u64 *out) {
        const u64 *p = lv + l * t.leaf_size;
        H::one([p](u32 i) { return p[i]; }, ls, out);
    };
    if (t.n_caps() == t.n_leaves) {
        parallel_for(t.n_leaves, [&](u64 l) { leaf_hash(l, cp + 4 * l); });
        return;
    }
    const u64 L = t.sub_leaves(), D = t.sub_digests();
    parallel_for(t.n_leaves, [&](u64 l) {
        const u64 s = l / L, i = l % L;
        leaf_hash(l, dg + 4 * (s * D + (D - L) + i));
    });
    u64 base = D - L;
    for (u64 nodes = L / 2; nodes >= 2; nodes /= 2) {
        base -= nodes;
        parallel_for(nodes * t.n_caps(), [&](u64 x) {
            const u64 s = x / nodes, node = base + x % nodes;
            u64 *sd = dg + 4 * s * D;
            H::two(sd + 4 * (2 * node + 2), sd + 4 * (2 * node + 3), sd + 4 * node);
        });
    }
    for (u64 s = 0; s < t.n_caps(); s++) H::two(dg + 4 * s * D, dg + 4 * s * D + 4, cp + 4 * s);
}
}  // extern "C++"

void fill_digests_buf_linear_cpu(void *digests_buf_ptr, void *cap_buf_ptr, const void *leaves_buf_ptr,
                                 u64 digests_buf_size, u64 cap_buf_size, u64 leaves_buf_size, u64 leaf_size,
                                 u64 cap_height, u64 hash_type)
{
    try {
        if (!hash_type_supported(hash_type))
            throw Error((int)hipErrorInvalidValue, "hash type " + std::to_string(hash_type) +
                                                       " is not implemented (Poseidon, Keccak, Poseidon2 are)");
        TreeShape t = tree_shape(digests_buf_size, cap_buf_size, leaves_buf_size, leaf_size, cap_height);
        u64 *dg = (u64 *)digests_buf_ptr, *cp = (u64 *)cap_buf_ptr;
        const u64 *lv = (const u64 *)leaves_buf_ptr;
        if (hash_type == HashKeccak)
            host_forest<KeccakHasher>(t, dg, cp, lv);
        else if (hash_type == HashPoseidon2)
            host_forest<Poseidon2Hasher>(t, dg, cp, lv);
        else
            host_forest<PoseidonHasher>(t, dg, cp, lv);
    } catch (const std::exception &e) {
        fprintf(stderr, "fill_digests_buf_linear_cpu: %s\n", e.what());
    }
}

// ------------------------------------------------------------- extensions
RustError zk_commit(size_t device_id, const uint64_t *coeffs, uint64_t *lde, uint64_t *leaves, uint64_t *digests,
                    uint64_t *cap, uint32_t lg_n, uint32_t rate_bits, uint32_t ncols, uint32_t cap_height,
                    uint64_t shift)
{
    return guarded([&] {
        const CommitJob j{coeffs, lde, leaves, digests, cap, lg_n, rate_bits, ncols, cap_height};
        commit_validate(j);
        if (gl_canon(shift) == 0) throw Error((int)hipErrorInvalidValue, "invalid args");
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        if (gl_canon(shift) != c.coset_gen) c.set_coset(shift);
        commit_launch(c, j, c.stream, true);
        sync(c);
    });
}

RustError zk_commit_noalloc(size_t device_id, const uint64_t *coeffs, uint64_t *lde, uint64_t *leaves,
                            uint64_t *digests, uint64_t *cap, uint32_t lg_n, uint32_t rate_bits, uint32_t ncols,
                            uint32_t cap_height, uint64_t shift)
{
    return guarded([&] {
        const CommitJob j{coeffs, lde, leaves, digests, cap, lg_n, rate_bits, ncols, cap_height};
        commit_validate(j);
        DeviceCtx &c = ctx_existing(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        // changing the coset would rebuild its table through host memory
        if (gl_canon(shift) != c.coset_gen)
            throw Error((int)hipErrorInvalidValue, "invalid args: shift differs from the device's coset generator "
                                                   "(set it with init_coset first)");
        commit_noalloc_launch(c, j, c.stream);
        sync(c);
    });
}

RustError zk_batch_commit(size_t device_id, const zk_commit_job *jobs, uint32_t njobs, uint32_t flags)
{
    return guarded([&] {
        if (njobs && !jobs) throw Error((int)hipErrorInvalidValue, "invalid args");
        for (uint32_t i = 0; i < njobs; i++) commit_validate(reinterpret_cast<const CommitJob *>(jobs)[i]);
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        batch_commit(c, reinterpret_cast<const CommitJob *>(jobs), njobs, (flags & ZK_BATCH_GRAPH) != 0);
        prof_flush();
    });
}

static_assert(sizeof(zk_serve_request) == sizeof(zk::ServeRequest), "zk_serve_request layout");
static_assert(sizeof(zk_serve_result) == sizeof(zk::ServeResult), "zk_serve_result layout");

RustError zk_server_create(size_t device_id, uint64_t budget_bytes, uint32_t max_inflight, zk_server **server)
{
    return guarded([&] {
        if (!server) throw Error((int)hipErrorInvalidValue, "invalid args");
        *server = nullptr;
        *server = reinterpret_cast<zk_server *>(server_create((int)dev_id(device_id), budget_bytes, max_inflight));
    });
}

RustError zk_server_submit(zk_server *server, const zk_serve_request *req, uint64_t *cap_out, uint64_t *ticket)
{
    return guarded([&] {
        if (!server || !req || !ticket) throw Error((int)hipErrorInvalidValue, "invalid args");
        *ticket = server_submit(reinterpret_cast<Server *>(server), *reinterpret_cast<const ServeRequest *>(req),
                                cap_out);
    });
}

RustError zk_server_wait(zk_server *server, uint64_t ticket, zk_serve_result *result)
{
    return guarded([&] {
        if (!server || !result) throw Error((int)hipErrorInvalidValue, "invalid args");
        server_wait(reinterpret_cast<Server *>(server), ticket, *reinterpret_cast<ServeResult *>(result));
    });
}

RustError zk_server_release(zk_server *server, uint64_t ticket)
{
    return guarded([&] {
        if (!server) throw Error((int)hipErrorInvalidValue, "invalid args");
        server_release(reinterpret_cast<Server *>(server), ticket);
    });
}

RustError zk_server_destroy(zk_server *server)
{
    return guarded([&] {
        if (server) server_destroy(reinterpret_cast<Server *>(server));
    });
}

RustError zk_merkle_proofs(size_t device_id, const uint64_t *digests, const uint64_t *leaves, uint64_t col_stride,
                           uint64_t n_leaves, uint64_t leaf_size, uint64_t cap_height, const uint64_t *indices,
                           uint64_t n_indices, uint64_t *siblings_out, uint64_t *rows_out)
{
    return guarded([&] {
        const u64 ncaps = cap_height < 63 ? (u64)1 << cap_height : 0;
        if (!ncaps || ncaps > n_leaves) throw Error((int)hipErrorInvalidValue, "invalid args");
        const TreeShape t = tree_shape(ncaps < n_leaves ? 2 * (n_leaves - ncaps) : 0, ncaps, n_leaves,
                                       leaf_size ? leaf_size : 1, cap_height);
        const bool has_layers = ncaps < n_leaves;
        if (n_indices && (!indices || (has_layers && (!digests || !siblings_out)) || (rows_out && !leaves)))
            throw Error((int)hipErrorInvalidValue, "invalid args");
        if (rows_out && col_stride && (col_stride < n_leaves || n_leaves > ((u64)1 << 32)))
            throw Error((int)hipErrorInvalidValue, "invalid args");
        for (uint64_t i = 0; i < n_indices; i++)
            if (indices[i] >= n_leaves) throw Error((int)hipErrorInvalidValue, "leaf index out of range");
        if (!n_indices) return;
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        u64 *di = (u64 *)c.get_aux(n_indices * sizeof(u64));
        ZK_HIP(hipMemcpyAsync(di, indices, n_indices * sizeof(u64), hipMemcpyHostToDevice, c.stream));
        proofs_launch(t, digests, leaves, col_stride, di, n_indices, siblings_out, rows_out, c.stream);
        sync(c);
    });
}

RustError zk_ntt_tune_set(uint32_t lg, uint32_t npasses, const uint32_t *bits, const int32_t *gbits,
                          const uint32_t *kernel16)
{
    return guarded([&] {
        if (npasses > 3 || (npasses && (!bits || !gbits || !kernel16)))
            throw Error((int)hipErrorInvalidValue, "invalid args");
        const u32 zb[3] = {0, 0, 0}, zr[3] = {0, 0, 0};
        const int32_t zg[3] = {-1, -1, -1};
        if (!ntt_tune_set(lg, npasses, npasses ? bits : zb, npasses ? gbits : zg, npasses ? kernel16 : zr))
            throw Error((int)hipErrorInvalidValue, "invalid args: not a runnable NTT plan");
    });
}

RustError zk_ntt_tune_get(uint32_t lg, uint32_t *npasses, uint32_t *bits, int32_t *gbits, uint32_t *kernel16)
{
    return guarded([&] {
        if (!npasses || !bits || !gbits || !kernel16 || lg == 0 || lg > 30)
            throw Error((int)hipErrorInvalidValue, "invalid args");
        ntt_tune_load_once();
        ntt_tune_read(lg, *npasses, bits, gbits, kernel16);
    });
}

RustError zk_batch_clear(size_t device_id)
{
    return guarded([&] {
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        batch_clear(c);
    });
}

RustError zk_fill_random(size_t device_id, uint64_t *out, uint64_t n, uint64_t seed)
{
    return guarded([&] {
        DeviceCtx &c = ctx(dev_id(device_id));
        std::lock_guard<std::mutex> lk(c.mu);
        fill_random_launch(out, n, seed, c.stream);
        sync(c);
    });
}

void zk_profile_enable(int on) { prof_set(on != 0); }

void zk_profile_reset(void) { prof_reset(); }

RustError zk_profile_get(int idx, char *name, size_t name_len, double *total_ms, uint64_t *launches)
{
    std::string n;
    double ms;
    uint64_t l;
    if (!prof_get(idx, n, ms, l)) return RustError{1, nullptr};
    if (name && name_len) {
        strncpy(name, n.c_str(), name_len - 1);
        name[name_len - 1] = 0;
    }
    if (total_ms) *total_ms = ms;
    if (launches) *launches = l;
    return ok();
}

}  // extern "C"
