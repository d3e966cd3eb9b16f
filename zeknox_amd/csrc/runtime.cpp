// Per-device runtime: contexts, tables, scratch.  See runtime.hpp.
#include "runtime.hpp"

#include <memory>
#include <vector>

namespace zk {

static std::mutex g_mu;
static std::vector<std::unique_ptr<DeviceCtx>> g_ctx;

int device_count()
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Powers x^i for i < kPowLen, computed on the host (exact, canonical).
static void fill_powers(u64 x, u64 *row)
{
    u64 v = 1;
    for (int i = 0; i < kPowLen; i++) {
        row[i] = v;
        v = gl_mul(v, x);
    }
}

void DeviceCtx::set_coset(u64 g)
{
    std::vector<u64> h(3 * kPowLen);
    u64 base = gl_canon(g);
    for (int k = 0; k < 3; k++) {
        fill_powers(base, h.data() + k * kPowLen);
        base = gl_pow(base, (u64)kPowLen);  // g^(2^(11(k+1)))
    }
    ZK_HIP(hipMemcpyAsync(gpw, h.data(), h.size() * sizeof(u64), hipMemcpyHostToDevice, stream));
    ZK_HIP(hipStreamSynchronize(stream));
    coset_gen = gl_canon(g);
    coset_gen_inv = gl_inv(coset_gen);
}

void *DeviceCtx::get_scratch(size_t bytes)
{
    if (bytes > scratch_bytes) {
        if (scratch) {
            ZK_HIP(hipStreamSynchronize(stream));
            ZK_HIP(hipFree(scratch));
            scratch = nullptr;
            scratch_bytes = 0;
        }
        ZK_HIP(hipMalloc(&scratch, bytes));
        scratch_bytes = bytes;
    }
    return scratch;
}

void *DeviceCtx::get_aux(size_t bytes)
{
    if (bytes > aux_bytes) {
        if (aux) {
            ZK_HIP(hipStreamSynchronize(stream));
            ZK_HIP(hipFree(aux));
            aux = nullptr;
            aux_bytes = 0;
        }
        ZK_HIP(hipMalloc(&aux, bytes));
        aux_bytes = bytes;
    }
    return aux;
}

void ntt_tune_load_once();  // ntt.hip

static std::unique_ptr<DeviceCtx> make_ctx(int dev)
{
    ntt_tune_load_once();
    auto c = std::make_unique<DeviceCtx>();
    c->dev = dev;
    ZK_HIP(hipSetDevice(dev));
    // Blocking w.r.t. the null stream, like the reference's per-GPU "zero"
    // stream: work the caller queued on the default stream completes first.
    ZK_HIP(hipStreamCreate(&c->stream));
    int lds_max = 0;
    if (hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess &&
        lds_max >= 2 * 66560)
        c->lds_budget = 66560;  // 64 KiB tiles plus one element of padding per row
    std::vector<u64> h((size_t)(kMaxLg + 1) * kPowLen);
    for (int dir = 0; dir < 2; dir++) {
        for (int L = 0; L <= kMaxLg; L++) {
            u64 w = gl_omega((u32)L);
            if (dir) w = gl_inv(w);
            fill_powers(w, h.data() + (size_t)L * kPowLen);
        }
        ZK_HIP(hipMalloc(&c->pw[dir], h.size() * sizeof(u64)));
        ZK_HIP(hipMemcpy(c->pw[dir], h.data(), h.size() * sizeof(u64), hipMemcpyHostToDevice));
        std::vector<u64> tb(16 * 64);
        const u64 w = h[(size_t)10 * kPowLen + 1];  // omega_1024 (or its inverse)
        for (u32 v = 0; v < 16; v++) {
            const u64 wv = gl_pow(w, zk_bitrev(v, 4));
            u64 x = 1;
            for (u32 j = 0; j < 64; j++) {
                tb[64 * v + j] = x;
                x = gl_mul(x, wv);
            }
        }
        ZK_HIP(hipMalloc(&c->r10tb[dir], tb.size() * sizeof(u64)));
        ZK_HIP(hipMemcpy(c->r10tb[dir], tb.data(), tb.size() * sizeof(u64), hipMemcpyHostToDevice));
    }
    ZK_HIP(hipMalloc(&c->gpw, 3 * kPowLen * sizeof(u64)));
    c->set_coset(7);
    fixed_scratch_setup(*c);
    return c;
}

DeviceCtx &ctx(size_t dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int n = device_count();
    if (n <= 0) throw Error(-(int)hipErrorNoDevice, "no HIP device available");
    if (dev >= (size_t)n)
        throw Error((int)hipErrorInvalidDevice, "invalid device id " + std::to_string(dev));
    if (g_ctx.size() < (size_t)n) g_ctx.resize(n);
    if (!g_ctx[dev]) g_ctx[dev] = make_ctx((int)dev);
    ZK_HIP(hipSetDevice((int)dev));
    return *g_ctx[dev];
}

DeviceCtx &ctx_existing(size_t dev)
{
    std::lock_guard<std::mutex> lk(g_mu);
    if (dev >= g_ctx.size() || !g_ctx[dev])
        throw Error((int)hipErrorInvalidValue, "invalid args: device " + std::to_string(dev) +
                                                   " is not initialised (call init_cuda or any other entry "
                                                   "point on it first)");
    ZK_HIP(hipSetDevice((int)dev));
    return *g_ctx[dev];
}

}  // namespace zk

// ------------------------------------------------------------- profiling
#include <map>

namespace zk {

struct ProfEntry {
    double ms = 0;
    uint64_t launches = 0;
};
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::map<std::string, ProfEntry> g_prof;
struct PendingPair {
    std::string name;
    hipEvent_t a, b;
};
static std::vector<PendingPair> g_pending;

bool prof_enabled() { return g_prof_on; }
void prof_set(bool on) { g_prof_on = on; }

ProfScope::ProfScope(const char *n, hipStream_t s) : name(n), st(s)
{
    if (!g_prof_on) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        a = b = nullptr;
        return;
    }
    hipEventRecord(a, st);
}

ProfScope::~ProfScope()
{
    if (!a) return;
    hipEventRecord(b, st);
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_pending.push_back({name, a, b});
}

void prof_flush()
{
    std::lock_guard<std::mutex> lk(g_prof_mu);
    std::vector<PendingPair> keep;
    for (auto &p : g_pending) {
        if (hipEventQuery(p.b) != hipSuccess) {
            keep.push_back(p);
            continue;
        }
        float ms = 0;
        hipEventElapsedTime(&ms, p.a, p.b);
        auto &e = g_prof[p.name];
        e.ms += ms;
        e.launches++;
        hipEventDestroy(p.a);
        hipEventDestroy(p.b);
    }
    g_pending.swap(keep);
}

void prof_reset()
{
    prof_flush();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.clear();
}

bool prof_get(int idx, std::string &name, double &ms, uint64_t &launches)
{
    prof_flush();
    std::lock_guard<std::mutex> lk(g_prof_mu);
    if (idx < 0 || (size_t)idx >= g_prof.size()) return false;
    auto it = g_prof.begin();
    std::advance(it, idx);
    name = it->first;
    ms = it->second.ms;
    launches = it->second.launches;
    return true;
}

}  // namespace zk
