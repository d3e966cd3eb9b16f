// Commit server: continuous batching of commit requests on one GPU.
//
// A worker thread owns the GPU side.  Requests wait in FIFO order until the
// number in flight is below max_inflight and the device-memory arena (one
// allocation of `budget` bytes, made at creation) has room for the request's
// coefficients (synthetic ones), LDE, digests and cap.  Whatever is admitted
// at one moment forms a group (at most max_inflight / 2 requests, so the next
// group can be admitted while one runs): each request's coset LDE goes to the
// least-loaded of kServeLanes streams, and the group's Merkle forests run as
// fused launches on one of kForestStreams streams once those LDEs are done
// (forest_launch_tables: one leaf-hash launch over every leaf of the group and
// one launch per tree height, as the batched commits do), followed by the cap
// copies and one event per request.  A group of one request runs the single-
// commit kernels instead (commit_launch: quad-lane leaves and tree tops for
// small trees, which cut its latency).  Groups admitted while others run
// overlap with them on the device.  The worker polls the events, copies each
// finished cap out of a pinned staging slot, returns the memory (unless the
// caller keeps the LDE and digests resident for openings) and admits whatever
// now fits.  The root of unity power tables are the device context's
// (immutable once built); the coset power rows are copied into server-owned
// device memory at creation, so an init_coset or a zk_commit with another
// shift on another thread cannot change the generator under requests in
// flight.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "arena.hpp"
#include "commit.hpp"
#include "forest_map.hpp"
#include "merkle.hpp"
#include "ntt.hpp"
#include "serve.hpp"

namespace zk {

constexpr int kServeLanes = 4;
constexpr int kForestStreams = 2;
constexpr u32 kMaxGroup = 16;  // requests per group
// Forest tables of one group: at most 3 G + 32 launches (two leaf launches
// and one per tree height, lg <= 30, plus the splits of launches beyond
// kForestMaxBlocks workgroups) of at most G entries of <= 72 bytes, each
// table 256-byte aligned: 112 KiB for G = 16.
constexpr size_t kGroupTableBytes = 128u << 10;
constexpr u32 kMaxServeCapHeight = 12;
constexpr u64 kCapSlotElems = 4ull << kMaxServeCapHeight;

static uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

static u64 align256(u64 x) { return (x + 255) & ~(u64)255; }

struct Req {
    enum State { Pending, Running, Done, Failed };
    uint64_t id = 0;
    ServeRequest r{};
    uint64_t *cap_out = nullptr;
    u64 bytes = 0, off = 0;
    bool holds_memory = false;
    u64 *coeffs = nullptr, *lde = nullptr, *digests = nullptr, *cap = nullptr;
    int lane = -1, slot = -1, group = -1;
    hipEvent_t done = nullptr, lde_done = nullptr;
    uint64_t submit_ns = 0, admit_ns = 0, done_ns = 0;
    State state = Pending;
    std::string error;
};

class Server {
public:
    Server(int dev, uint64_t budget, uint32_t max_inflight);
    ~Server();
    uint64_t submit(const ServeRequest &r, uint64_t *cap_out);
    void wait(uint64_t ticket, ServeResult &res);
    void release(uint64_t ticket);

private:
    void run();
    void place(Req *q);
    void launch_group(const std::vector<Req *> &g, int gslot);
    void finish(Req *q, hipError_t e);
    static u64 request_bytes(const ServeRequest &r);

    DeviceCtx &c_;
    u64 coset_;        // the coset generator, fixed at creation
    u64 *gpw_ = nullptr;  // server-owned copy of its 3 x 2^11 coset power rows
    uint64_t budget_;
    uint32_t max_inflight_;
    uint8_t *base_ = nullptr;
    Arena arena_;
    hipStream_t lanes_[kServeLanes] = {};
    uint32_t lane_load_[kServeLanes] = {};
    hipStream_t forest_[kForestStreams] = {};
    u32 next_forest_ = 0;
    u64 *pinned_ = nullptr;
    std::vector<int> free_slots_;
    std::vector<hipEvent_t> free_events_;
    // groups in flight: per slot a device and a pinned host table buffer and
    // the number of its requests not yet finished
    u32 group_max_ = 1;
    std::vector<void *> table_dev_, table_host_;
    std::vector<u32> group_left_;
    std::vector<int> free_groups_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Req *> pending_;
    std::vector<Req *> running_;
    std::unordered_map<uint64_t, std::unique_ptr<Req>> reqs_;
    uint64_t next_id_ = 1;
    bool stop_ = false;
    std::thread worker_;
};

Server::Server(int dev, uint64_t budget, uint32_t max_inflight)
    : c_(ctx((size_t)dev)), budget_(budget), max_inflight_(max_inflight), arena_(budget)
{
    if (budget == 0 || max_inflight == 0 || max_inflight > 1024)
        throw Error((int)hipErrorInvalidValue, "invalid args");
    try {
        ZK_HIP(hipMalloc(&gpw_, 3 * kPowLen * sizeof(u64)));
        {
            // snapshot of the device's coset rows, taken under the context lock
            std::lock_guard<std::mutex> lk(c_.mu);
            coset_ = c_.coset_gen;
            ZK_HIP(hipMemcpyAsync(gpw_, c_.gpw, 3 * kPowLen * sizeof(u64), hipMemcpyDeviceToDevice, c_.stream));
            ZK_HIP(hipStreamSynchronize(c_.stream));
        }
        ZK_HIP(hipMalloc(&base_, budget));
        ZK_HIP(hipHostMalloc(&pinned_, (size_t)max_inflight * kCapSlotElems * sizeof(u64), hipHostMallocDefault));
        for (int l = 0; l < kServeLanes; l++) ZK_HIP(hipStreamCreateWithFlags(&lanes_[l], hipStreamNonBlocking));
        for (int l = 0; l < kForestStreams; l++) ZK_HIP(hipStreamCreateWithFlags(&forest_[l], hipStreamNonBlocking));
        for (uint32_t i = 0; i < 2 * max_inflight; i++) {  // a done and an LDE event per request
            hipEvent_t e;
            ZK_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
            free_events_.push_back(e);
        }
        for (uint32_t i = 0; i < max_inflight; i++) free_slots_.push_back((int)i);
        group_max_ = std::max<u32>(1, std::min<u32>(kMaxGroup, max_inflight / 2));
        // groups of two or more requests: at most max_inflight / 2 in flight
        const u32 ngroups = std::max<u32>(1, max_inflight / 2);
        for (u32 i = 0; i < ngroups; i++) {
            void *d = nullptr, *h = nullptr;
            ZK_HIP(hipMalloc(&d, kGroupTableBytes));
            table_dev_.push_back(d);
            ZK_HIP(hipHostMalloc(&h, kGroupTableBytes, hipHostMallocDefault));
            table_host_.push_back(h);
            free_groups_.push_back((int)i);
        }
        group_left_.assign(ngroups, 0);
    } catch (...) {
        for (auto e : free_events_) (void)hipEventDestroy(e);
        for (auto s : lanes_)
            if (s) (void)hipStreamDestroy(s);
        for (auto s : forest_)
            if (s) (void)hipStreamDestroy(s);
        for (auto d : table_dev_) (void)hipFree(d);
        for (auto h : table_host_) (void)hipHostFree(h);
        if (pinned_) (void)hipHostFree(pinned_);
        if (base_) (void)hipFree(base_);
        if (gpw_) (void)hipFree(gpw_);
        throw;
    }
    worker_ = std::thread([this] { run(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        // kept results are dropped with the server: their memory may be what
        // a queued request is waiting for
        for (auto &kv : reqs_) {
            Req *q = kv.second.get();
            if (q->state == Req::Done && q->holds_memory) {
                arena_.release(q->off, q->bytes);
                q->holds_memory = false;
            }
        }
    }
    cv_work_.notify_all();
    worker_.join();  // returns once nothing is pending or running
    for (auto e : free_events_) (void)hipEventDestroy(e);
    for (auto s : lanes_) (void)hipStreamDestroy(s);
    for (auto s : forest_) (void)hipStreamDestroy(s);
    for (auto d : table_dev_) (void)hipFree(d);
    for (auto h : table_host_) (void)hipHostFree(h);
    (void)hipHostFree(pinned_);
    (void)hipFree(base_);
    (void)hipFree(gpw_);
}

u64 Server::request_bytes(const ServeRequest &r)
{
    const u64 n = (u64)1 << r.lg_n, N = (u64)1 << (r.lg_n + r.rate_bits), ncaps = (u64)1 << r.cap_height;
    const u64 nd = ncaps < N ? 2 * (N - ncaps) : 0;
    return (r.coeffs ? 0 : align256(8 * (u64)r.ncols * n)) + align256(8 * (u64)r.ncols * N) + align256(32 * nd) +
           align256(32 * ncaps);
}

uint64_t Server::submit(const ServeRequest &r, uint64_t *cap_out)
{
    const u32 lg = r.lg_n + r.rate_bits;
    if (!cap_out || r.lg_n == 0 || lg > 30 || r.ncols == 0 || r.cap_height > lg || r.cap_height > kMaxServeCapHeight)
        throw Error((int)hipErrorInvalidValue, "invalid args");
    // caller coefficients must not live in the server's own memory: a range
    // the arena hands out again could be overwritten while the request waits
    if (r.coeffs) {
        const uint8_t *cb = reinterpret_cast<const uint8_t *>(r.coeffs);
        const u64 cbytes = 8 * (u64)r.ncols << r.lg_n;
        if (cb < base_ + budget_ && base_ < cb + cbytes)
            throw Error((int)hipErrorInvalidValue, "invalid args: coefficients inside the server's memory");
    }
    const u64 bytes = request_bytes(r);
    if (arena_.round(bytes) > budget_)
        throw Error((int)hipErrorOutOfMemory, "request larger than the server's memory budget");
    auto q = std::make_unique<Req>();
    q->r = r;
    q->cap_out = cap_out;
    q->bytes = bytes;
    q->submit_ns = now_ns();
    std::lock_guard<std::mutex> lk(mu_);
    if (stop_) throw Error((int)hipErrorInvalidValue, "server is shutting down");
    q->id = next_id_++;
    pending_.push_back(q.get());
    const uint64_t id = q->id;
    reqs_.emplace(id, std::move(q));
    cv_work_.notify_one();
    return id;
}

// Called by the worker with mu_ held: the request's buffers in its arena
// range, a cap staging slot, its events and its LDE lane.
void Server::place(Req *q)
{
    const ServeRequest &r = q->r;
    const u64 n = (u64)1 << r.lg_n, N = (u64)1 << (r.lg_n + r.rate_bits), ncaps = (u64)1 << r.cap_height;
    const u64 nd = ncaps < N ? 2 * (N - ncaps) : 0;
    uint8_t *p = base_ + q->off;
    if (r.coeffs) {
        q->coeffs = const_cast<u64 *>(r.coeffs);
    } else {
        q->coeffs = (u64 *)p;
        p += align256(8 * (u64)r.ncols * n);
    }
    q->lde = (u64 *)p;
    p += align256(8 * (u64)r.ncols * N);
    q->digests = nd ? (u64 *)p : nullptr;
    p += align256(32 * nd);
    q->cap = (u64 *)p;
    int lane = 0;
    for (int l = 1; l < kServeLanes; l++)
        if (lane_load_[l] < lane_load_[lane]) lane = l;
    q->slot = free_slots_.back();
    free_slots_.pop_back();
    q->done = free_events_.back();
    free_events_.pop_back();
    q->lde_done = free_events_.back();
    free_events_.pop_back();
    q->lane = lane;
    lane_load_[lane]++;
    q->admit_ns = now_ns();
    q->state = Req::Running;
}

// Called by the worker with mu_ held: enqueues a group (see the top of the
// file).  gslot: the group's table buffers (groups of two or more).
void Server::launch_group(const std::vector<Req *> &g, int gslot)
{
    std::vector<ForestJob> jobs;
    for (Req *q : g) {
        place(q);
        const ServeRequest &r = q->r;
        const hipStream_t st = lanes_[q->lane];
        if (!r.coeffs) fill_random_launch(q->coeffs, (u64)r.ncols << r.lg_n, r.seed, st);
        const CommitJob j{q->coeffs, q->lde, nullptr, q->digests, q->cap, r.lg_n, r.rate_bits, r.ncols, r.cap_height};
        if (g.size() == 1) {
            commit_launch(c_, j, st, false, gpw_);
            ZK_HIP(hipMemcpyAsync(pinned_ + (u64)q->slot * kCapSlotElems, q->cap, 32ull << r.cap_height,
                                  hipMemcpyDeviceToHost, st));
            ZK_HIP(hipEventRecord(q->done, st));
            return;
        }
        lde_launch(c_, j, st, false, gpw_);
        ZK_HIP(hipEventRecord(q->lde_done, st));
        jobs.push_back(ForestJob{q->lde, nullptr, q->digests, q->cap, r.lg_n + r.rate_bits, r.ncols, r.cap_height});
    }
    std::vector<unsigned char> bytes;
    std::vector<ForestLaunch> launches;
    if (!forest_tables(jobs, bytes, launches)) throw Error((int)hipErrorInvalidValue, "invalid forest job");
    if (bytes.size() > kGroupTableBytes) throw Error((int)hipErrorInvalidValue, "forest tables too large");
    const hipStream_t fs = forest_[next_forest_++ % kForestStreams];
    memcpy(table_host_[gslot], bytes.data(), bytes.size());
    ZK_HIP(hipMemcpyAsync(table_dev_[gslot], table_host_[gslot], bytes.size(), hipMemcpyHostToDevice, fs));
    for (Req *q : g) ZK_HIP(hipStreamWaitEvent(fs, q->lde_done, 0));
    forest_launch_tables(launches, table_dev_[gslot], fs);
    for (Req *q : g) {
        ZK_HIP(hipMemcpyAsync(pinned_ + (u64)q->slot * kCapSlotElems, q->cap, 32ull << q->r.cap_height,
                              hipMemcpyDeviceToHost, fs));
        ZK_HIP(hipEventRecord(q->done, fs));
    }
}

// Called by the worker with mu_ held, once the request's event has resolved
// (or its launch failed: e != hipSuccess).
void Server::finish(Req *q, hipError_t e)
{
    q->done_ns = now_ns();
    if (e == hipSuccess) {
        memcpy(q->cap_out, pinned_ + (u64)q->slot * kCapSlotElems, 32ull << q->r.cap_height);
        q->state = Req::Done;
    } else {
        q->state = Req::Failed;
        if (q->error.empty()) q->error = std::string("commit request failed: ") + hipGetErrorString(e);
    }
    if (q->slot >= 0) free_slots_.push_back(q->slot);
    if (q->done) free_events_.push_back(q->done);
    if (q->lde_done) free_events_.push_back(q->lde_done);
    if (q->lane >= 0) lane_load_[q->lane]--;
    if (q->group >= 0 && --group_left_[q->group] == 0) free_groups_.push_back(q->group);
    q->slot = q->lane = q->group = -1;
    q->done = q->lde_done = nullptr;
    // a kept result finishing while the server shuts down is dropped like the
    // ones already done: its memory may be what a queued request waits for
    if (q->holds_memory && (!q->r.keep || q->state == Req::Failed || stop_)) {
        arena_.release(q->off, q->bytes);
        q->holds_memory = false;
    }
}

void Server::run()
{
    (void)hipSetDevice(c_.dev);
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        bool finished = false;
        for (size_t i = 0; i < running_.size();) {
            Req *q = running_[i];
            const hipError_t e = hipEventQuery(q->done);
            if (e == hipErrorNotReady) {
                i++;
                continue;
            }
            finish(q, e);
            running_[i] = running_.back();
            running_.pop_back();
            finished = true;
        }
        while (!pending_.empty() && running_.size() < max_inflight_) {
            // a group: the requests at the head of the queue that fit now
            std::vector<Req *> g;
            const u32 gmax = free_groups_.empty() ? 1 : group_max_;
            while (!pending_.empty() && g.size() < gmax && running_.size() + g.size() < max_inflight_) {
                Req *q = pending_.front();
                if (!arena_.alloc(q->bytes, q->off)) break;  // first in line waits for memory
                pending_.pop_front();
                q->holds_memory = true;
                g.push_back(q);
            }
            if (g.empty()) break;
            int gslot = -1;
            if (g.size() > 1) {
                gslot = free_groups_.back();
                free_groups_.pop_back();
                group_left_[gslot] = (u32)g.size();
                for (Req *q : g) q->group = gslot;
            }
            try {
                launch_group(g, gslot);
                for (Req *q : g) running_.push_back(q);
            } catch (const std::exception &ex) {
                // work already enqueued for the group must end before its
                // memory, slots and tables go to other requests
                for (auto st : lanes_) (void)hipStreamSynchronize(st);
                for (auto st : forest_) (void)hipStreamSynchronize(st);
                for (Req *q : g) {
                    q->error = ex.what();
                    finish(q, hipErrorLaunchFailure);
                }
                finished = true;
            }
        }
        if (finished) cv_done_.notify_all();
        if (running_.empty() && pending_.empty()) {
            if (stop_) break;
            cv_work_.wait(lk);
            continue;
        }
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
    }
}

void Server::wait(uint64_t ticket, ServeResult &res)
{
    std::unique_lock<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) throw Error((int)hipErrorInvalidValue, "unknown ticket");
    Req *q = it->second.get();
    cv_done_.wait(lk, [q] { return q->state == Req::Done || q->state == Req::Failed; });
    res.submit_ns = q->submit_ns;
    res.admit_ns = q->admit_ns;
    res.done_ns = q->done_ns;
    res.device_bytes = q->bytes;
    const bool kept = q->state == Req::Done && q->r.keep;
    res.lde = kept ? q->lde : nullptr;
    res.digests = kept ? q->digests : nullptr;
    if (q->state == Req::Failed) {
        const std::string msg = q->error;
        reqs_.erase(it);
        throw Error((int)hipErrorLaunchFailure, msg);
    }
    if (!kept) reqs_.erase(it);
}

void Server::release(uint64_t ticket)
{
    std::lock_guard<std::mutex> lk(mu_);
    auto it = reqs_.find(ticket);
    if (it == reqs_.end()) throw Error((int)hipErrorInvalidValue, "unknown ticket");
    Req *q = it->second.get();
    if (q->state != Req::Done) throw Error((int)hipErrorInvalidValue, "request not complete");
    if (q->holds_memory) arena_.release(q->off, q->bytes);
    reqs_.erase(it);
    cv_work_.notify_one();  // memory freed: the worker may admit more
}

Server *server_create(int dev, uint64_t budget, uint32_t max_inflight) { return new Server(dev, budget, max_inflight); }
uint64_t server_submit(Server *s, const ServeRequest &r, uint64_t *cap_out) { return s->submit(r, cap_out); }
void server_wait(Server *s, uint64_t ticket, ServeResult &res) { s->wait(ticket, res); }
void server_release(Server *s, uint64_t ticket) { s->release(ticket); }
void server_destroy(Server *s) { delete s; }

}  // namespace zk
