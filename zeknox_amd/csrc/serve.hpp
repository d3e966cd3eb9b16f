// Commit server (serve.hip): admits commit requests by a device-memory budget
// as they arrive; the requests admitted together run as one group (their
// LDEs on several streams, their Merkle forests as fused launches), and
// groups admitted while others are on the GPU overlap with them (continuous
// batching).
#pragma once
#include <stdint.h>

#include "runtime.hpp"

namespace zk {

// Same layouts as zk_serve_request / zk_serve_result of include/zeknox_ext.h.
struct ServeRequest {
    uint32_t lg_n, rate_bits, ncols, cap_height;
    const uint64_t *coeffs;  // device, ncols x 2^lg_n column-major; nullptr: synthetic from seed
    uint64_t seed;
    uint32_t keep;  // keep the LDE and digests on the device until release
    uint32_t reserved;
};

struct ServeResult {
    uint64_t submit_ns, admit_ns, done_ns;  // host steady clock
    uint64_t *lde, *digests;                // device, when kept (else nullptr)
    uint64_t device_bytes;                  // arena bytes the request held
};

class Server;

Server *server_create(int dev, uint64_t budget_bytes, uint32_t max_inflight);
// Queues a request; cap_out (host, 4 << cap_height u64) is written before
// wait() returns for its ticket.
uint64_t server_submit(Server *s, const ServeRequest &r, uint64_t *cap_out);
// Blocks until the request is complete; throws its error if it failed.
void server_wait(Server *s, uint64_t ticket, ServeResult &res);
void server_release(Server *s, uint64_t ticket);
void server_destroy(Server *s);

}  // namespace zk
