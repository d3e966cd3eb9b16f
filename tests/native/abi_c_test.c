/* C11 caller of libzeknox.so through include/lib.h, the way a cgo binding
 * sees it (okx/zeknox wrappers/go/lib/lib.go:7-12 includes the reference's
 * lib.h from C; bool is a one-byte type there).  It links -lzeknox unchanged
 * and checks:
 *   cpu  the argument-error conventions that need no device (code 1 with
 *        "invalid args" for a bad LDE, code 1 and no message for lg 0 NTT,
 *        RustError messages owned by the caller) and the host Merkle entry
 *        point against the Go cap known answers;
 *   gpu  the Go cap known answers through fill_digests_buf_linear_gpu_with_gpu_ptr
 *        and its multi-GPU name (wrappers/go/lib/lib_test.go:194-214), a
 *        coset LDE from host buffers against its closed form
 *        (p(x) = a + b x at g w^i) and compute_transpose_rev against its
 *        definition.
 * Usage: abi_c_test cpu|gpu; prints one line per failed check, exit code =
 * number of failures. */
#include <hip/hip_runtime_api.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lib.h"
#include "zeknox_ext.h"

static const uint64_t P = 0xFFFFFFFF00000001ull;
static int failures = 0;

#define CHECK(cond, ...)                                        \
    do {                                                        \
        if (!(cond)) {                                          \
            failures++;                                         \
            printf("FAIL %s:%d: ", __FILE__, __LINE__);         \
            printf(__VA_ARGS__);                                \
            printf("\n");                                       \
        }                                                       \
    } while (0)

static uint64_t mulmod(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a * b) % P); }
static uint64_t addmod(uint64_t a, uint64_t b) { return (uint64_t)(((unsigned __int128)a + b) % P); }
static uint64_t powmod(uint64_t a, uint64_t e)
{
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = mulmod(r, a);
        a = mulmod(a, a);
        e >>= 1;
    }
    return r;
}
static uint32_t bitrev(uint32_t x, uint32_t bits)
{
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) r |= ((x >> i) & 1u) << (bits - 1 - i);
    return r;
}

/* A RustError's message belongs to the caller (the Rust wrapper frees it,
 * wrappers/rust/src/error.rs:11-19). */
static int take(RustError e, const char **msg_out, char *buf, size_t n)
{
    buf[0] = 0;
    if (e.message) {
        strncpy(buf, e.message, n - 1);
        buf[n - 1] = 0;
        free(e.message);
    }
    if (msg_out) *msg_out = e.message ? buf : NULL;
    return e.code;
}

/* wrappers/go/lib/lib_test.go:194-214: a 7-element leaf repeated over every
 * leaf; every cap entry equals the expected hash. */
static const uint64_t kLeaf[7] = {8395359103262935841ull, 1377884553022145855ull, 2370707998790318766ull,
                                  3651132590097252162ull, 1141848076261006345ull, 12736915248278257710ull,
                                  9898074228282442027ull};
struct KatCase {
    uint64_t n_leaves, n_caps;
    uint64_t cap[4];
};
static const struct KatCase kKats[3] = {
    {8, 8, {7544909477878586743ull, 7431000548126831493ull, 17815668806142634286ull, 13168106265494210017ull}},
    {8, 4, {7320113784629306061ull, 6779895905614599055ull, 11363085033710007978ull, 10376477785629036773ull}},
    {16, 2, {15643713602922591410ull, 15764162833178839023ull, 7963799186968795376ull, 10933634831850922421ull}},
};

static uint32_t log2u(uint64_t x)
{
    uint32_t r = 0;
    while (x > 1) {
        x >>= 1;
        r++;
    }
    return r;
}

static void check_caps(const char *what, const struct KatCase *k, const uint64_t *caps)
{
    for (uint64_t j = 0; j < k->n_caps; j++)
        for (int i = 0; i < 4; i++)
            CHECK(caps[4 * j + i] == k->cap[i], "%s: %llu leaves, %llu caps: cap[%llu][%d] = %llu", what,
                  (unsigned long long)k->n_leaves, (unsigned long long)k->n_caps, (unsigned long long)j, i,
                  (unsigned long long)caps[4 * j + i]);
}

static void cpu_checks(void)
{
    char buf[256];
    const char *msg;
    NTT_Config cfg;
    memset(&cfg, 0, sizeof cfg);
    cfg.batches = 1;
    cfg.extension_rate_bits = 3;
    /* ntt.cu:478-481: a zero-size LDE is code 1 "invalid args" */
    int code = take(compute_batched_lde(0, NULL, NULL, 0, forward, cfg), &msg, buf, sizeof buf);
    CHECK(code == 1 && msg && strcmp(msg, "invalid args") == 0, "lde lg 0: code %d msg %s", code, msg ? msg : "-");
    /* ntt.cu:255-256: a zero-size NTT is code 1 without a message */
    code = take(compute_batched_ntt(0, NULL, 0, forward, cfg), &msg, buf, sizeof buf);
    CHECK(code == 1 && msg == NULL, "ntt lg 0: code %d", code);
    /* an in-place commit is refused before any device work */
    code = take(zk_commit(0, (const uint64_t *)(uintptr_t)(1u << 20), (uint64_t *)(uintptr_t)(1u << 20), NULL,
                          (uint64_t *)(uintptr_t)(3u << 20), (uint64_t *)(uintptr_t)(4u << 20), 4, 1, 2, 1, 7),
                &msg, buf, sizeof buf);
    CHECK(code == 1 && msg && strstr(msg, "overlapping"), "overlapping commit: code %d", code);
    size_t n = 99;
    code = take(get_number_of_gpus(&n), NULL, buf, sizeof buf);
    CHECK(code == 0 && n < 99, "get_number_of_gpus: code %d n %zu", code, n);
    CHECK(sizeof(NTT_Config) == 24 && sizeof(NTT_TransposeConfig) == 8 && sizeof(RustError) == 16,
          "struct sizes %zu %zu %zu", sizeof(NTT_Config), sizeof(NTT_TransposeConfig), sizeof(RustError));
    /* host Merkle: the Go known answers (lib_test.go:265-277 runs the same
     * cases through the CPU entry point) */
    for (int c = 0; c < 3; c++) {
        const struct KatCase *k = &kKats[c];
        uint64_t *leaves = malloc(sizeof(uint64_t) * 7 * k->n_leaves);
        uint64_t nd = 2 * (k->n_leaves - k->n_caps);
        uint64_t *dg = calloc(4 * (nd ? nd : 1), sizeof(uint64_t));
        uint64_t *cp = calloc(4 * k->n_caps, sizeof(uint64_t));
        for (uint64_t l = 0; l < k->n_leaves; l++) memcpy(leaves + 7 * l, kLeaf, sizeof kLeaf);
        fill_digests_buf_linear_cpu(dg, cp, leaves, nd, k->n_caps, k->n_leaves, 7, log2u(k->n_caps), HashPoseidon);
        check_caps("fill_digests_buf_linear_cpu", k, cp);
        free(leaves);
        free(dg);
        free(cp);
    }
}

static void *dmalloc(size_t bytes)
{
    void *p = NULL;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        printf("FAIL hipMalloc(%zu)\n", bytes);
        exit(100);
    }
    return p;
}

static void gpu_checks(void)
{
    char buf[256];
    const char *msg;
    CHECK(cuda_available(), "cuda_available() is false");
    /* Go known answers through the device Merkle entry points */
    for (int c = 0; c < 3; c++) {
        const struct KatCase *k = &kKats[c];
        uint64_t nd = 2 * (k->n_leaves - k->n_caps);
        uint64_t *leaves = malloc(sizeof(uint64_t) * 7 * k->n_leaves);
        uint64_t cp[4 * 16];
        for (uint64_t l = 0; l < k->n_leaves; l++) memcpy(leaves + 7 * l, kLeaf, sizeof kLeaf);
        void *d_leaves = dmalloc(sizeof(uint64_t) * 7 * k->n_leaves);
        void *d_dg = dmalloc(sizeof(uint64_t) * 4 * (nd ? nd : 1));
        void *d_cp = dmalloc(sizeof(uint64_t) * 4 * k->n_caps);
        hipMemcpy(d_leaves, leaves, sizeof(uint64_t) * 7 * k->n_leaves, hipMemcpyHostToDevice);
        for (int multi = 0; multi < 2; multi++) {
            hipMemset(d_cp, 0, sizeof(uint64_t) * 4 * k->n_caps);
            if (multi)
                fill_digests_buf_linear_multigpu_with_gpu_ptr(d_dg, d_cp, d_leaves, nd, k->n_caps, k->n_leaves, 7,
                                                              log2u(k->n_caps), HashPoseidon);
            else
                fill_digests_buf_linear_gpu_with_gpu_ptr(d_dg, d_cp, d_leaves, nd, k->n_caps, k->n_leaves, 7,
                                                         log2u(k->n_caps), HashPoseidon, 0);
            hipMemcpy(cp, d_cp, sizeof(uint64_t) * 4 * k->n_caps, hipMemcpyDeviceToHost);
            check_caps(multi ? "fill_digests_buf_linear_multigpu_with_gpu_ptr"
                             : "fill_digests_buf_linear_gpu_with_gpu_ptr", k, cp);
        }
        hipFree(d_leaves);
        hipFree(d_dg);
        hipFree(d_cp);
        free(leaves);
    }
    /* coset LDE from host buffers: column b holds p_b(x) = a_b + c_b x, so
     * out[b][i] = a_b + c_b g w^i with w = 7^((p-1)/N), g = 7 */
    {
        const uint32_t lg_n = 10, r = 3, C = 3;
        const uint32_t n = 1u << lg_n, N = 1u << (lg_n + r);
        uint64_t *in = calloc((size_t)C * n, sizeof(uint64_t));
        uint64_t *out = calloc((size_t)C * N, sizeof(uint64_t));
        for (uint32_t b = 0; b < C; b++) {
            in[(size_t)b * n] = 1000 + b;
            in[(size_t)b * n + 1] = P - 1 - b;
        }
        NTT_Config cfg;
        memset(&cfg, 0, sizeof cfg);
        cfg.batches = C;
        cfg.extension_rate_bits = r;
        cfg.with_coset = true;
        int code = take(init_coset(0, lg_n + r, 7), &msg, buf, sizeof buf);
        CHECK(code == 0, "init_coset: %d %s", code, buf);
        code = take(compute_batched_lde(0, out, in, lg_n, forward, cfg), &msg, buf, sizeof buf);
        CHECK(code == 0, "compute_batched_lde: %d %s", code, buf);
        const uint64_t w = powmod(7, (P - 1) >> (lg_n + r));
        uint64_t x = 7;  /* g w^i */
        int bad = 0;
        for (uint32_t i = 0; i < N; i++) {
            for (uint32_t b = 0; b < C; b++) {
                const uint64_t want = addmod(1000 + b, mulmod(P - 1 - b, x));
                if (out[(size_t)b * N + i] != want && bad++ < 3)
                    CHECK(0, "lde out[%u][%u] = %llu, want %llu", b, i, (unsigned long long)out[(size_t)b * N + i],
                          (unsigned long long)want);
            }
            x = mulmod(x, w);
        }
        CHECK(bad == 0, "%d LDE values differ from the closed form", bad);
        /* transpose_rev of the LDE, host buffers: leaves[rev(i)][b] = lde[b][i] */
        uint64_t *tr = calloc((size_t)C * N, sizeof(uint64_t));
        NTT_TransposeConfig tc;
        memset(&tc, 0, sizeof tc);
        tc.batches = C;
        code = take(compute_transpose_rev(0, tr, out, lg_n + r, tc), &msg, buf, sizeof buf);
        CHECK(code == 0, "compute_transpose_rev: %d %s", code, buf);
        bad = 0;
        for (uint32_t i = 0; i < N; i++)
            for (uint32_t b = 0; b < C; b++)
                if (tr[(size_t)bitrev(i, lg_n + r) * C + b] != out[(size_t)b * N + i]) bad++;
        CHECK(bad == 0, "%d transposed values misplaced", bad);
        free(in);
        free(out);
        free(tr);
    }
}

int main(int argc, char **argv)
{
    const int gpu = argc > 1 && strcmp(argv[1], "gpu") == 0;
    cpu_checks();
    if (gpu) gpu_checks();
    printf("%s checks: %d failed\n", gpu ? "cpu+gpu" : "cpu", failures);
    return failures;
}
