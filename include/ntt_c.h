/*
 * ntt_c.h — C ABI of libntt_mi355x.so, the MI355X (gfx950) NTT library.
 *
 * Plain C: pointers, sizes and status codes; no C++ or torch types.  Every
 * entry point returns an ntt_status (0 = success) and never exits the
 * process.  Device pointers are HIP device memory; `stream` is a hipStream_t
 * passed as void* (NULL = the default stream).  All transforms are
 * asynchronous on that stream and allocate no device memory: tables are made
 * by ntt_c_init / ntt_c_batch_prepare / ntt_c_context_prepare (or a context's
 * first use of a (q, n, kind)), and the four-step workspace of a stream is
 * reserved once at set-up (ntt_c_init: the NULL stream; ntt_c_config_reserve,
 * ntt_c_context_reserve: any stream).  A transform on a stream without enough
 * workspace returns NTT_ERR_OOM; it never allocates, synchronises or frees.
 *
 * Each entry point names the reference interface it replaces
 * (MMRROOO/ntt-tensorcore, file:line).  include/ntt.cuh and
 * include/ntt_arbitrary.cuh wrap these in the reference's C++ names.
 */
#ifndef NTT_C_H_
#define NTT_C_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    NTT_OK = 0,
    NTT_ERR_INVALID_ARG = 1,  /* null pointer, n not a power of two, bad variant/flags */
    NTT_ERR_NO_ROOT = 2,      /* q prime-test failed or (2)n does not divide q-1 */
    NTT_ERR_OOM = 3,          /* device or host allocation failed / over budget */
    NTT_ERR_DEVICE = 4,       /* a HIP call failed */
    NTT_ERR_UNSUPPORTED = 5,  /* size or modulus outside what this build handles */
    NTT_ERR_BUSY = 6,         /* serving queue full / request not finished */
} ntt_status;

/* ModularParams, reference include/modular_arith.cuh:11-19 (same fields, same order). */
typedef struct {
    uint64_t q;
    uint64_t q_twice;
    uint64_t barrett_k;
    uint64_t barrett_m;
    uint64_t omega;      /* primitive n-th root g^((q-1)/n), g by the reference's generator rule */
    uint64_t omega_inv;
    uint64_t n_inv;
} ntt_c_modular_params;

/* NTTConfig, reference include/ntt.cuh:11-24.  The leading fields are the
 * reference's; `impl` holds this library's plans and workspace. */
typedef struct {
    uint64_t n;
    uint64_t log_n;
    uint64_t q;
    uint64_t mu;                 /* floor(2^64 / q) */
    uint64_t* d_twiddles;        /* device: omega^k, k < n */
    uint64_t* d_twiddles_inv;    /* device: omega^-k, k < n */
    ntt_c_modular_params params;
    void* impl;
} ntt_c_config;

/* NTTVariant, reference include/ntt.cuh:27-32. */
enum { NTT_VARIANT_BASIC = 0, NTT_VARIANT_FOUR_STEP = 1, NTT_VARIANT_TLMOP = 2, NTT_VARIANT_FULL_OPTIMIZED = 3 };

/* Transform flags for the native descriptor API. */
enum {
    NTT_FLAG_INVERSE = 1u,     /* inverse transform (includes n^-1) */
    NTT_FLAG_NEGACYCLIC = 2u,  /* Z_q[x]/(x^n+1) with psi = g^((q-1)/2n) */
    NTT_FLAG_U64 = 4u,         /* coefficients stored as uint64 (else uint32; q >= 2^31 requires U64) */
};

/* ---------------- 31-bit path: include/ntt.cuh ---------------- */

/* ntt_init, reference src/ntt_basic.cu:217-249 (include/ntt.cuh:35).  Builds
 * params by the reference's generator rule, the device twiddle tables, the
 * transform plans and the NULL stream's workspace (one call of any kind; a
 * batch runs in chunks that fit it).  q prime with n | q-1, either q < 2^62
 * or the Goldilocks prime 2^64 - 2^32 + 1 (its own reduction; the reference's
 * uint64 sums overflow there).  On failure nothing stays allocated. */
ntt_status ntt_c_init(uint64_t n, uint64_t q, ntt_c_config** out);
/* ntt_cleanup, reference src/ntt_basic.cu:251-257 (include/ntt.cuh:38). */
ntt_status ntt_c_cleanup(ntt_c_config* cfg);
/* Set-up call with no reference counterpart (the reference's optimized path
 * needs no workspace; its four-step allocates per call,
 * src/ntt_four_step.cu:212-228): reserves the workspace of `stream` for calls
 * of up to max_batch polynomials at once (larger batches run in chunks).
 * Needed once per stream other than NULL before transforms on it; it may
 * allocate and synchronise the stream, so it must not run while another
 * thread issues transforms on that stream. */
ntt_status ntt_c_config_reserve(const ntt_c_config* cfg, void* stream, uint64_t max_batch);

/* No reference counterpart; for callers that cannot allocate device memory
 * at all (code under hipGraph capture, tight serving loops, hosts that forbid
 * allocation).  The transform of ntt_c_forward / ntt_c_inverse /
 * ntt_c_forward_negacyclic / ntt_c_inverse_negacyclic (flags NTT_FLAG_INVERSE
 * | NTT_FLAG_NEGACYCLIC) on uint64 data in place, with its four-step
 * workspace in a static 32 MiB device array of the library instead of a
 * reserved buffer: no hipMalloc or hipFree, on any stream.  A transform needs
 * at most 8 bytes of workspace per coefficient, so every n <= 2^22 fits;
 * transforms whose workspace exceeds the 32 MiB (some plans at 2^23, all at
 * 2^24 and above) return NTT_ERR_UNSUPPORTED and leave the data alone.  All calls on one GPU share
 * that GPU's array: calls that overlap on one GPU (concurrent streams or host
 * threads) race; calls ordered on one stream do not.  Nothing in the array
 * needs to start at zero (every word is written before it is read). */
ntt_status ntt_c_config_transform_noalloc(uint64_t* d_data, const ntt_c_config* cfg, uint32_t flags, void* stream);

/* ntt_forward_{basic,four_step,tlmop,optimized} / ntt_inverse_*, reference
 * include/ntt.cuh:42-57 (src/ntt_basic.cu:268-334, src/ntt_four_step.cu:244-392,
 * src/ntt_optimized.cu:515-613).  In place, natural order in and out,
 * canonical values; every variant computes ntt_host's transform. */
ntt_status ntt_c_forward(uint64_t* d_data, const ntt_c_config* cfg, int variant, void* stream);
ntt_status ntt_c_inverse(uint64_t* d_data, const ntt_c_config* cfg, int variant, void* stream);

/* Negacyclic transform over Z_q[x]/(x^n+1): A_k = sum_j a_j psi^(j(2k+1)).
 * The reference names it (README.md:30) but has no code
 * (tests/test_ntt.cu:304-307); needs 2n | q-1. */
ntt_status ntt_c_forward_negacyclic(uint64_t* d_data, const ntt_c_config* cfg, void* stream);
ntt_status ntt_c_inverse_negacyclic(uint64_t* d_data, const ntt_c_config* cfg, void* stream);

/* ntt_forward_batch / ntt_inverse_batch, reference src/ntt_basic.cu:337-389
 * (include/ntt.cuh:70-75): batch_size polynomials at d_data + i*n, all in one
 * launch sequence instead of the reference's per-polynomial loop. */
ntt_status ntt_c_forward_batch(uint64_t* d_data, uint64_t batch_size, const ntt_c_config* cfg, int variant,
                               void* stream);
ntt_status ntt_c_inverse_batch(uint64_t* d_data, uint64_t batch_size, const ntt_c_config* cfg, int variant,
                               void* stream);

/* poly_multiply_ntt, reference src/ntt_basic.cu:392-409 (include/ntt.cuh:78-79). */
ntt_status ntt_c_poly_multiply(uint64_t* d_result, const uint64_t* d_a, const uint64_t* d_b,
                               const ntt_c_config* cfg, void* stream);

/* Negacyclic product a*b mod (x^n+1, q) into d_result; d_a, d_b are
 * overwritten with their forward transforms (the pipeline of
 * tests/test_ntt.cu:177-240 on the negacyclic transform).  The two forward
 * transforms run as one batch of two polynomials, and the pointwise product
 * is computed by the inverse transform's first pass as it loads its tiles.
 * d_result may alias d_a or d_b; d_a == d_b squares (one forward transform);
 * d_a and d_b otherwise must not overlap (NTT_ERR_INVALID_ARG). */
ntt_status ntt_c_negacyclic_multiply(uint64_t* d_result, uint64_t* d_a, uint64_t* d_b, const ntt_c_config* cfg,
                                     void* stream);

/* bit_reverse_permute, reference src/ntt_basic.cu:259-265 (include/ntt.cuh:82). */
ntt_status ntt_c_bit_reverse(uint64_t* d_data, uint64_t n, void* stream);

/* Coefficient width conversion between the reference ABI's uint64_t arrays
 * and this library's 32-bit storage (NTT_FLAG_U64 off), device to device on
 * the stream.  No reference counterpart: the reference stores only uint64_t
 * (include/ntt.cuh:35-85).  u64_to_u32 writes src[i] mod q (q < 2^32), so any
 * 64-bit input becomes a canonical residue; u32_to_u64 zero-extends.  Both
 * need 16-byte-aligned, non-overlapping buffers (NTT_ERR_INVALID_ARG). */
ntt_status ntt_c_u64_to_u32(uint32_t* d_dst, const uint64_t* d_src, uint64_t n, uint64_t q, void* stream);
ntt_status ntt_c_u32_to_u64(uint64_t* d_dst, const uint32_t* d_src, uint64_t n, void* stream);

/* ntt_forward_rowmaj, reference src/ntt_optimized.cu:940-965 (include/ntt.cuh:63):
 * input and output in the pre-transposed layout data[b*N1 + a] = canon[a*N2 + b]
 * with N1 = 2^ceil(log n / 2), N2 = n / N1 (N1 = N2 = 64 at n = 4096, as in the
 * reference).  Any power-of-two n. */
ntt_status ntt_c_forward_rowmaj(uint64_t* d_data, const ntt_c_config* cfg, void* stream);
/* rowmaj_to_pretransposed / rowmaj_from_pretransposed, reference
 * src/ntt_optimized.cu:922-934 (include/ntt.cuh:64-67). */
ntt_status ntt_c_rowmaj_to_pretransposed(uint64_t* d_dst, const uint64_t* d_src, int n1, int n2, void* stream);
ntt_status ntt_c_rowmaj_from_pretransposed(uint64_t* d_dst, const uint64_t* d_src, int n1, int n2, void* stream);

/* ntt_host, reference src/ntt_basic.cu:412-463 (include/ntt.cuh:85): the host
 * transform of the public API (radix-2 DIT on the CPU, natural order). */
ntt_status ntt_c_host(uint64_t* data, uint64_t n, const ntt_c_modular_params* params, int inverse);

/* create_params / find_primitive_root, reference include/modular_arith.cuh:130-170. */
ntt_status ntt_c_create_params(uint64_t q, uint64_t n, ntt_c_modular_params* out);
uint64_t ntt_c_find_primitive_root(uint64_t q, uint64_t n);

/* ---------------- native descriptor API (no reference counterpart: the
 * reference's only batching is the per-polynomial loop above) ---------------- */

typedef struct {
    void* data;       /* device pointer, n coefficients (uint32 or uint64 per NTT_FLAG_U64), in place */
    uint64_t q;       /* prime, q < 2^31 with uint32 storage; q < 2^62 or 2^64 - 2^32 + 1 with NTT_FLAG_U64 */
    uint32_t log_n;   /* 1 .. 26 */
    uint32_t flags;   /* NTT_FLAG_* */
} ntt_c_desc;

typedef struct ntt_c_context ntt_c_context;  /* device-resident plan/twiddle cache */
typedef struct ntt_c_batch ntt_c_batch;      /* a prepared batch (job tables, workspace, graph) */

ntt_status ntt_c_context_create(ntt_c_context** out);
ntt_status ntt_c_context_destroy(ntt_c_context* ctx);
/* Bytes of device memory the context's tables use. */
uint64_t ntt_c_context_table_bytes(const ntt_c_context* ctx);
/* Set-up calls of the descriptor API (no reference counterpart):
 * context_reserve makes the workspace of `stream` at least `bytes` (it may
 * allocate and synchronise the stream; not while other threads issue work on
 * it); context_prepare builds the plans of the descriptors' (q, n, kind) and
 * reserves the workspace one of them needs on `stream`, so that later
 * ntt_c_transform calls on them allocate nothing (graph capture, serving
 * loops).  workspace_bytes: what one ntt_c_transform of *desc needs (at most
 * 8 n bytes; 0 for one-pass sizes). */
ntt_status ntt_c_context_reserve(ntt_c_context* ctx, void* stream, uint64_t bytes);
ntt_status ntt_c_context_prepare(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count, void* stream);
ntt_status ntt_c_workspace_bytes(const ntt_c_desc* desc, uint64_t* bytes);

/* One transform, one polynomial, launched on its own (the per-polynomial
 * path); its workspace is the stream's reservation (NTT_ERR_OOM if short). */
ntt_status ntt_c_transform(ntt_c_context* ctx, const ntt_c_desc* desc, void* stream);
/* ntt_c_transform on each descriptor in turn, on one stream: the reference's
 * one-polynomial-per-launch batch loop (src/ntt_basic.cu:337-362) issued from
 * native code, the comparison point of the descriptor batch. */
ntt_status ntt_c_transform_each(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count, void* stream);

/* One transform with an explicit radix split n = 2^log_r[0] * ... * 2^log_r[npass-1]
 * (npass <= 4), a tile of 2^log_e elements per workgroup, and bit i of
 * mfma_mask asking pass i to run its radix-16 rounds on the INT8 matrix cores
 * (honoured for 2^log_r[i] >= 16, q < 2^31 and tiles of >= 2^8 elements;
 * log_r[i] mod 4 levels stay on the vector ALUs).  No reference counterpart (the
 * reference fixes its split at radix-64 rounds, src/ntt_optimized.cu:529-554);
 * used by the autotuner and by tests that cover every plan shape. */
ntt_status ntt_c_transform_split(ntt_c_context* ctx, const ntt_c_desc* desc, uint32_t npass, const uint32_t* log_r,
                                 uint32_t log_e, uint32_t mfma_mask, void* stream);

/* Prepare a batch of descriptors (mixed n, q, direction, transform type):
 * builds tables for every (q, n) in the batch, the device job lists and the
 * workspace.  The descriptors' data pointers are bound at prepare time. */
ntt_status ntt_c_batch_prepare(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count, ntt_c_batch** out);
/* The key-switching inner product fused into inverse transforms: the
 * co-optimization of IP kernels with the NTT that the reference names as
 * absent (COOP, src/ntt_optimized.cu:510-511).  A batch as above, except that
 * each descriptor i with ips[i].dnum > 0 (NTT_FLAG_INVERSE required) reads,
 * in place of its data, the inner product
 *     x[c] = sum_{k < dnum} a[c + k*kstride] * b[c + k*kstride] mod q
 * of canonical residues in the descriptor's storage, and its inverse
 * transform writes INTT(x) to descs[i].data: x is formed in the first pass's
 * load and never stored.  dnum = 1 is a pointwise product (the second half of
 * a negacyclic multiplication).  Needs dnum <= 64, n <= kstride <= 2^40
 * when dnum > 1, reserved = 0, and data not overlapping a or b; ips[i].dnum = 0
 * transforms descs[i].data as ntt_c_batch_prepare does.  The inputs are bound
 * at prepare time like the data. */
typedef struct ntt_c_ip {
    const void* a;
    const void* b;
    uint32_t dnum;     /* pairs in the inner product; 0: none */
    uint32_t reserved; /* 0 */
    int64_t kstride;   /* elements from pair k to pair k + 1 */
} ntt_c_ip;
ntt_status ntt_c_batch_prepare_ip(ntt_c_context* ctx, const ntt_c_desc* descs, const ntt_c_ip* ips, size_t count,
                                  ntt_c_batch** out);
/* Launch a prepared batch: one kernel launch per (pass, arithmetic, storage
 * pair, round kind) group of its jobs (two launches for a 32-bit mix of one-
 * and two-pass plans), or, after ntt_c_batch_set_queue(batch, 1, 0), the
 * whole batch as ONE persistent launch.  Either way the launch count depends
 * on the kinds in the mix, not on the number of polynomials
 * (ntt_c_batch_info reports it).  Replaces the reference's per-polynomial
 * loop (src/ntt_basic.cu:337-389). */
ntt_status ntt_c_batch_launch(ntt_c_batch* batch, void* stream);
/* Capture the batch's launches into a hipGraph once; later calls replay it
 * (a new capture after ntt_c_batch_set_queue changes the launch form). */
ntt_status ntt_c_batch_launch_graph(ntt_c_batch* batch, void* stream);
/* Choose the launch form of a prepared batch (no reference counterpart):
 * enable = 1, the single persistent work-queue launch: workgroups take tiles
 * from per-XCD queues by atomic tickets and a pass of a polynomial starts as
 * soon as that polynomial's previous pass is complete (device counters), each
 * polynomial kept on one XCD where a probe launch saw all XCDs (else
 * agent-scope fences hand the data over); 2, the same launch with agent-scope
 * fences in every case; 0, one launch per phase group (the default: it
 * measured faster on every batch workload).  The work queue needs every job
 * on one kernel instantiation (32-bit coefficients and arithmetic, one round
 * kind): NTT_ERR_UNSUPPORTED otherwise.
 * spin_limit bounds each wait of a dependent tile (0: the default, about a
 * second); a wait that gives up is reported by ntt_c_batch_check.  Calls that
 * overlap on one batch race on its device counters. */
ntt_status ntt_c_batch_set_queue(ntt_c_batch* batch, int enable, uint32_t spin_limit);
/* Synchronises `stream`, then returns NTT_ERR_DEVICE (and clears the flag) if a
 * work-queue wait of an earlier launch of this batch gave up before its inputs
 * were complete, else NTT_OK. */
ntt_status ntt_c_batch_check(ntt_c_batch* batch, void* stream);
ntt_status ntt_c_batch_destroy(ntt_c_batch* batch);
/* Launch count, workspace bytes and launched workgroups of a prepared batch
 * in its current launch form. */
ntt_status ntt_c_batch_info(const ntt_c_batch* batch, uint32_t* launches, uint64_t* workspace_bytes,
                            uint64_t* workgroups);

/* ---------------- autotuner (no reference counterpart; the reference's split
 * is fixed at radix-64 rounds, src/ntt_optimized.cu:529-554) ----------------
 * For each log_n, times candidate radix splits / tile sizes / VALU-or-MFMA
 * passes on a uniform batch of elems_in_flight elements (modulus q) and
 * installs the fastest for that (size, launch-size class); later plans use it. */
typedef struct {
    uint32_t log_n, npass, log_r[4], log_e, mfma_mask, candidates;
    double us_per_transform;  /* tuned choice */
    double heuristic_us;      /* the built-in heuristic on the same batch */
} ntt_c_tune_result;
ntt_status ntt_c_autotune(ntt_c_context* ctx, uint64_t q, const uint32_t* log_ns, size_t count,
                          uint64_t elems_in_flight, uint32_t reps, ntt_c_tune_result* out);
ntt_status ntt_c_autotune_save(const char* path);
ntt_status ntt_c_autotune_load(const char* path);
ntt_status ntt_c_autotune_clear(void);
/* The split now chosen for 2^log_n with elems_in_flight elements per launch:
 * out = {npass, log_r[0..3], log_e, mfma_mask}. */
ntt_status ntt_c_split_for(uint64_t q, uint32_t log_n, uint64_t elems_in_flight, uint32_t out[7]);
/* Process-wide kernel choice (no reference counterpart).  NTT_OPT_PIPELINE:
 * value 1 (the default) runs a pass (32-bit storage, matrix-core rounds, a
 * pipelined tile shape; a single transform, or the polynomials of a batch
 * group that share one plan and sit at equal spacing) as workgroups that walk
 * several tiles and load the next tile while transforming the current one,
 * where the pass has at least two tiles per workgroup slot of the device;
 * 2 does so for every such pass with at least 16 tiles (testing); 0 runs one
 * tile per workgroup.  Other values are refused.  Returns the previous value
 * in *old (may be NULL).  Results are identical either way. */
#define NTT_OPT_PIPELINE 1
ntt_status ntt_c_set_option(uint32_t option, int64_t value, int64_t* old);

/* ---------------- continuous-batching server (no reference counterpart) ----------------
 * Requests are lists of device-resident polynomials (ntt_c_desc, transformed in
 * place).  A scheduler thread admits queued requests in arrival order while
 * their four-step workspace fits the HBM budget, packs every admitted request
 * into one batch launch per free slot (stream), and overlaps up to
 * max_inflight batches on the device.  Twiddles come from the context's
 * device-resident plan cache.  Latency = submit -> completion seen by the host. */
typedef struct ntt_c_server ntt_c_server;
typedef struct {
    uint64_t hbm_budget_bytes;  /* arena for workspace (default 1 GiB) */
    uint32_t max_inflight;      /* concurrent batches / streams (1..16, default 4) */
    uint32_t max_batch_polys;   /* polynomials per batch (default 1024) */
    int device;                 /* set by create: the current HIP device */
} ntt_c_server_opts;
typedef struct {
    uint64_t completed, failed, batches, butterflies;
    double p50_us, p99_us, mean_us;
    double wall_s;                  /* first submit -> last completion */
    uint32_t max_concurrent_batches;
    uint64_t arena_free_bytes;
} ntt_c_server_stats;

ntt_status ntt_c_server_create(ntt_c_context* ctx, const ntt_c_server_opts* opts, ntt_c_server** out);
/* Queues one request; NTT_ERR_OOM if it can never fit the budget. */
ntt_status ntt_c_server_submit(ntt_c_server* srv, const ntt_c_desc* polys, size_t count, uint64_t* request_id);
/* Blocks until the request finished (timeout_s < 0: no timeout; NTT_ERR_BUSY on timeout). */
ntt_status ntt_c_server_wait(ntt_c_server* srv, uint64_t request_id, double timeout_s);
ntt_status ntt_c_server_poll(ntt_c_server* srv, uint64_t request_id, int* done, double* latency_us);
ntt_status ntt_c_server_get_stats(ntt_c_server* srv, ntt_c_server_stats* out);
ntt_status ntt_c_server_reset_stats(ntt_c_server* srv);
/* Submits request i at arrival_s[i] seconds after the call (polys packed
 * back to back, counts[i] per request), waits for all, writes latencies. */
ntt_status ntt_c_server_replay(ntt_c_server* srv, const ntt_c_desc* polys, const uint32_t* counts,
                               const double* arrival_s, size_t nreq, double* latency_us);
ntt_status ntt_c_server_destroy(ntt_c_server* srv);

/* ---------------- K-limb big-prime transforms (include/ntt_arbitrary.cuh)
 * A value is K uint32 limbs of 24 bits, little-endian (DScalar<K>, reference
 * include/ntt_arbitrary.cuh:51,573-576), 1 <= K <= 12, in Montgomery form
 * x*R mod p with R = 2^(24K).  `prime_limbs` (K limbs, host memory) is any
 * odd modulus; `np` = -p^-1 mod 2^24 (mont_np, :456-462).  Arrays are n = 2^log_n
 * values in natural order, transformed in place; d_tw[k] = w^k * R mod p for a
 * primitive n-th root w (the table setup_ntt builds, :2698-2793).  Products are
 * CIOS Montgomery (mont_mul, :1218-1298) with canonical results, so outputs
 * equal host_ntt_naive (:2651-2676) on the same values, limb for limb. */

/* Forward DIT NTT: run_forward_ct / run_forward_ct_tcu, reference :2547-2646. */
ntt_status ntt_c_arb_forward(uint32_t* d_data, const uint32_t* d_tw, uint32_t n, uint32_t log_n,
                             const uint32_t* prime_limbs, uint32_t np, uint32_t K, void* stream);
/* Inverse NTT (the reference has none): twiddles w^-k = d_tw[(n-k) % n], then a
 * product by n_inv_mont = n^-1 * R mod p (K limbs, host memory). */
ntt_status ntt_c_arb_inverse(uint32_t* d_data, const uint32_t* d_tw, uint32_t n, uint32_t log_n,
                             const uint32_t* prime_limbs, uint32_t np, const uint32_t* n_inv_mont, uint32_t K,
                             void* stream);
/* d_a[i] = d_a[i] * d_b[i] * R^-1 mod p, i < n: the pointwise step of a
 * polynomial product and the psi-twists of the negacyclic transform (no
 * reference counterpart for K limbs; poly_multiply_ntt, src/ntt_basic.cu:392-409,
 * is the 31-bit one). */
ntt_status ntt_c_arb_pointwise_mul(uint32_t* d_a, const uint32_t* d_b, uint32_t n, const uint32_t* prime_limbs,
                                   uint32_t np, uint32_t K, void* stream);
/* d_a[i] = d_a[i] * s * R^-1 mod p with s in host memory: into Montgomery form
 * with s = R^2 mod p, out of it with s = 1 (the conversions of setup_ntt and
 * run_for_K, reference :2731-2746, 2884-2888). */
ntt_status ntt_c_arb_scale(uint32_t* d_a, uint32_t n, const uint32_t* scalar_limbs, const uint32_t* prime_limbs,
                           uint32_t np, uint32_t K, void* stream);

/* K-limb transforms on the four-step pass kernels, for moduli that fit 64-bit
 * arithmetic: K = 1 or 2, or K = 3 with p < 2^62 or p = 2^64 - 2^32 + 1
 * (Goldilocks, FieldConfig<3>).  Same data and the same results, limb for limb,
 * as ntt_c_arb_forward / ntt_c_arb_inverse with d_tw[k] = w^k R mod p, where
 * root_mont = w R mod p (d_tw[1]; K limbs, host memory): the transform is
 * linear, so it runs on the Montgomery-form values as plain residues.  Other
 * moduli return NTT_ERR_UNSUPPORTED (use ntt_c_arb_forward).  The plan owns
 * its tables and workspace (allocated here, at set-up); calls with one plan
 * must not overlap (one stream, or synchronised).  Replaces the reference's
 * run_forward_ct / run_forward_ct_tcu (include/ntt_arbitrary.cuh:2547-2646)
 * for these fields; the inverse has no reference counterpart. */
typedef struct ntt_c_arb_plan ntt_c_arb_plan;
ntt_status ntt_c_arb_plan_create(ntt_c_arb_plan** plan, const uint32_t* prime_limbs, uint32_t np, uint32_t K,
                                 uint32_t log_n, const uint32_t* root_mont);
ntt_status ntt_c_arb_plan_forward(const ntt_c_arb_plan* plan, uint32_t* d_data, void* stream);
ntt_status ntt_c_arb_plan_inverse(const ntt_c_arb_plan* plan, uint32_t* d_data, void* stream);
ntt_status ntt_c_arb_plan_destroy(ntt_c_arb_plan* plan);

/* ---------------- RNS kernels around the transform.  No reference
 * counterpart: the reference names them as absent (src/ntt_optimized.cu:510-513,
 * "BConv/IP kernels (none exist)", "no PMAC in this codebase").  Residues are
 * 32-bit, limb-major (limb i of n coefficients at d[i*n .. i*n + n)); moduli
 * are odd primes below 2^31. ---------------- */
typedef struct ntt_c_bconv_plan ntt_c_bconv_plan;
typedef struct ntt_c_pmac_plan ntt_c_pmac_plan;

/* Fast base conversion from the L moduli q_from (distinct, L <= 32) to the K
 * moduli p_to (K <= 64): y_j = sum_i [x_i (Q/q_i)^-1]_{q_i} (Q/q_i) mod p_j,
 * i.e. (X + u Q) mod p_j for the CRT value X of x and some 0 <= u < L.  The
 * plan holds its constant tables on the device. */
ntt_status ntt_c_bconv_create(const uint64_t* q_from, uint32_t L, const uint64_t* p_to, uint32_t K,
                              ntt_c_bconv_plan** out);
/* d_out (K x n) from d_in (L x n, canonical residues); no overlap. */
ntt_status ntt_c_bconv(const ntt_c_bconv_plan* plan, uint32_t* d_out, const uint32_t* d_in, uint64_t n, void* stream);
ntt_status ntt_c_bconv_destroy(ntt_c_bconv_plan* plan);

/* Per-limb multiply-accumulate of an inner product over dnum <= 32 pairs:
 * out[l][c] = (accumulate ? out[l][c] : 0) + sum_k a[k][l][c] * b[k][l][c] mod q_l,
 * a and b laid out [k][l][c] (dnum x L x n), inputs canonical; out may not
 * overlap a or b. */
ntt_status ntt_c_pmac_create(const uint64_t* q, uint32_t L, ntt_c_pmac_plan** out);
ntt_status ntt_c_pmac(const ntt_c_pmac_plan* plan, uint32_t* d_out, const uint32_t* d_a, const uint32_t* d_b,
                      uint32_t dnum, uint64_t n, int accumulate, void* stream);
ntt_status ntt_c_pmac_destroy(ntt_c_pmac_plan* plan);

/* ---------------- device-memory and timing helpers for hosts without a HIP
 * binding (tests, benchmarks, FFI callers) ---------------- */
ntt_status ntt_c_device_count(int* count);
ntt_status ntt_c_malloc(void** ptr, uint64_t bytes);
ntt_status ntt_c_free(void* ptr);
ntt_status ntt_c_memcpy_h2d(void* dst, const void* src, uint64_t bytes);
ntt_status ntt_c_memcpy_d2h(void* dst, const void* src, uint64_t bytes);
ntt_status ntt_c_memset(void* dst, int value, uint64_t bytes);
ntt_status ntt_c_stream_create(void** stream);
ntt_status ntt_c_stream_destroy(void* stream);
ntt_status ntt_c_stream_sync(void* stream);
ntt_status ntt_c_device_sync(void);
ntt_status ntt_c_event_create(void** ev);
ntt_status ntt_c_event_destroy(void* ev);
ntt_status ntt_c_event_record(void* ev, void* stream);
ntt_status ntt_c_event_elapsed_ms(float* ms, void* start, void* end);
/* Stream capture into a hipGraph (global capture mode, so an allocation or
 * synchronisation inside the captured calls fails the capture), instantiate,
 * replay, destroy. */
ntt_status ntt_c_graph_begin(void* stream);
ntt_status ntt_c_graph_end(void* stream, void** graph_exec);
ntt_status ntt_c_graph_launch(void* graph_exec, void* stream);
ntt_status ntt_c_graph_destroy(void* graph_exec);
/* Human-readable text of the last error on this thread. */
const char* ntt_c_last_error(void);

#ifdef __cplusplus
}
#endif
#endif /* NTT_C_H_ */
