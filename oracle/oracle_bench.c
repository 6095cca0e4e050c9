/*
 * oracle_bench.c — times the oracle's ntt_host restatement (the CPU baseline
 * that bench.py reports beside the GPU numbers).  Test infrastructure only.
 */
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct {
    uint64_t q, q_twice, barrett_k, barrett_m, omega, omega_inv, n_inv;
} orc_params;
void orc_create_params(uint64_t q, uint64_t n, orc_params* p);
void orc_ntt_host(uint64_t* data, uint64_t n, const orc_params* prm, int inverse);
void orc_ntt_negacyclic(uint64_t* data, uint64_t n, uint64_t q, int inverse);
void orc_random_poly(uint64_t* out, uint64_t n, uint64_t q, uint64_t seed);

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + ts.tv_nsec * 1e-9;
}

/* One negacyclic forward+inverse round trip of size n per repetition, on one
 * thread.  Parameter setup is inside the timed call, as in ntt_host's own
 * per-stage mod_pow.  Returns seconds per round trip (median over reps). */
double orc_time_negacyclic_roundtrip(uint64_t n, uint64_t q, uint64_t seed, int reps) {
    uint64_t* x = (uint64_t*)malloc(n * sizeof(uint64_t));
    double best[64];
    if (reps > 64) reps = 64;
    for (int r = 0; r < reps; r++) {
        orc_random_poly(x, n, q, seed + r);
        double t0 = now_s();
        orc_ntt_negacyclic(x, n, q, 0);
        orc_ntt_negacyclic(x, n, q, 1);
        best[r] = now_s() - t0;
    }
    free(x);
    /* median */
    for (int i = 1; i < reps; i++) { double v = best[i]; int j = i - 1; while (j >= 0 && best[j] > v) { best[j + 1] = best[j]; j--; } best[j + 1] = v; }
    return best[reps / 2];
}

typedef struct { const uint64_t* ns; const uint64_t* qs; const int* dirs; int count; int next; pthread_mutex_t mu; } job_t;

static void* worker(void* arg) {
    job_t* J = (job_t*)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int i = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (i >= J->count) break;
        uint64_t n = J->ns[i];
        uint64_t* x = (uint64_t*)malloc(n * sizeof(uint64_t));
        orc_random_poly(x, n, J->qs[i], 1000 + i);
        orc_ntt_negacyclic(x, n, J->qs[i], J->dirs[i]);
        free(x);
    }
    return NULL;
}

/* A batch of negacyclic transforms (sizes ns, primes qs, directions dirs), one
 * polynomial per worker on `threads` pthreads.  Returns wall seconds. */
double orc_time_batch(const uint64_t* ns, const uint64_t* qs, const int* dirs, int count, int threads) {
    job_t J = {ns, qs, dirs, count, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    if (threads > 256) threads = 256;
    double t0 = now_s();
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, worker, &J);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    return now_s() - t0;
}

/* In-place transforms of `count` polynomials (data[i]: n[i] canonical
 * residues mod q[i]; flags[i] bit 0 = inverse, bit 1 = negacyclic), one
 * polynomial per worker on `threads` pthreads.  Each polynomial goes through
 * the single-thread oracle (orc_ntt / orc_ntt_negacyclic), so the result is
 * the oracle's whatever the thread count.  Returns wall seconds. */
typedef struct {
    uint64_t* const* data; const uint64_t* ns; const uint64_t* qs; const int* flags;
    int count, next; pthread_mutex_t mu;
} xform_t;
void orc_ntt(uint64_t* data, uint64_t n, uint64_t q, int inverse);

static void* xform_worker(void* arg) {
    xform_t* J = (xform_t*)arg;
    for (;;) {
        pthread_mutex_lock(&J->mu);
        int i = J->next++;
        pthread_mutex_unlock(&J->mu);
        if (i >= J->count) break;
        const int inv = J->flags[i] & 1, nega = (J->flags[i] >> 1) & 1;
        if (nega) orc_ntt_negacyclic(J->data[i], J->ns[i], J->qs[i], inv);
        else orc_ntt(J->data[i], J->ns[i], J->qs[i], inv);
    }
    return NULL;
}

double orc_transform_batch(uint64_t* const* data, const uint64_t* ns, const uint64_t* qs, const int* flags,
                           int count, int threads) {
    xform_t J = {data, ns, qs, flags, count, 0, PTHREAD_MUTEX_INITIALIZER};
    pthread_t th[256];
    if (threads < 1) threads = 1;
    if (threads > 256) threads = 256;
    double t0 = now_s();
    for (int t = 0; t < threads; t++) pthread_create(&th[t], NULL, xform_worker, &J);
    for (int t = 0; t < threads; t++) pthread_join(th[t], NULL);
    return now_s() - t0;
}
