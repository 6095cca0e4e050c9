/*
 * ntt_oracle.c — CPU restatement of the reference NTT, used ONLY as the
 * checker by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
 * Nothing in the shipped library links or calls this file.
 *
 * Every function cites the reference file:line it restates.  Arithmetic is
 * plain C with unsigned __int128, exactly the operations the reference host
 * code performs, so a result here is the reference's result on the same input.
 *
 * Pinning (what checks each function independently of this text — see
 * tests/test_oracle.py):
 *   orc_find_generator / orc_create_params : BabyBear g=31 and independently
 *        computed omega/psi values; brute-force order check.
 *   orc_ntt_host    : O(N^2) DFT (orc_dft_naive is textbook, not a restatement)
 *                     at N<=2^10; closed-form KATs (x_j=j+1, delta, constant);
 *                     round trip; the pow-twiddle DIT of tests/test_inner_ntt.cu.
 *   orc_ntt_negacyclic : schoolbook product mod x^N+1 (tests/test_ntt.cu:186-199)
 *                     and the direct sum  A_k = sum_j a_j psi^{j(2k+1)}.
 *   orc_arb_*       : K-limb Montgomery identities, Goldilocks/BN254 root orders,
 *                     and agreement with the 64-bit path on the same prime.
 *   orc_bconv, orc_pmac : no reference code to restate (the reference names
 *                     BConv/IP and PMAC at src/ntt_optimized.cu:510-513 as
 *                     absent), so parity with the reference is unpinned; they
 *                     restate the standard definitions and are pinned by big-
 *                     integer CRT (S = sum t_i Q/q_i is X + uQ with 0 <= u < L)
 *                     and by direct Python-integer sums in tests/test_oracle.py.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ModularParams, include/modular_arith.cuh:11-19 (same field order). */
typedef struct {
    uint64_t q, q_twice, barrett_k, barrett_m, omega, omega_inv, n_inv;
} orc_params;

/* mod_pow, include/modular_arith.cuh:110-121 */
uint64_t orc_mod_pow(uint64_t base, uint64_t exp, uint64_t mod) {
    uint64_t result = 1;
    base %= mod;
    while (exp > 0) {
        if (exp & 1) result = (uint64_t)((u128)result * base % mod);
        base = (uint64_t)((u128)base * base % mod);
        exp >>= 1;
    }
    return result;
}

/* mod_inverse (Fermat), include/modular_arith.cuh:123-125 */
uint64_t orc_mod_inverse(uint64_t a, uint64_t mod) { return orc_mod_pow(a, mod - 2, mod); }

/* Generator search of find_primitive_root, include/modular_arith.cuh:130-147:
 * smallest g >= 2 with g^((q-1)/p) != 1 for every p in {2,3,5,7} dividing q-1. */
uint64_t orc_find_generator(uint64_t q) {
    static const uint64_t ps[4] = {2, 3, 5, 7};
    uint64_t g = 2;
    for (;;) {
        int ok = 1;
        uint64_t order = q - 1;
        for (int i = 0; i < 4; i++) {
            if (order % ps[i] == 0 && orc_mod_pow(g, order / ps[i], q) == 1) { ok = 0; break; }
        }
        if (ok) return g;
        g++;
    }
}

/* find_primitive_root, include/modular_arith.cuh:148-150: omega = g^((q-1)/n). */
uint64_t orc_primitive_root(uint64_t q, uint64_t n) {
    return orc_mod_pow(orc_find_generator(q), (q - 1) / n, q);
}

/* create_params, include/modular_arith.cuh:153-170 */
void orc_create_params(uint64_t q, uint64_t n, orc_params* p) {
    p->q = q;
    p->q_twice = 2 * q;
    int k = 64;
    while (((uint64_t)1 << (k - 64)) < q && k < 96) k++;  /* k-64 < 64 for every q < 2^63 */
    p->barrett_k = (uint64_t)k;
    p->barrett_m = (uint64_t)(((u128)1 << k) / q);
    p->omega = orc_primitive_root(q, n);
    p->omega_inv = orc_mod_inverse(p->omega, q);
    p->n_inv = orc_mod_inverse(n, q);
}

/* mu = floor(2^64/q), src/ntt_basic.cu:226-229 */
uint64_t orc_mu(uint64_t q) { return (uint64_t)(((u128)1 << 64) / q); }

/* Negacyclic root psi = g^((q-1)/(2n)) (psi^2 = omega of create_params, psi^n = -1).
 * The reference defines no negacyclic transform (tests/test_ntt.cu:304-307);
 * this is the definition the library adopts, built on the same g-rule. */
uint64_t orc_psi(uint64_t q, uint64_t n) {
    return orc_mod_pow(orc_find_generator(q), (q - 1) / (2 * n), q);
}

static int ilog2_u64(uint64_t n) { int l = 0; while (((uint64_t)1 << l) < n) l++; return l; }

/* ntt_host, src/ntt_basic.cu:412-463: in-place bit reversal, radix-2 DIT with
 * per-stage w_base = omega^(n/2^(s+1)), canonical add/sub, n^-1 scale on inverse.
 * The sum u + v is formed in 128 bits: the same result for q < 2^63, and exact
 * for the Goldilocks prime, where the reference's uint64 sum would wrap. */
void orc_ntt_host(uint64_t* data, uint64_t n, const orc_params* prm, int inverse) {
    int log_n = ilog2_u64(n);
    uint64_t q = prm->q;
    uint64_t omega = inverse ? prm->omega_inv : prm->omega;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t rev = 0, tmp = i;
        for (int j = 0; j < log_n; j++) { rev = (rev << 1) | (tmp & 1); tmp >>= 1; }
        if (i < rev) { uint64_t t = data[i]; data[i] = data[rev]; data[rev] = t; }
    }
    for (int stage = 0; stage < log_n; stage++) {
        uint64_t half = (uint64_t)1 << stage, full = half << 1, tw_stride = n / full;
        uint64_t w_base = orc_mod_pow(omega, tw_stride, q);
        for (uint64_t group = 0; group < n / full; group++) {
            uint64_t w = 1;
            for (uint64_t pos = 0; pos < half; pos++) {
                uint64_t i = group * full + pos, j = i + half;
                uint64_t u = data[i];
                uint64_t v = (uint64_t)((u128)data[j] * w % q);
                data[i] = (uint64_t)(((u128)u + v) % q);
                data[j] = (u >= v) ? (u - v) : (u + q - v);  /* < q: exact modulo 2^64 */
                w = (uint64_t)((u128)w * w_base % q);
            }
        }
    }
    if (inverse)
        for (uint64_t i = 0; i < n; i++) data[i] = (uint64_t)((u128)data[i] * prm->n_inv % q);
}

/* Convenience: parameters from (q, n), then ntt_host — what tests/test_ntt.cu:129-137 does. */
void orc_ntt(uint64_t* data, uint64_t n, uint64_t q, int inverse) {
    orc_params p;
    orc_create_params(q, n, &p);
    orc_ntt_host(data, n, &p, inverse);
}

/* Negacyclic transform over Z_q[x]/(x^n+1), natural order both ways:
 *   forward  A = ntt_host(a_j * psi^j)
 *   inverse  a_j = psi^-j * ntt_host_inverse(A)_j
 * ntt_host as above (src/ntt_basic.cu:412-463); psi per orc_psi. */
void orc_ntt_negacyclic(uint64_t* data, uint64_t n, uint64_t q, int inverse) {
    orc_params p;
    orc_create_params(q, n, &p);
    uint64_t psi = orc_psi(q, n);
    if (!inverse) {
        uint64_t t = 1;
        for (uint64_t j = 0; j < n; j++) { data[j] = (uint64_t)((u128)data[j] * t % q); t = (uint64_t)((u128)t * psi % q); }
        orc_ntt_host(data, n, &p, 0);
    } else {
        orc_ntt_host(data, n, &p, 1);
        uint64_t psi_inv = orc_mod_inverse(psi, q), t = 1;
        for (uint64_t j = 0; j < n; j++) { data[j] = (uint64_t)((u128)data[j] * t % q); t = (uint64_t)((u128)t * psi_inv % q); }
    }
}

/* poly_mul_kernel semantics, src/ntt_basic.cu:392-398: r_i = a_i*b_i mod q. */
void orc_pointwise_mul(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; i++) r[i] = (uint64_t)((u128)a[i] * b[i] % q);
}

/* Schoolbook product mod x^n+1, tests/test_ntt.cu:186-199. */
void orc_negacyclic_schoolbook(const uint64_t* a, const uint64_t* b, uint64_t* out, uint64_t n, uint64_t q) {
    for (uint64_t i = 0; i < n; i++) out[i] = 0;
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t j = 0; j < n; j++) {
            uint64_t prod = (uint64_t)((u128)a[i] * b[j] % q);
            uint64_t idx = (i + j) % n;
            if (i + j >= n) out[idx] = (uint64_t)(((u128)out[idx] + q - prod) % q);
            else            out[idx] = (uint64_t)(((u128)out[idx] + prod) % q);
        }
}

/* Textbook O(n^2) DFT X_k = sum_j x_j w^(jk) — an independent definition used
 * to pin orc_ntt_host, not a restatement of reference code. */
void orc_dft_naive(const uint64_t* x, uint64_t* X, uint64_t n, uint64_t q, uint64_t w) {
    for (uint64_t k = 0; k < n; k++) {
        uint64_t wk = orc_mod_pow(w, k, q), t = 1;
        u128 acc = 0;
        for (uint64_t j = 0; j < n; j++) {
            acc += (u128)x[j] * t % q;
            t = (uint64_t)((u128)t * wk % q);
            if ((j & 1023) == 1023) acc %= q;
        }
        X[k] = (uint64_t)(acc % q);
    }
}

/* Independent DIT with per-butterfly pow twiddles, tests/test_inner_ntt.cu:12-75. */
void orc_ntt_pow_twiddle(uint64_t* a, uint64_t n, uint64_t q, uint64_t omega) {
    int log_n = ilog2_u64(n);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t r = 0;
        for (int b = 0; b < log_n; b++) r |= ((i >> b) & 1) << (log_n - 1 - b);
        if (i < r) { uint64_t t = a[i]; a[i] = a[r]; a[r] = t; }
    }
    for (uint64_t len = 2; len <= n; len <<= 1) {
        uint64_t half = len >> 1;
        for (uint64_t s = 0; s < n; s += len)
            for (uint64_t k = 0; k < half; k++) {
                uint64_t w = orc_mod_pow(omega, k * (n / len), q);
                uint64_t u = a[s + k], v = (uint64_t)((u128)a[s + k + half] * w % q);
                a[s + k] = (uint64_t)(((u128)u + v) % q);
                a[s + k + half] = (uint64_t)(((u128)u + q - v) % q);
            }
    }
}

/* ---- portable input generator (identical on every host) ----
 * splitmix64 (Steele, Lea, Flood 2014) stream; value drawn by rejection into
 * [0,q): accept x < floor(2^64/q)*q, return x mod q. */
static uint64_t splitmix64_next(uint64_t* s) {
    uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
void orc_random_poly(uint64_t* out, uint64_t n, uint64_t q, uint64_t seed) {
    uint64_t s = seed;
    uint64_t lim = (uint64_t)((((u128)1 << 64) / q) * q);  /* q < 2^63 so lim fits */
    for (uint64_t i = 0; i < n; i++) {
        uint64_t x;
        do { x = splitmix64_next(&s); } while (x >= lim);
        out[i] = x % q;
    }
}
void orc_random_poly_u32(uint32_t* out, uint64_t n, uint64_t q, uint64_t seed) {
    uint64_t s = seed;
    uint64_t lim = (uint64_t)((((u128)1 << 64) / q) * q);
    for (uint64_t i = 0; i < n; i++) {
        uint64_t x;
        do { x = splitmix64_next(&s); } while (x >= lim);
        out[i] = (uint32_t)(x % q);
    }
}

/* Deterministic primality (Miller-Rabin with the 7 bases that are exact below 2^64). */
int orc_is_prime(uint64_t n) {
    if (n < 2) return 0;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (int i = 0; i < 12; i++) { if (n == small[i]) return 1; if (n % small[i] == 0) return 0; }
    uint64_t d = n - 1; int r = 0;
    while (!(d & 1)) { d >>= 1; r++; }
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (int i = 0; i < 7; i++) {
        uint64_t a = bases[i] % n;
        if (a == 0) continue;
        uint64_t x = orc_mod_pow(a, d, n);
        if (x == 1 || x == n - 1) continue;
        int comp = 1;
        for (int k = 1; k < r; k++) { x = (uint64_t)((u128)x * x % n); if (x == n - 1) { comp = 0; break; } }
        if (comp) return 0;
    }
    return 1;
}

/* The `count` largest primes below `below` with q = 1 (mod 2^two_adic), descending. */
int orc_ntt_primes(uint64_t below, int two_adic, int count, uint64_t* out) {
    uint64_t step = (uint64_t)1 << two_adic;
    uint64_t q = ((below - 1) / step) * step + 1;
    if (q >= below) q -= step;
    int got = 0;
    while (got < count && q > step) {
        if (orc_is_prime(q)) out[got++] = q;
        q -= step;
    }
    return got;
}

/* ===================== K-limb (include/ntt_arbitrary.cuh) =====================
 * Values are K little-endian limbs of LIMB_BITS=24 bits, one per uint32
 * (DScalar<K>, ntt_arbitrary.cuh:51,573-576).  K <= ORC_KMAX. */
#define ORC_LIMB_BITS 24
#define ORC_LIMB_MASK ((1ULL << ORC_LIMB_BITS) - 1)
#define ORC_KMAX 12

/* mont_np, ntt_arbitrary.cuh:456-462 */
uint32_t orc_arb_mont_np(const uint32_t* p) {
    uint32_t p0 = p[0], x = 1;
    for (int k = 0; k < 6; k++) x = x * (2u - p0 * x);
    return (0u - x) & (uint32_t)ORC_LIMB_MASK;
}

/* mont_mul<K> (CIOS), ntt_arbitrary.cuh:1218-1298: out = a*b*R^-1 mod p, R=2^(24K). */
void orc_arb_mont_mul(uint32_t* out, const uint32_t* a, const uint32_t* b, const uint32_t* p,
                      uint32_t np, int K) {
    uint64_t T[ORC_KMAX + 2];
    for (int i = 0; i < K + 2; i++) T[i] = 0;
    for (int i = 0; i < K; i++) {
        uint64_t ai = a[i], carry = 0;
        for (int j = 0; j < K; j++) {
            uint64_t v = T[j] + ai * b[j] + carry;
            T[j] = v & ORC_LIMB_MASK; carry = v >> ORC_LIMB_BITS;
        }
        { uint64_t v = T[K] + carry; T[K] = v & ORC_LIMB_MASK; T[K + 1] += v >> ORC_LIMB_BITS; }
        uint64_t m = (T[0] * (uint64_t)np) & ORC_LIMB_MASK;
        carry = 0;
        for (int j = 0; j < K; j++) {
            uint64_t v = T[j] + m * (uint64_t)p[j] + carry;
            T[j] = v & ORC_LIMB_MASK; carry = v >> ORC_LIMB_BITS;
        }
        { uint64_t v = T[K] + carry; T[K] = v & ORC_LIMB_MASK; T[K + 1] += v >> ORC_LIMB_BITS; }
        for (int j = 0; j < K + 1; j++) T[j] = T[j + 1];
        T[K + 1] = 0;
    }
    int ge = (T[K] != 0);
    if (!ge) {
        ge = 1;
        for (int i = K - 1; i >= 0; i--) if (T[i] != p[i]) { ge = T[i] > p[i]; break; }
    }
    if (ge) {
        int64_t borrow = 0;
        for (int i = 0; i < K; i++) {
            int64_t s = (int64_t)T[i] - (int64_t)p[i] - borrow;
            if (s < 0) { s += (int64_t)(1LL << ORC_LIMB_BITS); borrow = 1; } else borrow = 0;
            out[i] = (uint32_t)s;
        }
    } else {
        for (int i = 0; i < K; i++) out[i] = (uint32_t)T[i];
    }
}

static int arb_cmp(const uint32_t* a, const uint32_t* b, int K) {
    for (int i = K - 1; i >= 0; i--) { if (a[i] < b[i]) return -1; if (a[i] > b[i]) return 1; }
    return 0;
}
/* mod_add / mod_sub, ntt_arbitrary.cuh:300-328 (canonical inputs). */
static void arb_add(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* p, int K) {
    uint32_t t[ORC_KMAX]; uint64_t c = 0;
    for (int i = 0; i < K; i++) { uint64_t s = (uint64_t)a[i] + b[i] + c; t[i] = (uint32_t)(s & ORC_LIMB_MASK); c = s >> ORC_LIMB_BITS; }
    if (c || arb_cmp(t, p, K) >= 0) {
        int64_t br = 0;
        for (int i = 0; i < K; i++) { int64_t s = (int64_t)t[i] - p[i] - br; if (s < 0) { s += 1LL << ORC_LIMB_BITS; br = 1; } else br = 0; t[i] = (uint32_t)s; }
    }
    memcpy(r, t, sizeof(uint32_t) * K);
}
static void arb_sub(uint32_t* r, const uint32_t* a, const uint32_t* b, const uint32_t* p, int K) {
    uint32_t t[ORC_KMAX]; int64_t br = 0;
    for (int i = 0; i < K; i++) { int64_t s = (int64_t)a[i] - b[i] - br; if (s < 0) { s += 1LL << ORC_LIMB_BITS; br = 1; } else br = 0; t[i] = (uint32_t)s; }
    if (br) {
        uint64_t c = 0;
        for (int i = 0; i < K; i++) { uint64_t s = (uint64_t)t[i] + p[i] + c; t[i] = (uint32_t)(s & ORC_LIMB_MASK); c = s >> ORC_LIMB_BITS; }
    }
    memcpy(r, t, sizeof(uint32_t) * K);
}

/* host_ntt_naive<K>, ntt_arbitrary.cuh:2651-2676, restated on Montgomery-form
 * data (R = 2^(24K)) with Montgomery-form twiddles tw_m[k] = omega^k * R:
 * mont_mul(a*R, w*R) = a*w*R, so every value stays in Montgomery form and the
 * result, converted out, equals host_ntt_naive on the natural-form values. */
void orc_arb_ntt_mont(uint32_t* a, const uint32_t* tw_m, uint32_t n, const uint32_t* p, uint32_t np,
                      int K) {
    int log_n = ilog2_u64(n);
    uint32_t t[ORC_KMAX], u[ORC_KMAX];
    for (uint32_t i = 0; i < n; i++) {
        uint32_t j = 0;
        for (int b = 0; b < log_n; b++) j = (j << 1) | ((i >> b) & 1);
        if (i < j) for (int l = 0; l < K; l++) { uint32_t x = a[i * K + l]; a[i * K + l] = a[j * K + l]; a[j * K + l] = x; }
    }
    for (int s = 0; s < log_n; s++) {
        uint32_t m = 1u << (s + 1), half = m >> 1, tw_stride = n / m;
        for (uint32_t g = 0; g < n; g += m)
            for (uint32_t k = 0; k < half; k++) {
                const uint32_t* w = tw_m + (size_t)((uint64_t)k * tw_stride % n) * K;
                orc_arb_mont_mul(t, a + (size_t)(g + k + half) * K, w, p, np, K);
                memcpy(u, a + (size_t)(g + k) * K, sizeof(uint32_t) * K);
                arb_add(a + (size_t)(g + k) * K, u, t, p, K);
                arb_sub(a + (size_t)(g + k + half) * K, u, t, p, K);
            }
    }
}

/* Fast RNS base conversion from base {q_i} (i < L) to base {p_j} (j < K), the
 * BConv of the COOP step named at reference src/ntt_optimized.cu:510-513
 * (no code there): with Q = prod q_i and qh_i = Q / q_i,
 *   t_i = x_i * (qh_i^-1 mod q_i) mod q_i,   y_j = sum_i t_i * (qh_i mod p_j) mod p_j,
 * so y_j = (X + u*Q) mod p_j for the CRT value X of x and some 0 <= u < L.
 * Limb-major layout: in[i*n + c], out[j*n + c]; moduli are primes. */
void orc_bconv(const uint32_t* in, uint32_t* out, uint64_t n, const uint64_t* q, uint32_t L, const uint64_t* p,
               uint32_t K) {
    uint64_t* w = (uint64_t*)malloc(sizeof(uint64_t) * (L ? L : 1));
    uint64_t* m = (uint64_t*)malloc(sizeof(uint64_t) * (L && K ? (uint64_t)L * K : 1));
    for (uint32_t i = 0; i < L; i++) {
        uint64_t h = 1;
        for (uint32_t k = 0; k < L; k++)
            if (k != i) h = (uint64_t)((u128)h * (q[k] % q[i]) % q[i]);
        w[i] = orc_mod_inverse(h, q[i]);
        for (uint32_t j = 0; j < K; j++) {
            uint64_t g = 1;
            for (uint32_t k = 0; k < L; k++)
                if (k != i) g = (uint64_t)((u128)g * (q[k] % p[j]) % p[j]);
            m[(uint64_t)j * L + i] = g;
        }
    }
    uint64_t* t = (uint64_t*)malloc(sizeof(uint64_t) * (L ? L : 1));
    for (uint64_t c = 0; c < n; c++) {
        for (uint32_t i = 0; i < L; i++) t[i] = (uint64_t)((u128)in[(uint64_t)i * n + c] * w[i] % q[i]);
        for (uint32_t j = 0; j < K; j++) {
            u128 acc = 0;
            for (uint32_t i = 0; i < L; i++) acc += (u128)t[i] * m[(uint64_t)j * L + i];
            out[(uint64_t)j * n + c] = (uint32_t)(acc % p[j]);
        }
    }
    free(t);
    free(m);
    free(w);
}

/* Multiply-accumulate over dnum polynomial pairs per RNS limb, the inner
 * product (IP) / PMAC step named at reference src/ntt_optimized.cu:510-513
 * (no code there):  out[l][c] = (acc ? out[l][c] : 0) + sum_k a[k][l][c] b[k][l][c]  mod q_l,
 * layouts a[(k*L + l)*n + c], out[l*n + c]; inputs canonical. */
void orc_pmac(uint32_t* out, const uint32_t* a, const uint32_t* b, uint32_t dnum, uint32_t L, uint64_t n,
              const uint64_t* q, int acc) {
    for (uint32_t l = 0; l < L; l++)
        for (uint64_t c = 0; c < n; c++) {
            u128 s = acc ? out[(uint64_t)l * n + c] : 0;
            for (uint32_t k = 0; k < dnum; k++) {
                const uint64_t at = ((uint64_t)k * L + l) * n + c;
                s += (u128)a[at] * b[at];
            }
            out[(uint64_t)l * n + c] = (uint32_t)(s % q[l]);
        }
}
