// ntt_arbitrary.cuh — drop-in for the reference's include/ntt_arbitrary.cuh
// (namespace ntt_arb) on MI355X.
//
// Same names, template parameter K (number of 24-bit limbs), value layout
// (DScalar<K>: K uint32 limbs, little-endian, Montgomery form x*R mod p with
// R = 2^(24K)) and host helpers as the reference, with hipStream_t in place of
// cudaStream_t.  Device work forwards to libntt_mi355x.so (ntt_c_arb_*,
// ntt_c.h); the host arithmetic below is plain C++ for setup and checking.
//
// Differences from the reference, all deliberate:
//  * FieldConfig<2>: the reference's 49-bit prime 281476520214529 cannot be
//    held in two 24-bit limbs and is truncated there to 1543503873, which is
//    composite (reference :488-507).  Here K=2 uses the 48-bit prime
//    0xffffd8000001 = 16777176 * 2^24 + 1 (generator 5), MAX_LOG_N 24.
//  * Inverse and negacyclic transforms exist (run_inverse_ct,
//    run_forward_negacyclic, run_inverse_negacyclic); the reference is
//    forward-only.
//  * NTTBuffers keeps d_tfm8 / d_hada64 for source compatibility but leaves
//    them null: the library's kernels need no per-size MMA tables, and
//    run_forward_ct_tcu accepts and ignores them.
//  * setup_ntt also builds a plan (ntt_c_arb_plan_create) when p fits 64-bit
//    arithmetic (K <= 3: the K = 2 prime, Goldilocks), and run_forward /
//    run_inverse, run_for_K and the negacyclic helpers run it on the
//    four-step pass kernels; run_forward_ct / run_forward_ct_tcu keep the
//    limb kernels.  Both give the same limbs.
//  * Failed library calls print ntt_c_last_error() and abort() unless
//    NTT_NO_ABORT is defined (the reference does no error checking here).
#pragma once

#include <hip/hip_runtime_api.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iomanip>
#include <iostream>
#include <random>
#include <vector>

#include "ntt_c.h"

namespace ntt_arb {

constexpr int LIMB_BITS = 24;  // reference :51-53
constexpr uint64_t LIMB_BASE = 1ULL << LIMB_BITS;
constexpr uint64_t LIMB_MASK = LIMB_BASE - 1;

inline ntt_status& last_status() {
    static thread_local ntt_status s = NTT_OK;
    return s;
}

inline void check_(ntt_status s, const char* what) {
    last_status() = s;
    if (s != NTT_OK) {
        std::fprintf(stderr, "ntt_arb: %s failed (%d): %s\n", what, (int)s, ntt_c_last_error());
#ifndef NTT_NO_ABORT
        std::abort();
#endif
    }
}

inline void hip_check_(hipError_t e, const char* what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "ntt_arb: %s: %s\n", what, hipGetErrorString(e));
        last_status() = NTT_ERR_DEVICE;
#ifndef NTT_NO_ABORT
        std::abort();
#endif
    }
}

// ---------------------------------------------------------------- BigInt<K>
// K-limb unsigned integer, limbs[0] least significant (reference :149-183).
template <int K>
struct BigInt {
    uint64_t limbs[K];

    BigInt() {
        for (int i = 0; i < K; i++) limbs[i] = 0;
    }
    static BigInt zero() { return BigInt(); }
    static BigInt one() {
        BigInt r;
        r.limbs[0] = 1;
        return r;
    }
    static BigInt from_u64(uint64_t v) {
        BigInt r;
        for (int i = 0; i < K && v; i++, v >>= LIMB_BITS) r.limbs[i] = v & LIMB_MASK;
        return r;
    }
    bool is_zero() const {
        for (int i = 0; i < K; i++)
            if (limbs[i]) return false;
        return true;
    }
    int cmp(const BigInt& o) const {
        for (int i = K - 1; i >= 0; i--)
            if (limbs[i] != o.limbs[i]) return limbs[i] < o.limbs[i] ? -1 : 1;
        return 0;
    }
    // bit i (0 = least significant)
    int bit(int i) const { return (int)((limbs[i / LIMB_BITS] >> (i % LIMB_BITS)) & 1); }
};

// Hex string (most significant digit first, optional 0x) -> BigInt<K>;
// digits beyond K limbs are dropped (reference :194-222).
template <int K>
inline BigInt<K> from_hex(const char* hex) {
    if (hex[0] == '0' && (hex[1] == 'x' || hex[1] == 'X')) hex += 2;
    BigInt<K> r;
    const int len = (int)std::strlen(hex);
    for (int d = 0; d < len; d++) {
        const char c = hex[len - 1 - d];
        const int v = (c >= '0' && c <= '9') ? c - '0'
                      : (c >= 'a' && c <= 'f') ? c - 'a' + 10
                      : (c >= 'A' && c <= 'F') ? c - 'A' + 10
                                               : 0;
        const int bitpos = 4 * d, li = bitpos / LIMB_BITS, sh = bitpos % LIMB_BITS;
        // a hex digit never straddles a limb boundary: 24 is a multiple of 4
        if (li < K) r.limbs[li] |= (uint64_t)v << sh;
    }
    return r;
}

// r = a + b; returns the carry out of the top limb (reference :225-237).
template <int K>
inline uint64_t big_add(BigInt<K>& r, const BigInt<K>& a, const BigInt<K>& b) {
    uint64_t c = 0;
    for (int i = 0; i < K; i++) {
        const uint64_t s = a.limbs[i] + b.limbs[i] + c;
        r.limbs[i] = s & LIMB_MASK;
        c = s >> LIMB_BITS;
    }
    return c;
}

// r = a - b; returns the borrow (reference :239-252).
template <int K>
inline uint64_t big_sub(BigInt<K>& r, const BigInt<K>& a, const BigInt<K>& b) {
    uint64_t br = 0;
    for (int i = 0; i < K; i++) {
        const uint64_t sub = b.limbs[i] + br;
        br = a.limbs[i] < sub;
        r.limbs[i] = (a.limbs[i] + (br << LIMB_BITS) - sub) & LIMB_MASK;
    }
    return br;
}

// Full product into 2K limbs (reference :267-294).
template <int K>
inline void big_mul(BigInt<2 * K>& r, const BigInt<K>& a, const BigInt<K>& b) {
    r = BigInt<2 * K>();
    for (int i = 0; i < K; i++) {
        uint64_t c = 0;
        for (int j = 0; j < K; j++) {
            const uint64_t v = r.limbs[i + j] + a.limbs[i] * b.limbs[j] + c;
            r.limbs[i + j] = v & LIMB_MASK;
            c = v >> LIMB_BITS;
        }
        r.limbs[i + K] = c;
    }
}

// a mod p for canonical-or-below-2p input (reference :300-308).
template <int K>
inline BigInt<K> mod_reduce_once(const BigInt<K>& a, const BigInt<K>& p) {
    if (a.cmp(p) < 0) return a;
    BigInt<K> r;
    big_sub<K>(r, a, p);
    return r;
}

template <int K>
inline BigInt<K> mod_add(const BigInt<K>& a, const BigInt<K>& b, const BigInt<K>& p) {
    BigInt<K> r;
    const uint64_t c = big_add<K>(r, a, b);
    if (c || r.cmp(p) >= 0) big_sub<K>(r, r, p);
    return r;
}

template <int K>
inline BigInt<K> mod_sub(const BigInt<K>& a, const BigInt<K>& b, const BigInt<K>& p) {
    BigInt<K> r;
    if (big_sub<K>(r, a, b)) big_add<K>(r, r, p);
    return r;
}

// Full reduction of a K_IN-limb value modulo a K_OUT-limb p, by binary long
// division: r <- 2r + bit, minus p when r >= p (reference mod_slow :331-383).
template <int K_OUT, int K_IN>
inline BigInt<K_OUT> mod_slow(const BigInt<K_IN>& a, const BigInt<K_OUT>& p) {
    BigInt<K_OUT + 1> r, pp;
    for (int i = 0; i < K_OUT; i++) pp.limbs[i] = p.limbs[i];
    for (int bitpos = K_IN * LIMB_BITS - 1; bitpos >= 0; bitpos--) {
        uint64_t c = (uint64_t)a.bit(bitpos);
        for (int i = 0; i < K_OUT + 1; i++) {
            const uint64_t v = (r.limbs[i] << 1) | c;
            r.limbs[i] = v & LIMB_MASK;
            c = v >> LIMB_BITS;
        }
        if (r.cmp(pp) >= 0) big_sub<K_OUT + 1>(r, r, pp);
    }
    BigInt<K_OUT> out;
    for (int i = 0; i < K_OUT; i++) out.limbs[i] = r.limbs[i];
    return out;
}

template <int K>
inline BigInt<K> mod_mul(const BigInt<K>& a, const BigInt<K>& b, const BigInt<K>& p) {
    BigInt<2 * K> w;
    big_mul<K>(w, a, b);
    return mod_slow<K, 2 * K>(w, p);
}

template <int K>
inline BigInt<K> mod_pow(const BigInt<K>& base, uint64_t exp, const BigInt<K>& p) {
    BigInt<K> r = mod_reduce_once<K>(BigInt<K>::one(), p), b = base;
    for (; exp; exp >>= 1) {
        if (exp & 1) r = mod_mul<K>(r, b, p);
        b = mod_mul<K>(b, b, p);
    }
    return r;
}

template <int K>
inline BigInt<K> mod_pow_big(const BigInt<K>& base, const BigInt<K>& exp, const BigInt<K>& p) {
    BigInt<K> r = mod_reduce_once<K>(BigInt<K>::one(), p);
    for (int i = K * LIMB_BITS - 1; i >= 0; i--) {
        r = mod_mul<K>(r, r, p);
        if (exp.bit(i)) r = mod_mul<K>(r, base, p);
    }
    return r;
}

// a^-1 = a^(p-2) mod p for prime p (reference :422-430).
template <int K>
inline BigInt<K> mod_inv(const BigInt<K>& a, const BigInt<K>& p) {
    BigInt<K> e;
    big_sub<K>(e, p, BigInt<K>::from_u64(2));
    return mod_pow_big<K>(a, e, p);
}

// ---------------------------------------------------------------- Montgomery
// -p^-1 mod 2^24 by Newton iteration x <- x(2 - p x) (reference :456-462).
template <int K>
inline uint32_t mont_np(const BigInt<K>& p) {
    const uint32_t p0 = (uint32_t)p.limbs[0];
    uint32_t x = p0;  // correct to 3 bits for odd p0; each step doubles the bits
    for (int it = 0; it < 4; it++) x *= 2u - p0 * x;
    return (uint32_t)((0u - x) & LIMB_MASK);
}

// R^2 mod p, R = 2^(24K) (reference :466-474).
template <int K>
inline BigInt<K> mont_R2(const BigInt<K>& p) {
    BigInt<2 * K + 1> r2;
    r2.limbs[2 * K] = 1;
    return mod_slow<K, 2 * K + 1>(r2, p);
}

// a*b*R^-1 mod p on the host, CIOS (the device product's definition,
// reference mont_mul :1218-1298 / mont_mul_big :1301-1311).
template <int K>
inline BigInt<K> mont_mul_big(const BigInt<K>& a, const BigInt<K>& b, const BigInt<K>& p, uint32_t np) {
    uint64_t t[K + 2] = {};
    for (int i = 0; i < K; i++) {
        uint64_t c = 0;
        for (int j = 0; j < K; j++) {
            const uint64_t v = t[j] + a.limbs[i] * b.limbs[j] + c;
            t[j] = v & LIMB_MASK;
            c = v >> LIMB_BITS;
        }
        uint64_t v = t[K] + c;
        t[K] = v & LIMB_MASK;
        t[K + 1] += v >> LIMB_BITS;
        const uint64_t m = (t[0] * np) & LIMB_MASK;
        c = 0;
        for (int j = 0; j < K; j++) {
            const uint64_t w = t[j] + m * p.limbs[j] + c;
            t[j] = w & LIMB_MASK;
            c = w >> LIMB_BITS;
        }
        v = t[K] + c;
        t[K] = v & LIMB_MASK;
        t[K + 1] += v >> LIMB_BITS;
        for (int j = 0; j <= K; j++) t[j] = t[j + 1];
        t[K + 1] = 0;
    }
    BigInt<K> r;
    for (int j = 0; j < K; j++) r.limbs[j] = t[j];
    if (t[K] || r.cmp(p) >= 0) big_sub<K>(r, r, p);
    return r;
}

// ---------------------------------------------------------------- fields
// FieldConfig<K>: prime and a primitive 2^MAX_LOG_N-th root (reference :479-568).
template <int K_>
struct FieldConfig;

template <>
struct FieldConfig<2> {
    static constexpr int K = 2;
    static constexpr int MAX_LOG_N = 24;
    static BigInt<2> prime() { return BigInt<2>::from_u64(0xFFFFD8000001ULL); }
    // 5^((p-1)/2^24) mod p
    static BigInt<2> omega_2pow_max() { return BigInt<2>::from_u64(280117225980523ULL); }
};

template <>
struct FieldConfig<3> {  // Goldilocks 2^64 - 2^32 + 1
    static constexpr int K = 3;
    static constexpr int MAX_LOG_N = 24;
    static BigInt<3> prime() { return BigInt<3>::from_u64(0xFFFFFFFF00000001ULL); }
    static BigInt<3> omega_2pow_max() { return BigInt<3>::from_u64(0x86CDCC31C307E171ULL); }
};

template <>
struct FieldConfig<11> {  // BN254 scalar field
    static constexpr int K = 11;
    static constexpr int MAX_LOG_N = 24;
    static BigInt<11> prime() {
        return from_hex<11>("30644E72E131A029B85045B68181585D2833E84879B9709143E1F593F0000001");
    }
    // the field's 2^28-th root of unity squared four times
    static BigInt<11> omega_2pow_max() {
        BigInt<11> w = from_hex<11>("2A3C09F0A58A7E8500E0A7EB8EF62ABC402D111E41112ED49BD61B6E725B19F0");
        for (int i = 0; i < 28 - MAX_LOG_N; i++) w = mod_mul<11>(w, w, prime());
        return w;
    }
};

// ---------------------------------------------------------------- device values
template <int K>
struct DScalar {
    uint32_t limbs[K];
};

template <int K>
inline DScalar<K> load_scalar(const BigInt<K>& x) {
    DScalar<K> r;
    for (int i = 0; i < K; i++) r.limbs[i] = (uint32_t)x.limbs[i];
    return r;
}

template <int K>
inline BigInt<K> store_scalar(const DScalar<K>& x) {
    BigInt<K> r;
    for (int i = 0; i < K; i++) r.limbs[i] = x.limbs[i];
    return r;
}

// ---------------------------------------------------------------- device transforms
// Forward NTT, natural order in and out, in place (reference :2547-2590).
template <int K>
inline void run_forward_ct(DScalar<K>* d_data, const DScalar<K>* d_tw, uint32_t n, int log_n, DScalar<K> prime_d,
                           uint32_t np, hipStream_t stream = 0) {
    check_(ntt_c_arb_forward(d_data->limbs, d_tw->limbs, n, (uint32_t)log_n, prime_d.limbs, np, K, stream),
           "run_forward_ct");
}

// Same transform; the MMA tables are not needed (reference :2599-2646).
template <int K>
inline void run_forward_ct_tcu(DScalar<K>* d_data, const DScalar<K>* d_tw, const DScalar<K>* /*d_tfm8*/,
                               const DScalar<K>* /*d_hada64*/, uint32_t n, int log_n, DScalar<K> prime_d, uint32_t np,
                               hipStream_t stream = 0) {
    run_forward_ct<K>(d_data, d_tw, n, log_n, prime_d, np, stream);
}

// Inverse NTT (no reference counterpart): twiddles w^-k = d_tw[(n-k) % n],
// then the product by n_inv_mont = n^-1 R mod p.
template <int K>
inline void run_inverse_ct(DScalar<K>* d_data, const DScalar<K>* d_tw, uint32_t n, int log_n, DScalar<K> prime_d,
                           uint32_t np, DScalar<K> n_inv_mont, hipStream_t stream = 0) {
    check_(ntt_c_arb_inverse(d_data->limbs, d_tw->limbs, n, (uint32_t)log_n, prime_d.limbs, np, n_inv_mont.limbs, K,
                             stream),
           "run_inverse_ct");
}

// d_a[i] = d_a[i] * d_b[i] * R^-1 mod p
template <int K>
inline void pointwise_mul_mont(DScalar<K>* d_a, const DScalar<K>* d_b, uint32_t n, DScalar<K> prime_d, uint32_t np,
                               hipStream_t stream = 0) {
    check_(ntt_c_arb_pointwise_mul(d_a->limbs, d_b->limbs, n, prime_d.limbs, np, K, stream), "pointwise_mul_mont");
}

// d_a[i] = d_a[i] * s * R^-1 mod p (s = R^2 mod p: into Montgomery form; s = 1: out of it)
template <int K>
inline void scale_mont(DScalar<K>* d_a, uint32_t n, DScalar<K> s, DScalar<K> prime_d, uint32_t np,
                       hipStream_t stream = 0) {
    check_(ntt_c_arb_scale(d_a->limbs, n, s.limbs, prime_d.limbs, np, K, stream), "scale_mont");
}

// ---------------------------------------------------------------- host reference
// Radix-2 DIT on natural-form values: bit reversal, then stages with
// w = tw[k * n/m] (reference host_ntt_naive :2651-2676).
template <int K>
inline void host_ntt_naive(std::vector<BigInt<K>>& a, const std::vector<BigInt<K>>& tw, const BigInt<K>& p) {
    const size_t n = a.size();
    int log_n = 0;
    while (((size_t)1 << log_n) < n) log_n++;
    for (size_t i = 0; i < n; i++) {
        size_t j = 0;
        for (int b = 0; b < log_n; b++) j |= ((i >> b) & 1) << (log_n - 1 - b);
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t m = 2; m <= n; m <<= 1) {
        const size_t half = m / 2, stride = n / m;
        for (size_t g = 0; g < n; g += m)
            for (size_t k = 0; k < half; k++) {
                const BigInt<K> t = mod_mul<K>(a[g + k + half], tw[k * stride], p);
                const BigInt<K> u = a[g + k];
                a[g + k] = mod_add<K>(u, t, p);
                a[g + k + half] = mod_sub<K>(u, t, p);
            }
    }
}

// ---------------------------------------------------------------- buffers
// Owns the device input and twiddle table (reference NTTBuffers :2681-2696),
// plus what the inverse and negacyclic transforms need.
template <int K>
struct NTTBuffers {
    int n = 0;
    int log_n = 0;
    BigInt<K> prime{};
    BigInt<K> R2{};
    uint32_t np = 0;
    BigInt<K> omega_n{};
    std::vector<BigInt<K>> tw_host;    // w^k, natural form
    std::vector<BigInt<K>> data_host;  // input, natural form
    DScalar<K>* d_data = nullptr;      // Montgomery form
    DScalar<K>* d_tw = nullptr;        // w^k R mod p
    DScalar<K>* d_tfm8 = nullptr;      // unused here (kept for source compatibility)
    DScalar<K>* d_hada64 = nullptr;    // unused here
    DScalar<K> prime_d{};
    DScalar<K> n_inv_mont{};           // n^-1 R mod p
    BigInt<K> psi{};                   // primitive 2n-th root (zero if 2n exceeds the field's 2-adic range)
    DScalar<K>* d_psi = nullptr;       // psi^j R mod p
    DScalar<K>* d_psi_inv = nullptr;   // psi^-j R mod p
    ntt_c_arb_plan* plan = nullptr;    // four-step pass kernels, when p fits 64-bit arithmetic
};

template <int K>
inline DScalar<K> to_mont_d(const BigInt<K>& x, const NTTBuffers<K>& B) {
    return load_scalar<K>(mont_mul_big<K>(x, B.R2, B.prime, B.np));
}

template <int K>
inline void upload_mont_(DScalar<K>* d, const std::vector<BigInt<K>>& xs, const NTTBuffers<K>& B) {
    std::vector<DScalar<K>> h(xs.size());
    for (size_t i = 0; i < xs.size(); i++) h[i] = to_mont_d<K>(xs[i], B);
    hip_check_(hipMemcpy(d, h.data(), h.size() * sizeof(DScalar<K>), hipMemcpyHostToDevice), "upload");
}

// Twiddles and input on the host, Montgomery copies on the device
// (reference setup_ntt :2698-2793: input limbs from std::mt19937_64(seed),
// each masked to 24 bits, then reduced mod p).
template <int K>
inline NTTBuffers<K> setup_ntt(int log_n, uint64_t seed = 0xC0FFEE) {
    using FC = FieldConfig<K>;
    NTTBuffers<K> B;
    B.log_n = log_n;
    B.n = 1 << log_n;
    B.prime = FC::prime();
    B.np = mont_np<K>(B.prime);
    B.R2 = mont_R2<K>(B.prime);
    BigInt<K> w = FC::omega_2pow_max();
    if (log_n < FC::MAX_LOG_N) {
        for (int i = 0; i < FC::MAX_LOG_N - log_n - 1; i++) w = mod_mul<K>(w, w, B.prime);
        B.psi = w;
        w = mod_mul<K>(w, w, B.prime);
    } else {
        for (int i = 0; i < FC::MAX_LOG_N - log_n; i++) w = mod_mul<K>(w, w, B.prime);
    }
    B.omega_n = w;
    B.tw_host.resize(B.n);
    B.tw_host[0] = BigInt<K>::one();
    for (int i = 1; i < B.n; i++) B.tw_host[i] = mod_mul<K>(B.tw_host[i - 1], w, B.prime);
    std::mt19937_64 rng(seed);
    B.data_host.resize(B.n);
    for (int i = 0; i < B.n; i++) {
        BigInt<K> r;
        for (int j = 0; j < K; j++) r.limbs[j] = rng() & LIMB_MASK;
        B.data_host[i] = mod_slow<K, K>(r, B.prime);
    }
    B.prime_d = load_scalar<K>(B.prime);
    B.n_inv_mont = to_mont_d<K>(mod_inv<K>(BigInt<K>::from_u64((uint64_t)B.n), B.prime), B);
    const size_t bytes = (size_t)B.n * sizeof(DScalar<K>);
    hip_check_(hipMalloc((void**)&B.d_data, bytes), "hipMalloc");
    hip_check_(hipMalloc((void**)&B.d_tw, bytes), "hipMalloc");
    upload_mont_<K>(B.d_data, B.data_host, B);
    upload_mont_<K>(B.d_tw, B.tw_host, B);
    if (!B.psi.is_zero()) {
        std::vector<BigInt<K>> fw(B.n), iv(B.n);
        const BigInt<K> pinv = mod_inv<K>(B.psi, B.prime);
        fw[0] = iv[0] = BigInt<K>::one();
        for (int j = 1; j < B.n; j++) {
            fw[j] = mod_mul<K>(fw[j - 1], B.psi, B.prime);
            iv[j] = mod_mul<K>(iv[j - 1], pinv, B.prime);
        }
        hip_check_(hipMalloc((void**)&B.d_psi, bytes), "hipMalloc");
        hip_check_(hipMalloc((void**)&B.d_psi_inv, bytes), "hipMalloc");
        upload_mont_<K>(B.d_psi, fw, B);
        upload_mont_<K>(B.d_psi_inv, iv, B);
    }
    if (K <= 3) {  // moduli that fit 64-bit arithmetic run on the four-step pass kernels
        const DScalar<K> root = to_mont_d<K>(w, B);
        const ntt_status st = ntt_c_arb_plan_create(&B.plan, B.prime_d.limbs, B.np, K, (uint32_t)log_n, root.limbs);
        if (st != NTT_ERR_UNSUPPORTED) check_(st, "setup_ntt (plan)");
    }
    return B;
}

template <int K>
inline void teardown_ntt(NTTBuffers<K>& B) {
    for (DScalar<K>** p : {&B.d_data, &B.d_tw, &B.d_tfm8, &B.d_hada64, &B.d_psi, &B.d_psi_inv}) {
        if (*p) hip_check_(hipFree(*p), "hipFree");
        *p = nullptr;
    }
    if (B.plan) check_(ntt_c_arb_plan_destroy(B.plan), "teardown_ntt (plan)");
    B.plan = nullptr;
}

// Forward / inverse transform of B's size on d, limb for limb the results of
// run_forward_ct / run_inverse_ct with B's table: the four-step pass kernels
// when B holds a plan (p fits 64-bit arithmetic), else the limb kernels.
template <int K>
inline void run_forward(NTTBuffers<K>& B, DScalar<K>* d, hipStream_t stream = 0) {
    if (B.plan) check_(ntt_c_arb_plan_forward(B.plan, d->limbs, stream), "run_forward");
    else run_forward_ct<K>(d, B.d_tw, (uint32_t)B.n, B.log_n, B.prime_d, B.np, stream);
}

template <int K>
inline void run_inverse(NTTBuffers<K>& B, DScalar<K>* d, hipStream_t stream = 0) {
    if (B.plan) check_(ntt_c_arb_plan_inverse(B.plan, d->limbs, stream), "run_inverse");
    else run_inverse_ct<K>(d, B.d_tw, (uint32_t)B.n, B.log_n, B.prime_d, B.np, B.n_inv_mont, stream);
}

// Negacyclic transforms (no reference counterpart): A_k = sum_j a_j psi^(j(2k+1)),
// as psi-twist then forward NTT; the inverse untwists after the inverse NTT.
template <int K>
inline void run_forward_negacyclic(NTTBuffers<K>& B, DScalar<K>* d, hipStream_t stream = 0) {
    if (!B.d_psi) check_(NTT_ERR_UNSUPPORTED, "run_forward_negacyclic (no 2n-th root in this field)");
    pointwise_mul_mont<K>(d, B.d_psi, (uint32_t)B.n, B.prime_d, B.np, stream);
    run_forward<K>(B, d, stream);
}

template <int K>
inline void run_inverse_negacyclic(NTTBuffers<K>& B, DScalar<K>* d, hipStream_t stream = 0) {
    if (!B.d_psi_inv) check_(NTT_ERR_UNSUPPORTED, "run_inverse_negacyclic (no 2n-th root in this field)");
    run_inverse<K>(B, d, stream);
    pointwise_mul_mont<K>(d, B.d_psi_inv, (uint32_t)B.n, B.prime_d, B.np, stream);
}

// ---------------------------------------------------------------- driver
// Times the forward transform and checks it against host_ntt_naive after
// conversion out of Montgomery form (reference run_for_K :2808-2906).
template <int K>
inline int run_for_K(int log_n, bool do_host_check = true) {
    using FC = FieldConfig<K>;
    if (log_n > FC::MAX_LOG_N) {
        std::cerr << "log_n=" << log_n << " exceeds MAX_LOG_N=" << FC::MAX_LOG_N << " for K=" << K << "\n";
        return 1;
    }
    NTTBuffers<K> B = setup_ntt<K>(log_n);
    std::cout << "K=" << K << " (LIMB_BITS=" << LIMB_BITS << "), N=2^" << log_n << "=" << B.n << "\n";
    auto run = [&] { run_forward<K>(B, B.d_data); };
    for (int it = 0; it < 3; it++) run();
    hip_check_(hipDeviceSynchronize(), "sync");
    hipEvent_t s, e;
    hip_check_(hipEventCreate(&s), "event");
    hip_check_(hipEventCreate(&e), "event");
    const int iters = 20;
    hip_check_(hipEventRecord(s, 0), "record");
    for (int it = 0; it < iters; it++) run();
    hip_check_(hipEventRecord(e, 0), "record");
    hip_check_(hipEventSynchronize(e), "sync");
    float ms = 0;
    hip_check_(hipEventElapsedTime(&ms, s, e), "elapsed");
    std::cout << "  NTT: " << (ms * 1000.0 / iters) << " us/iter\n";
    int mismatches = 0;
    if (do_host_check) {
        upload_mont_<K>(B.d_data, B.data_host, B);
        run();
        hip_check_(hipDeviceSynchronize(), "sync");
        std::vector<DScalar<K>> got((size_t)B.n);
        hip_check_(hipMemcpy(got.data(), B.d_data, got.size() * sizeof(DScalar<K>), hipMemcpyDeviceToHost), "d2h");
        std::vector<BigInt<K>> ref = B.data_host;
        host_ntt_naive<K>(ref, B.tw_host, B.prime);
        for (int i = 0; i < B.n; i++) {
            const BigInt<K> v = mont_mul_big<K>(store_scalar<K>(got[i]), BigInt<K>::one(), B.prime, B.np);
            if (v.cmp(ref[i]) != 0) {
                if (mismatches < 5) std::cout << "  MISMATCH @ i=" << i << "\n";
                mismatches++;
            }
        }
        if (mismatches == 0)
            std::cout << "  CORRECT (matches host reference at all " << B.n << " positions)\n";
        else
            std::cout << "  FAILED (" << mismatches << " mismatches)\n";
    }
    hip_check_(hipEventDestroy(s), "event");
    hip_check_(hipEventDestroy(e), "event");
    teardown_ntt<K>(B);
    return mismatches == 0 ? 0 : 1;
}

}  // namespace ntt_arb
