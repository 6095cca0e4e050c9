// ntt.cuh — drop-in for the reference's include/ntt.cuh on MI355X.
//
// Same namespace, names, argument meaning and defaults as the reference
// (include/ntt.cuh:8-87), with hipStream_t in place of cudaStream_t.  Every
// function forwards to the C ABI of libntt_mi355x.so (ntt_c.h).  The
// reference's functions return void and exit(EXIT_FAILURE) on CUDA errors
// (src/ntt_basic.cu:13-21); these return void too and, on a failed call, print
// ntt_c_last_error() and abort() unless NTT_NO_ABORT is defined (then the
// status of the last call is in ntt::last_status()).  Use the ntt_c_* API
// directly for status codes without aborts.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "modular_arith.cuh"
#include "ntt_c.h"

namespace ntt {

// NTTConfig, reference include/ntt.cuh:11-24 (same leading fields; `impl` is
// this library's plan/workspace handle).
typedef ntt_c_config NTTConfig;

// NTTVariant, reference include/ntt.cuh:27-32.  All variants compute the same
// transform (ntt_host's); the library picks its kernels itself.
enum class NTTVariant { BASIC = NTT_VARIANT_BASIC, FOUR_STEP = NTT_VARIANT_FOUR_STEP, TLMOP = NTT_VARIANT_TLMOP,
                        FULL_OPTIMIZED = NTT_VARIANT_FULL_OPTIMIZED };

inline ntt_status& last_status() {
    static thread_local ntt_status s = NTT_OK;
    return s;
}

inline void check_(ntt_status s, const char* what) {
    last_status() = s;
    if (s != NTT_OK) {
        std::fprintf(stderr, "ntt: %s failed (%d): %s\n", what, (int)s, ntt_c_last_error());
#ifndef NTT_NO_ABORT
        std::abort();
#endif
    }
}

// reference include/ntt.cuh:35 / src/ntt_basic.cu:217-249
inline NTTConfig* ntt_init(uint64_t n, uint64_t q = DEFAULT_PRIME) {
    NTTConfig* c = nullptr;
    check_(ntt_c_init(n, q, &c), "ntt_init");
    return c;
}
// reference :38 / src/ntt_basic.cu:251-257
inline void ntt_cleanup(NTTConfig* config) { check_(ntt_c_cleanup(config), "ntt_cleanup"); }
// No reference counterpart: reserves the workspace of stream s (once, at
// set-up) for calls of up to max_batch polynomials.  ntt_init reserves the
// default stream's; calls on another stream without a reservation fail with
// NTT_ERR_OOM instead of allocating.
inline void ntt_reserve(const NTTConfig* c, hipStream_t s, uint64_t max_batch = 1) {
    check_(ntt_c_config_reserve(c, s, max_batch), "ntt_reserve");
}

// reference :42-57
inline void ntt_forward_basic(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward(d, c, NTT_VARIANT_BASIC, s), "ntt_forward_basic");
}
inline void ntt_inverse_basic(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_inverse(d, c, NTT_VARIANT_BASIC, s), "ntt_inverse_basic");
}
inline void ntt_forward_four_step(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward(d, c, NTT_VARIANT_FOUR_STEP, s), "ntt_forward_four_step");
}
inline void ntt_inverse_four_step(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_inverse(d, c, NTT_VARIANT_FOUR_STEP, s), "ntt_inverse_four_step");
}
inline void ntt_forward_tlmop(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward(d, c, NTT_VARIANT_TLMOP, s), "ntt_forward_tlmop");
}
inline void ntt_inverse_tlmop(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_inverse(d, c, NTT_VARIANT_TLMOP, s), "ntt_inverse_tlmop");
}
inline void ntt_forward_optimized(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward(d, c, NTT_VARIANT_FULL_OPTIMIZED, s), "ntt_forward_optimized");
}
inline void ntt_inverse_optimized(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_inverse(d, c, NTT_VARIANT_FULL_OPTIMIZED, s), "ntt_inverse_optimized");
}

// Negacyclic transform (named at reference README.md:30, absent from its code).
inline void ntt_forward_negacyclic(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward_negacyclic(d, c, s), "ntt_forward_negacyclic");
}
inline void ntt_inverse_negacyclic(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_inverse_negacyclic(d, c, s), "ntt_inverse_negacyclic");
}
// The product pipeline of reference tests/test_ntt.cu:177-240 (forward
// negacyclic a and b, poly_multiply_ntt, inverse negacyclic) as one call: r =
// a*b mod (x^n+1, q); a and b are left holding their forward transforms.
inline void ntt_negacyclic_multiply(uint64_t* r, uint64_t* a, uint64_t* b, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_negacyclic_multiply(r, a, b, c, s), "ntt_negacyclic_multiply");
}

// reference :63-67 / src/ntt_optimized.cu:922-965
inline void ntt_forward_rowmaj(uint64_t* d, const NTTConfig* c, hipStream_t s = 0) {
    check_(ntt_c_forward_rowmaj(d, c, s), "ntt_forward_rowmaj");
}
inline void rowmaj_to_pretransposed(uint64_t* dst, const uint64_t* src, int N1, int N2, hipStream_t s = 0) {
    check_(ntt_c_rowmaj_to_pretransposed(dst, src, N1, N2, s), "rowmaj_to_pretransposed");
}
inline void rowmaj_from_pretransposed(uint64_t* dst, const uint64_t* src, int N1, int N2, hipStream_t s = 0) {
    check_(ntt_c_rowmaj_from_pretransposed(dst, src, N1, N2, s), "rowmaj_from_pretransposed");
}

// reference :70-75 / src/ntt_basic.cu:337-389
inline void ntt_forward_batch(uint64_t* d, uint64_t batch_size, const NTTConfig* c,
                              NTTVariant v = NTTVariant::BASIC, hipStream_t s = 0) {
    check_(ntt_c_forward_batch(d, batch_size, c, (int)v, s), "ntt_forward_batch");
}
inline void ntt_inverse_batch(uint64_t* d, uint64_t batch_size, const NTTConfig* c,
                              NTTVariant v = NTTVariant::BASIC, hipStream_t s = 0) {
    check_(ntt_c_inverse_batch(d, batch_size, c, (int)v, s), "ntt_inverse_batch");
}

// reference :78-79 / src/ntt_basic.cu:392-409
inline void poly_multiply_ntt(uint64_t* r, const uint64_t* a, const uint64_t* b, const NTTConfig* c,
                              hipStream_t s = 0) {
    check_(ntt_c_poly_multiply(r, a, b, c, s), "poly_multiply_ntt");
}

// reference :82 / src/ntt_basic.cu:259-265
inline void bit_reverse_permute(uint64_t* d, uint64_t n, hipStream_t s = 0) {
    check_(ntt_c_bit_reverse(d, n, s), "bit_reverse_permute");
}

// reference :85 / src/ntt_basic.cu:412-463
inline void ntt_host(std::vector<uint64_t>& data, const ModularParams& params, bool inverse = false) {
    check_(ntt_c_host(data.data(), data.size(), &params, inverse ? 1 : 0), "ntt_host");
}

}  // namespace ntt
