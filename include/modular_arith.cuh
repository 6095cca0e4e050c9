// modular_arith.cuh — drop-in for the reference's include/modular_arith.cuh
// (host side): ModularParams, the FHE-friendly prime constants and the host
// helpers mod_pow / mod_inverse / find_primitive_root / create_params /
// compute_twiddle_factors, with the reference's semantics.  The device-side
// helpers of the reference (mod_add/mod_sub at :34-43, the unused Barrett and
// Montgomery32 at :45-106) live inside libntt_mi355x.so's kernels; callers of
// the reference never invoked them outside its own kernels.
#pragma once

#include <cstdint>

#include "ntt_c.h"

namespace ntt {

// ModularParams, reference include/modular_arith.cuh:11-19 (identical layout).
typedef ntt_c_modular_params ModularParams;

// Prime constants, reference include/modular_arith.cuh:23-30.
constexpr uint64_t PRIME_32_1 = 0xFFFFFFFF00000001ULL;
constexpr uint64_t PRIME_30_1 = 1073479681ULL;
constexpr uint64_t PRIME_30_2 = 1073610753ULL;
constexpr uint64_t PRIME_31_1 = 2013265921ULL;
constexpr uint64_t PRIME_32_2 = 4293918721ULL;
constexpr uint64_t DEFAULT_PRIME = 2013265921ULL;

// mod_pow, reference :110-121.
inline uint64_t mod_pow(uint64_t base, uint64_t exp, uint64_t mod) {
    uint64_t result = 1;
    base %= mod;
    while (exp > 0) {
        if (exp & 1) result = (uint64_t)((unsigned __int128)result * base % mod);
        base = (uint64_t)((unsigned __int128)base * base % mod);
        exp >>= 1;
    }
    return result;
}

// mod_inverse, reference :123-125.
inline uint64_t mod_inverse(uint64_t a, uint64_t mod) { return mod_pow(a, mod - 2, mod); }

// find_primitive_root, reference :130-151 (computed by the library).
inline uint64_t find_primitive_root(uint64_t q, uint64_t n) { return ntt_c_find_primitive_root(q, n); }

// create_params, reference :153-170.
inline ModularParams create_params(uint64_t q, uint64_t n) {
    ModularParams p{};
    ntt_c_create_params(q, n, &p);
    return p;
}

// compute_twiddle_factors, reference :173-178.
inline void compute_twiddle_factors(uint64_t* twiddles, uint64_t n, uint64_t omega, uint64_t q) {
    twiddles[0] = 1;
    for (uint64_t i = 1; i < n; i++) twiddles[i] = (uint64_t)((unsigned __int128)twiddles[i - 1] * omega % q);
}

}  // namespace ntt
