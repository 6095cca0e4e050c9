// field_host.hpp — host-side modular arithmetic and twiddle-table builders.
//
// The primitive-root rule is the reference's: g is the smallest g >= 2 with
// g^((q-1)/p) != 1 for every p in {2,3,5,7} dividing q-1, and
// omega = g^((q-1)/n) (include/modular_arith.cuh:130-151).  Every table this
// library uploads is a function of that omega (or psi = g^((q-1)/2n), the
// negacyclic root, with psi^2 = omega), so outputs equal ntt_host's.
#pragma once
#include <stdint.h>

#include <vector>

namespace ntth {

typedef unsigned __int128 u128;

inline uint64_t mulmod(uint64_t a, uint64_t b, uint64_t q) { return (uint64_t)((u128)a * b % q); }

// mod_pow, reference include/modular_arith.cuh:110-121.
inline uint64_t powmod(uint64_t b, uint64_t e, uint64_t q) {
    uint64_t r = 1 % q;
    b %= q;
    while (e) {
        if (e & 1) r = mulmod(r, b, q);
        b = mulmod(b, b, q);
        e >>= 1;
    }
    return r;
}

inline uint64_t invmod(uint64_t a, uint64_t q) { return powmod(a, q - 2, q); }

// Generator search of find_primitive_root, reference include/modular_arith.cuh:132-146.
inline uint64_t generator(uint64_t q) {
    static const uint64_t ps[4] = {2, 3, 5, 7};
    for (uint64_t g = 2;; g++) {
        bool ok = true;
        for (uint64_t p : ps)
            if ((q - 1) % p == 0 && powmod(g, (q - 1) / p, q) == 1) { ok = false; break; }
        if (ok) return g;
    }
}

// Deterministic Miller-Rabin for 64-bit n.
inline bool is_prime(uint64_t n) {
    if (n < 2) return false;
    static const uint64_t small[] = {2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37};
    for (uint64_t p : small) {
        if (n == p) return true;
        if (n % p == 0) return false;
    }
    uint64_t d = n - 1;
    int r = 0;
    while (!(d & 1)) { d >>= 1; r++; }
    static const uint64_t bases[] = {2, 325, 9375, 28178, 450775, 9780504, 1795265022};
    for (uint64_t a0 : bases) {
        uint64_t a = a0 % n;
        if (!a) continue;
        uint64_t x = powmod(a, d, n);
        if (x == 1 || x == n - 1) continue;
        bool comp = true;
        for (int k = 1; k < r; k++) {
            x = mulmod(x, x, n);
            if (x == n - 1) { comp = false; break; }
        }
        if (comp) return false;
    }
    return true;
}

inline uint32_t bitrev(uint32_t x, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t i = 0; i < bits; i++) r = (r << 1) | ((x >> i) & 1u);
    return r;
}

// Shoup companion floor(w * 2^bits / q) for bits = 32 or 64.
inline uint64_t shoup(uint64_t w, uint64_t q, int bits) {
    return (uint64_t)(((u128)w << bits) / q);
}

// Merged Cooley-Tukey table for an R-point transform, natural order in,
// bit-reversed order out: Z[2^l + i] = root^((R>>(l+1)) * br_l(i)) * fac^(R>>(l+1)).
// fac = 1 gives the cyclic DFT with root `root` (a primitive R-th root);
// fac = psi_R (psi_R^2 = root) gives the negacyclic transform.
inline std::vector<uint64_t> ct_table(uint32_t log_r, uint64_t root, uint64_t fac, uint64_t q) {
    const uint32_t R = 1u << log_r;
    std::vector<uint64_t> z(R, 1);
    for (uint32_t l = 0; l < log_r; l++) {
        const uint64_t step = R >> (l + 1);
        const uint64_t wl = powmod(root, step, q);  // primitive 2^(l+1)-th root
        const uint64_t fl = powmod(fac, step, q);
        for (uint32_t i = 0; i < (1u << l); i++) z[(1u << l) + i] = mulmod(powmod(wl, bitrev(i, l), q), fl, q);
    }
    z[0] = 1;
    return z;
}

// base^i for i < count, times `mult`.
inline std::vector<uint64_t> powers(uint64_t base, uint64_t count, uint64_t q, uint64_t mult = 1) {
    std::vector<uint64_t> v(count);
    uint64_t t = mult % q;
    for (uint64_t i = 0; i < count; i++) { v[i] = t; t = mulmod(t, base, q); }
    return v;
}

}  // namespace ntth
