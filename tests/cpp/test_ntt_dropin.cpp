// C++ caller of the drop-in include/ntt.cuh, written the way the reference's
// own suite uses the API (tests/test_ntt.cu:59-174, 242-313): round trips of
// every variant, GPU vs ntt_host, and the negacyclic polynomial product that
// the reference defined but never ran (tests/test_ntt.cu:177-240).  Inputs use
// a fixed splitmix64 stream instead of libstdc++'s uniform_int_distribution.
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <vector>

#include "ntt.cuh"

using namespace ntt;

static std::vector<uint64_t> random_polynomial(uint64_t n, uint64_t q, uint64_t seed) {
    std::vector<uint64_t> v(n);
    uint64_t s = seed;
    const unsigned __int128 lim = ((((unsigned __int128)1) << 64) / q) * q;
    for (uint64_t i = 0; i < n; i++) {
        uint64_t x;
        do {
            uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
            x = z ^ (z >> 31);
        } while ((unsigned __int128)x >= lim);
        v[i] = x % q;
    }
    return v;
}

#define HCHECK(x)                                                              \
    do {                                                                       \
        if ((x) != hipSuccess) { std::printf("HIP error at %d\n", __LINE__); return false; } \
    } while (0)

static void run(uint64_t* d, NTTConfig* c, NTTVariant v, bool inv) {
    switch (v) {
        case NTTVariant::BASIC: inv ? ntt_inverse_basic(d, c) : ntt_forward_basic(d, c); break;
        case NTTVariant::FOUR_STEP: inv ? ntt_inverse_four_step(d, c) : ntt_forward_four_step(d, c); break;
        case NTTVariant::TLMOP: inv ? ntt_inverse_tlmop(d, c) : ntt_forward_tlmop(d, c); break;
        case NTTVariant::FULL_OPTIMIZED: inv ? ntt_inverse_optimized(d, c) : ntt_forward_optimized(d, c); break;
    }
}

static bool roundtrip_and_host(uint64_t n, NTTVariant v, bool vs_host) {
    NTTConfig* c = ntt_init(n, DEFAULT_PRIME);
    auto x = random_polynomial(n, c->q, 42);
    uint64_t* d = nullptr;
    HCHECK(hipMalloc(&d, n * 8));
    HCHECK(hipMemcpy(d, x.data(), n * 8, hipMemcpyHostToDevice));
    run(d, c, v, false);
    HCHECK(hipDeviceSynchronize());
    bool ok = true;
    if (vs_host) {
        auto h = x;
        ntt_host(h, c->params, false);
        std::vector<uint64_t> g(n);
        HCHECK(hipMemcpy(g.data(), d, n * 8, hipMemcpyDeviceToHost));
        ok = ok && g == h;
    }
    run(d, c, v, true);
    std::vector<uint64_t> y(n);
    HCHECK(hipMemcpy(y.data(), d, n * 8, hipMemcpyDeviceToHost));
    ok = ok && y == x;
    HCHECK(hipFree(d));
    ntt_cleanup(c);
    return ok;
}

static bool negacyclic_product(uint64_t n, bool fused) {
    NTTConfig* c = ntt_init(n, DEFAULT_PRIME);
    auto a = random_polynomial(n, c->q, 42), b = random_polynomial(n, c->q, 123);
    std::vector<uint64_t> ref(n, 0);
    for (uint64_t i = 0; i < n; i++)
        for (uint64_t j = 0; j < n; j++) {
            uint64_t p = (uint64_t)((unsigned __int128)a[i] * b[j] % c->q);
            uint64_t k = (i + j) % n;
            ref[k] = (i + j >= n) ? (ref[k] + c->q - p) % c->q : (ref[k] + p) % c->q;
        }
    uint64_t *da, *db, *dr;
    HCHECK(hipMalloc(&da, n * 8));
    HCHECK(hipMalloc(&db, n * 8));
    HCHECK(hipMalloc(&dr, n * 8));
    HCHECK(hipMemcpy(da, a.data(), n * 8, hipMemcpyHostToDevice));
    HCHECK(hipMemcpy(db, b.data(), n * 8, hipMemcpyHostToDevice));
    if (fused) {
        ntt_negacyclic_multiply(dr, da, db, c);
    } else {
        ntt_forward_negacyclic(da, c);
        ntt_forward_negacyclic(db, c);
        poly_multiply_ntt(dr, da, db, c);
        ntt_inverse_negacyclic(dr, c);
    }
    std::vector<uint64_t> r(n);
    HCHECK(hipMemcpy(r.data(), dr, n * 8, hipMemcpyDeviceToHost));
    HCHECK(hipFree(da));
    HCHECK(hipFree(db));
    HCHECK(hipFree(dr));
    ntt_cleanup(c);
    return r == ref;
}

int main() {
    int passed = 0, total = 0;
    auto rec = [&](bool ok, const char* what, uint64_t n) {
        total++;
        passed += ok;
        std::printf("%-34s N=%-8llu %s\n", what, (unsigned long long)n, ok ? "PASSED" : "FAILED");
    };
    for (uint64_t n : {64ull, 256ull, 1024ull, 4096ull}) rec(roundtrip_and_host(n, NTTVariant::BASIC, true), "Basic roundtrip + vs host", n);
    for (uint64_t n : {4096ull, 8192ull, 16384ull}) rec(roundtrip_and_host(n, NTTVariant::FOUR_STEP, true), "FourStep roundtrip + vs host", n);
    for (uint64_t n = 4096; n <= (1ull << 20); n <<= 1) rec(roundtrip_and_host(n, NTTVariant::FULL_OPTIMIZED, n <= 65536), "Optimized roundtrip (+ vs host)", n);
    for (uint64_t n : {64ull, 1024ull}) rec(negacyclic_product(n, false), "Negacyclic product vs schoolbook", n);
    for (uint64_t n : {64ull, 4096ull}) rec(negacyclic_product(n, true), "Fused negacyclic product", n);
    std::printf("Passed: %d / %d\n", passed, total);
    return passed == total ? 0 : 1;
}
