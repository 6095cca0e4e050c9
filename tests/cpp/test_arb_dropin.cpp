// C++ caller of the drop-in include/ntt_arbitrary.cuh, written the way the
// reference drives it (run_for_K, include/ntt_arbitrary.cuh:2808-2906, called
// from src/ntt_mma_arbitrary.cu): forward transforms for K = 2, 3, 11 checked
// against host_ntt_naive, then the additions (inverse round trip, negacyclic
// product against the schoolbook product mod x^n + 1).
//
// `test_arb_dropin --host` runs only the host arithmetic checks (no GPU).
#include <hip/hip_runtime_api.h>

#include <cstdio>
#include <cstring>
#include <vector>

#include "ntt_arbitrary.cuh"

using namespace ntt_arb;

static int g_fail = 0;
#define EXPECT(c)                                                   \
    do {                                                            \
        if (!(c)) {                                                 \
            std::printf("FAILED: %s (line %d)\n", #c, __LINE__);    \
            g_fail++;                                               \
        }                                                           \
    } while (0)

template <int K>
static BigInt<K> pow2_big(int e, const BigInt<K>& p) {
    return mod_pow<K>(BigInt<K>::from_u64(2), (uint64_t)e, p);
}

template <int K>
static void host_checks() {
    using FC = FieldConfig<K>;
    const BigInt<K> p = FC::prime(), one = BigInt<K>::one();
    const uint32_t np = mont_np<K>(p);
    EXPECT((((uint64_t)p.limbs[0] * np + 1) & LIMB_MASK) == 0);
    // the root has order exactly 2^MAX_LOG_N: w^(2^(m-1)) = p - 1
    BigInt<K> w = FC::omega_2pow_max(), pm1;
    big_sub<K>(pm1, p, one);
    for (int i = 0; i < FC::MAX_LOG_N - 1; i++) w = mod_mul<K>(w, w, p);
    EXPECT(w.cmp(pm1) == 0);
    // R2 = R^2 mod p, with R = 2^(24K)
    EXPECT(mont_R2<K>(p).cmp(pow2_big<K>(48 * K, p)) == 0);
    // Montgomery round trip and a product
    const BigInt<K> R2 = mont_R2<K>(p);
    BigInt<K> x = mod_slow<K, K>(from_hex<K>("123456789abcdef0fedcba9876543210"), p);
    BigInt<K> y = mod_pow<K>(BigInt<K>::from_u64(7), 1000003, p);
    const BigInt<K> xm = mont_mul_big<K>(x, R2, p, np), ym = mont_mul_big<K>(y, R2, p, np);
    EXPECT(mont_mul_big<K>(xm, one, p, np).cmp(x) == 0);
    EXPECT(mont_mul_big<K>(mont_mul_big<K>(xm, ym, p, np), one, p, np).cmp(mod_mul<K>(x, y, p)) == 0);
    // inverse
    EXPECT(mod_mul<K>(mod_inv<K>(y, p), y, p).cmp(one) == 0);
    // mod_add / mod_sub wrap
    EXPECT(mod_add<K>(pm1, one, p).is_zero());
    EXPECT(mod_sub<K>(BigInt<K>::zero(), one, p).cmp(pm1) == 0);
    std::printf("host checks K=%d done\n", K);
}

template <int K>
static void roundtrip(int log_n) {
    NTTBuffers<K> B = setup_ntt<K>(log_n, 12345);
    const size_t bytes = (size_t)B.n * sizeof(DScalar<K>);
    std::vector<DScalar<K>> before(B.n), after(B.n);
    hip_check_(hipMemcpy(before.data(), B.d_data, bytes, hipMemcpyDeviceToHost), "d2h");
    run_forward_ct<K>(B.d_data, B.d_tw, (uint32_t)B.n, log_n, B.prime_d, B.np);
    run_inverse_ct<K>(B.d_data, B.d_tw, (uint32_t)B.n, log_n, B.prime_d, B.np, B.n_inv_mont);
    hip_check_(hipMemcpy(after.data(), B.d_data, bytes, hipMemcpyDeviceToHost), "d2h");
    EXPECT(std::memcmp(before.data(), after.data(), bytes) == 0);
    std::printf("roundtrip K=%d log_n=%d done\n", K, log_n);
    teardown_ntt<K>(B);
}

// run_forward / run_inverse (the four-step pass kernels when setup_ntt made a
// plan: K <= 3 here) against the limb kernels on the same data.
template <int K>
static void plan_vs_limb_kernels(int log_n, bool expect_plan) {
    NTTBuffers<K> B = setup_ntt<K>(log_n, 4242);
    EXPECT((B.plan != nullptr) == expect_plan);
    const size_t bytes = (size_t)B.n * sizeof(DScalar<K>);
    DScalar<K>* d2 = nullptr;
    hip_check_(hipMalloc((void**)&d2, bytes), "hipMalloc");
    hip_check_(hipMemcpy(d2, B.d_data, bytes, hipMemcpyDeviceToDevice), "d2d");
    std::vector<DScalar<K>> x(B.n), a(B.n), b(B.n);
    hip_check_(hipMemcpy(x.data(), B.d_data, bytes, hipMemcpyDeviceToHost), "d2h");
    run_forward<K>(B, B.d_data);
    run_forward_ct<K>(d2, B.d_tw, (uint32_t)B.n, log_n, B.prime_d, B.np);
    hip_check_(hipMemcpy(a.data(), B.d_data, bytes, hipMemcpyDeviceToHost), "d2h");
    hip_check_(hipMemcpy(b.data(), d2, bytes, hipMemcpyDeviceToHost), "d2h");
    EXPECT(std::memcmp(a.data(), b.data(), bytes) == 0);
    run_inverse<K>(B, B.d_data);
    hip_check_(hipMemcpy(a.data(), B.d_data, bytes, hipMemcpyDeviceToHost), "d2h");
    EXPECT(std::memcmp(a.data(), x.data(), bytes) == 0);
    std::printf("run_forward/run_inverse K=%d log_n=%d (%s) done\n", K, log_n, B.plan ? "plan" : "limb kernels");
    hip_check_(hipFree(d2), "hipFree");
    teardown_ntt<K>(B);
}

template <int K>
static void negacyclic_product(int log_n) {
    NTTBuffers<K> B = setup_ntt<K>(log_n, 777);
    const int n = B.n;
    const BigInt<K> p = B.prime;
    // second operand: the input reversed
    std::vector<BigInt<K>> a = B.data_host, b(a.rbegin(), a.rend());
    DScalar<K>* d_b = nullptr;
    hip_check_(hipMalloc((void**)&d_b, (size_t)n * sizeof(DScalar<K>)), "hipMalloc");
    upload_mont_<K>(d_b, b, B);
    run_forward_negacyclic<K>(B, B.d_data);
    run_forward_negacyclic<K>(B, d_b);
    pointwise_mul_mont<K>(B.d_data, d_b, (uint32_t)n, B.prime_d, B.np);
    run_inverse_negacyclic<K>(B, B.d_data);
    std::vector<DScalar<K>> got(n);
    hip_check_(hipMemcpy(got.data(), B.d_data, (size_t)n * sizeof(DScalar<K>), hipMemcpyDeviceToHost), "d2h");
    std::vector<BigInt<K>> ref(n);
    for (int i = 0; i < n; i++)
        for (int j = 0; j < n; j++) {
            const BigInt<K> t = mod_mul<K>(a[i], b[j], p);
            const int k = (i + j) % n;
            ref[k] = (i + j < n) ? mod_add<K>(ref[k], t, p) : mod_sub<K>(ref[k], t, p);
        }
    int bad = 0;
    for (int i = 0; i < n; i++)
        bad += mont_mul_big<K>(store_scalar<K>(got[i]), BigInt<K>::one(), p, B.np).cmp(ref[i]) != 0;
    EXPECT(bad == 0);
    std::printf("negacyclic product K=%d n=%d: %d mismatches\n", K, n, bad);
    hip_check_(hipFree(d_b), "hipFree");
    teardown_ntt<K>(B);
}

int main(int argc, char** argv) {
    host_checks<2>();
    host_checks<3>();
    host_checks<11>();
    if (argc > 1 && std::strcmp(argv[1], "--host") == 0) {
        std::printf(g_fail ? "FAILED\n" : "ALL PASSED\n");
        return g_fail ? 1 : 0;
    }
    g_fail += run_for_K<2>(12);
    g_fail += run_for_K<3>(12);
    g_fail += run_for_K<3>(16, true);
    g_fail += run_for_K<11>(10);
    roundtrip<3>(18);
    roundtrip<11>(12);
    roundtrip<2>(14);
    plan_vs_limb_kernels<2>(20, true);
    plan_vs_limb_kernels<3>(20, true);
    plan_vs_limb_kernels<3>(5, true);
    plan_vs_limb_kernels<11>(10, false);
    negacyclic_product<3>(6);
    negacyclic_product<11>(5);
    std::printf(g_fail ? "FAILED\n" : "ALL PASSED\n");
    return g_fail ? 1 : 0;
}
