// ntt_lib.hip — plans, launches and the C ABI of libntt_mi355x.so.
//
// A Plan holds the device-resident tables for one (q, log n, cyclic or
// negacyclic, forward or inverse) transform and the pass descriptions that
// use them.  Every public transform — the reference's single-polynomial
// entry points, its batch loop, and this library's descriptor batches — is a
// list of PassJobs fed to one kernel template (ntt_kernels.hpp).
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/ntt_c.h"
#include "field_host.hpp"
#include "ntt_kernels.hpp"
#include "ntt_launch.hpp"

using namespace nttk;
using namespace ntth;

// ------------------------------------------------------------------ errors
static thread_local std::string g_last_error;

static ntt_status fail(ntt_status st, const std::string& msg) {
    g_last_error = msg;
    return st;
}

#define HIP_TRY(expr)                                                                              \
    do {                                                                                           \
        hipError_t e_ = (expr);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(NTT_ERR_DEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));       \
    } while (0)

extern "C" const char* ntt_c_last_error(void) { return g_last_error.c_str(); }

// ------------------------------------------------------------------ device kernels (small)
namespace nttk {

__global__ void pointwise_mul62_kernel(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t q,
                                       uint64_t qneg_inv, uint64_t r2) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) r[i] = mont64(mont64(a[i], b[i], q, qneg_inv), r2, q, qneg_inv);
}

// Goldilocks pointwise product r = a*b mod q (q = 2^64 - 2^32 + 1).
__global__ void pointwise_mulgold_kernel(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) r[i] = LoadMul<ModGold>(0, 0)(a[i], b[i]);
}

__global__ void bitrev_kernel(uint64_t* d, uint64_t n, uint32_t log_n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint64_t r = (uint64_t)bitrev_bits((uint32_t)i, log_n);
        if (i < r) {
            const uint64_t t = d[i];
            d[i] = d[r];
            d[r] = t;
        }
    }
}

// out[b*n1 + a] = in[a*n2 + b] for a < n1, b < n2 (a 32x32 LDS tile per workgroup).
__global__ void transpose_u64_kernel(uint64_t* out, const uint64_t* in, uint32_t n1, uint32_t n2) {
    __shared__ uint64_t t[32][33];
    const uint32_t a0 = blockIdx.y * 32, b0 = blockIdx.x * 32;
    for (uint32_t i = threadIdx.y; i < 32; i += blockDim.y) {
        const uint32_t a = a0 + i, b = b0 + threadIdx.x;
        if (a < n1 && b < n2) t[i][threadIdx.x] = in[(uint64_t)a * n2 + b];
    }
    __syncthreads();
    for (uint32_t i = threadIdx.y; i < 32; i += blockDim.y) {
        const uint32_t b = b0 + i, a = a0 + threadIdx.x;
        if (a < n1 && b < n2) out[(uint64_t)b * n1 + a] = t[threadIdx.x][i];
    }
}

// Coefficient width conversion between the reference's uint64_t arrays and
// the native 32-bit storage, 4 elements per thread and step (16-byte aligned
// bulk, scalar tail).  Narrowing reduces mod q (q < 2^32) by Barrett with
// mu = floor(2^64 / q): x - hi(x*mu)*q lies in [0, 2q).
__global__ void narrow_u64_kernel(uint32_t* dst, const uint64_t* src, uint64_t n, uint64_t q, uint64_t mu) {
    auto red = [&](uint64_t x) -> uint32_t {
        const uint64_t t = x - __umul64hi(x, mu) * q;
        return (uint32_t)(t >= q ? t - q : t);
    };
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n >> 2;
    for (uint64_t i = tid; i < n4; i += stride) {
        const ulonglong2 a = reinterpret_cast<const ulonglong2*>(src)[2 * i];
        const ulonglong2 b = reinterpret_cast<const ulonglong2*>(src)[2 * i + 1];
        reinterpret_cast<uint4*>(dst)[i] = make_uint4(red(a.x), red(a.y), red(b.x), red(b.y));
    }
    for (uint64_t i = 4 * n4 + tid; i < n; i += stride) dst[i] = red(src[i]);
}
__global__ void widen_u32_kernel(uint64_t* dst, const uint32_t* src, uint64_t n) {
    const uint64_t tid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t n4 = n >> 2;
    for (uint64_t i = tid; i < n4; i += stride) {
        const uint4 a = reinterpret_cast<const uint4*>(src)[i];
        reinterpret_cast<ulonglong2*>(dst)[2 * i] = make_ulonglong2(a.x, a.y);
        reinterpret_cast<ulonglong2*>(dst)[2 * i + 1] = make_ulonglong2(a.z, a.w);
    }
    for (uint64_t i = 4 * n4 + tid; i < n; i += stride) dst[i] = src[i];
}

// out[k] = lo[k & mask] * hi[k >> bits] (Shoup on hi): powers omega^k from two tables.
template <class M>
__global__ void powers_kernel(uint64_t* out, uint64_t n, const TwP<typename M::V>* lo, const TwP<typename M::V>* hi,
                              uint32_t bits, uint64_t q) {
    const M mod{(decltype(M::q))q};
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) out[i] = mod.mul(lo[i & ((1ull << bits) - 1)].w, hi[i >> bits]);
}

}  // namespace nttk

// ------------------------------------------------------------------ device limits
// LDS budget of one workgroup (tile plus, when it fits, the CT table).
static const uint32_t kLdsMax = kPassLdsMax;

// ------------------------------------------------------------------ plans
// M32: q < 2^31; M64: q < 2^62 (Shoup); GOLD: q = 2^64 - 2^32 + 1.
enum class Arith { M32, M64, GOLD };
static Arith arith_of(uint64_t q) {
    return q < (1ull << 31) ? Arith::M32 : q == kGoldilocks ? Arith::GOLD : Arith::M64;
}
enum class Store { U32, U64 };

static uint32_t vbytes(Arith a) { return a == Arith::M32 ? 4u : 8u; }
static uint32_t sbytes(Store s) { return s == Store::U32 ? 4u : 8u; }

// A radix split: P passes of 2^log_r[i] points, tile 2^log_e elements.
struct Split {
    uint32_t npass = 1;
    uint32_t log_r[4] = {0, 0, 0, 0};
    uint32_t log_e = 14;
    uint32_t mfma = 0;  // bit i: pass i runs on the INT8 matrix cores (R = 64 or 4096)
    bool operator<(const Split& o) const {
        return std::make_tuple(npass, log_r[0], log_r[1], log_r[2], log_r[3], log_e, mfma) <
               std::make_tuple(o.npass, o.log_r[0], o.log_r[1], o.log_r[2], o.log_r[3], o.log_e, o.mfma);
    }
};

static uint32_t max_log_e(Arith a) { return a == Arith::M32 ? 14u : 13u; }

// Autotuned choices, keyed by (arith, log_n, log2 of the elements in flight);
// filled by ntt_c_autotune, consulted before the heuristic.
static std::mutex g_tune_mu;
static std::map<std::tuple<int, uint32_t, uint32_t>, Split> g_tuned;

static uint32_t ilog2_floor(uint64_t v) { return v ? 63u - (uint32_t)__builtin_clzll(v) : 0u; }

// Heuristic split: the fewest passes whose radices leave >= 16 columns per tile
// (64 B row segments), with the tile as large as keeps >= 256 workgroups busy
// for the elements in flight in the launch.
static Split heuristic_split(uint32_t log_n, Arith ar, uint64_t elems_in_flight) {
    Split sp;
    const uint32_t emax = max_log_e(ar);
    const uint32_t lf = ilog2_floor(std::max<uint64_t>(elems_in_flight, 1ull << log_n));
    // largest tile that still leaves >= 256 workgroups (never below 2^11 elements)
    const uint32_t le_par = std::min(emax, std::max(lf > 8 ? lf - 8 : 0u, 11u));
    if (log_n <= emax - 2) {
        sp.npass = 1;
        sp.log_r[0] = log_n;
        sp.log_e = std::max(log_n, le_par);
        return sp;
    }
    for (uint32_t P = 2; P <= 4; P++) {
        const uint32_t lr_max = (log_n + P - 1) / P;
        if (lr_max + 4 > le_par && P < 4) continue;  // keep >= 16 columns per tile
        sp.npass = P;
        for (uint32_t i = 0; i < P; i++) sp.log_r[i] = log_n / P + (i < log_n % P ? 1u : 0u);
        sp.log_e = std::min(emax, std::max(le_par, lr_max + 4));
        return sp;
    }
    return sp;
}

// The tuned entry for (arith, log_n, elements in flight), else the tuned
// entry for the same transform at the nearest number of elements in flight
// within a factor of 4 (a split that won at a similar load), else, for a load
// beyond every tuned one, the entry of the largest tuned load if that load
// already fills the chip (2^22 elements or more: more tiles than workgroup
// slots, so a larger batch only adds more of the same tiles), else the
// heuristic.  (A plan tuned for a 2^24-element batch can be much slower for
// one small transform, whose few tiles cannot fill the chip.)
static Split choose_split(uint32_t log_n, Arith ar, uint64_t elems_in_flight) {
    {
        std::lock_guard<std::mutex> lk(g_tune_mu);
        const uint32_t le = ilog2_floor(std::max<uint64_t>(elems_in_flight, 1));
        auto it = g_tuned.find(std::make_tuple((int)ar, log_n, le));
        if (it != g_tuned.end()) return it->second;
        const Split* best = nullptr;
        const Split* top = nullptr;
        uint32_t best_d = 3;  // at most 2 apart in log2
        uint32_t top_e = 0;
        for (const auto& kv : g_tuned) {
            if (std::get<0>(kv.first) != (int)ar || std::get<1>(kv.first) != log_n) continue;
            const uint32_t e = std::get<2>(kv.first), d = e > le ? e - le : le - e;
            if (d < best_d) best = &kv.second, best_d = d;
            if (!top || e > top_e) top = &kv.second, top_e = e;
        }
        if (best) return *best;
        if (top && le > top_e && top_e >= 22) return *top;
    }
    return heuristic_split(log_n, ar, elems_in_flight);
}

struct PassSpec {
    uint32_t log_r = 0, log_c = 0, log_l = 0;  // rows, tile columns, product of earlier radices
    bool mfma = false;                         // radix-16 rounds on the matrix cores (mz_off)
    bool cyc = false;                          // CT table without the negacyclic factor (Z[2^l] = 1)
    int z_off = -1;
    int mz_off = -1;                           // matrix block of the radix-16 MFMA rounds (mfma)
    uint32_t post = POST_NONE;
    int lo_off = -1, hi_off = -1, mid_off = -1;
    int row_off = -1;        // row table of the post (see TableBuilder::rows), -1 if none
    int col_off = -1;        // column table of a stepped first-pass post (TableBuilder::cols)
    uint32_t row_shift = 0;  // tile row index u = first column >> row_shift
    uint32_t e_a = 0, e_b = 0, e_d = 0, e_shift = 0, e_mask = 0, lo_bits = 0, mid_bits = 0;
    uint64_t scale_w = 0, scale_wp = 0;
};

struct Plan {
    uint64_t q = 0;
    uint32_t log_n = 0;
    bool nega = false, inv = false;
    Arith arith = Arith::M32;
    Split split;
    int npass = 1;
    PassSpec pass[4];
    void* block = nullptr;  // device tables
    uint64_t bytes = 0;
    ~Plan() {
        if (block) (void)hipFree(block);
    }
};

// Validates (q, n) the way the transforms need: prime q, 2^log_n | q-1
// (2^(log_n+1) for the negacyclic root).
static ntt_status check_modulus(uint64_t q, uint32_t log_n, bool nega) {
    if (q < 3 || (q >= (1ull << 62) && q != kGoldilocks))
        return fail(NTT_ERR_UNSUPPORTED, "modulus must satisfy 3 <= q < 2^62 or be 2^64 - 2^32 + 1");
    if (!is_prime(q)) return fail(NTT_ERR_NO_ROOT, "modulus is not prime");
    const uint32_t need = log_n + (nega ? 1u : 0u);
    if (need >= 63 || ((q - 1) & ((1ull << need) - 1)) != 0)
        return fail(NTT_ERR_NO_ROOT, nega ? "2n does not divide q-1" : "n does not divide q-1");
    return NTT_OK;
}

struct TableBuilder {
    Arith arith;
    uint64_t q;
    std::vector<unsigned char> host;
    int add(const std::vector<uint64_t>& vals) {
        // Stored as Shoup pairs of the arithmetic width, 16-byte aligned.
        size_t off = (host.size() + 15) & ~size_t(15);
        const int bits = arith == Arith::M32 ? 32 : 64;
        const size_t esz = arith == Arith::M32 ? 8 : 16;
        host.resize(off + vals.size() * esz);
        for (size_t i = 0; i < vals.size(); i++) {
            const uint64_t w = vals[i], wp = shoup(w, q, bits);
            if (arith == Arith::M32) {
                uint32_t p[2] = {(uint32_t)w, (uint32_t)wp};
                memcpy(&host[off + i * esz], p, 8);
            } else {
                uint64_t p[2] = {w, wp};
                memcpy(&host[off + i * esz], p, 16);
            }
        }
        return (int)off;
    }
    int add_raw(const std::vector<unsigned char>& raw) {
        size_t off = (host.size() + 255) & ~size_t(255);
        host.resize(off + raw.size());
        memcpy(&host[off], raw.data(), raw.size());
        return (int)off;
    }
    // Two-level table of base^e, e < 2^mbits: low 2^lob entries and high part.
    // base^e * mult for e < 2^mbits as a product of table entries: two tables
    // of <= 2^9 entries up to 18 bits, else three (<= 2^9 each up to 27 bits),
    // small enough to stage in LDS.
    void pow_tables(PassSpec& s, uint64_t base, uint32_t mbits, uint64_t mult) {
        s.e_mask = (uint32_t)((1ull << mbits) - 1);
        if (mbits <= 18) {
            const uint32_t lob = (mbits + 1) / 2;
            s.lo_bits = lob;
            s.mid_bits = 0;
            s.lo_off = add(powers(base, 1ull << lob, q));
            s.mid_off = -1;
            s.hi_off = add(powers(powmod(base, 1ull << lob, q), 1ull << (mbits - lob), q, mult));
            return;
        }
        const uint32_t b0 = (mbits + 2) / 3, b1 = (mbits - b0 + 1) / 2, b2 = mbits - b0 - b1;
        s.lo_bits = b0;
        s.mid_bits = b1;
        s.lo_off = add(powers(base, 1ull << b0, q));
        s.mid_off = add(powers(powmod(base, 1ull << b0, q), 1ull << b1, q));
        s.hi_off = add(powers(powmod(base, 1ull << (b0 + b1), q), 1ull << b2, q, mult));
    }
    // Row table of a post whose exponent is constant across a tile's columns
    // (u = c >> e_shift fixed per tile, or no u term): entry u*R + k is
    // base^e(k, u) * mult, so the store multiplies by one staged entry per
    // row.  base has order dividing 2^mbits, so the exponent needs no mask.
    void rows(PassSpec& s, uint64_t base, uint64_t mult, uint32_t R, uint32_t U, uint32_t row_shift) {
        std::vector<uint64_t> v((size_t)U * R);
        for (uint32_t u = 0; u < U; u++) {
            uint64_t f = mulmod(powmod(base, (uint64_t)s.e_d * u, q), mult, q);
            const uint64_t st = powmod(base, (uint64_t)s.e_a * u + s.e_b, q);
            for (uint32_t k = 0; k < R; k++) {
                v[(size_t)u * R + k] = f;
                f = mulmod(f, st, q);
            }
        }
        s.row_off = add(v);
        s.row_shift = row_shift;
    }
    // Column table of a first-pass post with exponent s(c)*k + e_d*c,
    // s(c) = e_a*c + e_b (the stepped store, store_loop_step): per column
    // base^s, base^2s, base^3s, base^(s*kstep).  Also 2^32 mod q as the pass's
    // scale constant (the Montgomery factor of the stepped product).
    void cols(PassSpec& s, uint64_t base, uint32_t ncols, uint32_t kstep) {
        std::vector<uint64_t> v((size_t)ncols * 4);
        const uint64_t ga = powmod(base, s.e_a, q), gak = powmod(ga, kstep, q);
        uint64_t b = powmod(base, s.e_b, q), bk = powmod(b, kstep, q);
        for (uint32_t c = 0; c < ncols; c++) {
            const uint64_t b2 = mulmod(b, b, q);
            v[(size_t)c * 4 + 0] = b;
            v[(size_t)c * 4 + 1] = b2;
            v[(size_t)c * 4 + 2] = mulmod(b2, b, q);
            v[(size_t)c * 4 + 3] = bk;
            b = mulmod(b, ga, q);
            bk = mulmod(bk, gak, q);
        }
        s.col_off = add(v);
        s.scale_w = (1ull << 32) % q;
        s.scale_wp = shoup(s.scale_w, q, 32);
    }
};

// Row-table size limit (entries): larger tables keep the power-table post.
static const uint64_t kRowTableMax = 1ull << 20;

// Balanced base-256 digits of v (|v| < 2^30): v = sum_i d_i 256^i, d_i in [-128, 127].
static void balanced_digits(int64_t v, int8_t d[4]) {
    for (int i = 0; i < 4; i++) {
        int64_t r = ((v + 128) % 256 + 256) % 256 - 128;
        d[i] = (int8_t)r;
        v = (v - r) / 256;
    }
}

static int64_t centred(uint64_t v, uint64_t q) { return v > q / 2 ? (int64_t)v - (int64_t)q : (int64_t)v; }

// Matrix block of a pass whose radix-16 rounds run on the matrix cores
// (layout in ntt_mf16.hpp).  From the merged CT table Z of the pass, the
// round at level l0 maps unit i0's 16 slots by the 4-level butterfly network
// of ct_unit<4>: x -> U_i0 x.  Every U_i0 factors as F diag(d_i0) with one F
// for all rounds and units (F[p][v] = w16^(v*br4(p)) times a per-column
// factor); d_i0[v] = U_i0[0][v] / F[0][v].  Where the first round is a
// matrix-core round (log_r mod 4 = 0), F is that round's own U, so its d is
// all ones; otherwise F[0][v] = 1.  Both are read off the simulated network
// and the factorisation is checked for every unit, so the MFMA rounds
// reproduce the VALU rounds' outputs exactly.
static bool build_mf16_block(uint32_t log_r, const std::vector<uint64_t>& Z, uint64_t q,
                             std::vector<unsigned char>& blk) {
    const uint32_t m0 = log_r & 3u;
    std::vector<uint64_t> F;  // 16 x 16, F[p*16 + v]
    std::vector<uint64_t> tw;
    for (uint32_t l0 = m0; l0 + 4 <= log_r; l0 += 4) {
        for (uint32_t i0 = 0; i0 < (1u << l0); i0++) {
            uint64_t U[16][16];  // U[p][v]
            for (uint32_t v = 0; v < 16; v++) {
                uint64_t x[16] = {};
                x[v] = 1;
                for (uint32_t lam = 0; lam < 4; lam++) {
                    const uint32_t half = 1u << (3 - lam);
                    for (uint32_t gg = 0; gg < (1u << lam); gg++) {
                        const uint64_t wt = Z[(1u << (l0 + lam)) + (i0 << lam) + gg];
                        for (uint32_t v0 = 0; v0 < half; v0++) {
                            const uint32_t j = gg * 2 * half + v0;
                            const uint64_t a = x[j], bb = mulmod(x[j + half], wt, q);
                            x[j] = (a + bb) % q;
                            x[j + half] = (a + q - bb) % q;
                        }
                    }
                }
                for (uint32_t p = 0; p < 16; p++) U[p][v] = x[p];
            }
            if (F.empty()) {
                // With no VALU round first (m0 = 0), F is round 0's own
                // matrix, so that round's twiddles are all 1 and the kernel
                // skips them (a negacyclic first pass's weights ride in F);
                // otherwise F is normalised to a first row of ones.
                F.resize(256);
                for (uint32_t v = 0; v < 16; v++) {
                    if (U[0][v] == 0) return false;
                    const uint64_t dinv = invmod(U[0][v], q);
                    for (uint32_t p = 0; p < 16; p++) F[p * 16 + v] = m0 == 0 ? U[p][v] : mulmod(U[p][v], dinv, q);
                }
            }
            uint64_t d[16];
            for (uint32_t v = 0; v < 16; v++) {
                d[v] = mulmod(U[0][v], invmod(F[v], q), q);  // U = F diag(d): d = U's first row / F's
                tw.push_back(d[v]);
            }
            for (uint32_t p = 0; p < 16; p++)
                for (uint32_t v = 0; v < 16; v++)
                    if (mulmod(F[p * 16 + v], d[v], q) != U[p][v]) return false;
        }
    }
    if (F.empty()) return false;
    const bool pack = mf_has_pack(log_r);
    blk.assign(pack ? mf_pack_off(log_r) + kMfPackBytes : kMfTwOff + tw.size() * 8, 0);
    uint32_t* hdr = (uint32_t*)(blk.data() + kMfHdrOff);
    // Row p: y = lo + r + L0 with r = hi * 2^16 (mod q) must represent
    // sum_K G[p][K] * byte_K; the digits are byte - 128, so plane 2 starts at
    // T'_p (centred) with L0 + 2^16 T'_p = 128 * rowsum_p (mod q).
    const uint64_t L0 = (1ull << 28) + (1ull << 21) + (q * 67 + 99) / 100;
    const uint64_t inv16 = invmod(65536 % q, q);
    uint32_t qinv = (uint32_t)q;  // q^-1 mod 2^32 by Newton's iteration
    for (int i = 0; i < 5; i++) qinv *= 2u - (uint32_t)q * qinv;
    const int64_t cm = centred(((u128)1 << 48) % q, q);
    hdr[0] = (uint32_t)L0;
    hdr[1] = (uint32_t)(int32_t)cm;
    hdr[2] = (uint32_t)(int32_t)cm * qinv;
    // Fragments and plane-2 initial values of a 16 x 16 matrix A at byte
    // offset off: G[p][(v, b)] = A[p][v] 2^(8b) mod q, K slot 4v + b; lane l
    // holds K = 16 (l >> 4) + i, i < 16, and output rows 4 (l >> 4) .. + 3.
    auto emit = [&](const std::vector<uint64_t>& A, uint32_t off) {
        int8_t* frag = (int8_t*)(blk.data() + off);
        int32_t* h = (int32_t*)(blk.data() + off + kMfFragBytes);
        auto G = [&](uint32_t p, uint32_t k) -> uint64_t { return mulmod(A[p * 16 + (k >> 2)], 1ull << (8 * (k & 3)), q); };
        for (uint32_t lane = 0; lane < 64; lane++)
            for (uint32_t i = 0; i < 16; i++) {
                int8_t d[4];
                balanced_digits(centred(G(lane & 15u, 16 * (lane >> 4) + i), q), d);
                for (int c = 0; c < 4; c++) frag[(c * 64 + lane) * 16 + i] = d[c];
            }
        for (uint32_t p = 0; p < 16; p++) {
            uint64_t rs = 0;
            for (uint32_t k = 0; k < 64; k++) rs = (rs + G(p, k)) % q;
            const uint64_t T = mulmod(128, rs, q);
            const int64_t t = centred(mulmod((T + q - L0 % q) % q, inv16, q), q);
            for (uint32_t lane = 0; lane < 64; lane++)
                if (4 * (lane >> 4) <= p && p < 4 * (lane >> 4) + 4) h[lane * 4 + (p & 3)] = (int32_t)t;
        }
    };
    emit(F, 0);
    if (pack) {
        // The VALU remainder round (m0 levels at level 0, one block) as a
        // block-diagonal matrix: 16 >> m0 units of 2^m0 slots per MFMA
        // column, slot w = unit h * 2^m0 + v (the kernel's address map puts
        // v in the high and h in the low bits of the 4-bit slot field).
        const uint32_t S = 1u << m0;
        std::vector<uint64_t> Um(S * S);
        for (uint32_t v = 0; v < S; v++) {
            uint64_t x[8] = {};
            x[v] = 1;
            for (uint32_t lam = 0; lam < m0; lam++) {
                const uint32_t half = 1u << (m0 - 1 - lam);
                for (uint32_t gg = 0; gg < (1u << lam); gg++) {
                    const uint64_t wt = Z[(1u << lam) + gg];
                    for (uint32_t v0 = 0; v0 < half; v0++) {
                        const uint32_t j = gg * 2 * half + v0;
                        const uint64_t a = x[j], bb = mulmod(x[j + half], wt, q);
                        x[j] = (a + bb) % q;
                        x[j + half] = (a + q - bb) % q;
                    }
                }
            }
            for (uint32_t p = 0; p < S; p++) Um[p * S + v] = x[p];
        }
        std::vector<uint64_t> A(256, 0);
        for (uint32_t p = 0; p < 16; p++)
            for (uint32_t w = 0; w < 16; w++)
                if ((p >> m0) == (w >> m0)) A[p * 16 + w] = Um[(p & (S - 1)) * S + (w & (S - 1))];
        emit(A, mf_pack_off(log_r));
    }
    uint32_t* twp = (uint32_t*)(blk.data() + kMfTwOff);
    for (size_t i = 0; i < tw.size(); i++) {
        twp[2 * i] = (uint32_t)tw[i];
        twp[2 * i + 1] = (uint32_t)shoup(tw[i], q, 32);
    }
    return tw.size() == mf_tw_entries(log_r);
}

// Passes that may run their radix-16 rounds on the matrix cores: 32-bit
// arithmetic, R >= 16 and at least 2^8 elements per tile (one 16-unit group).
static bool mfma_ok(uint32_t log_r, Arith ar) { return ar == Arith::M32 && log_r >= 4; }

// Builds the tables of one transform for a radix split N = R_0 * ... * R_{P-1}.
// Pass i runs R_i-point transforms; L_i = R_0*...*R_{i-1} and M_i = N / L_i.
//   pass 0     : CT table over R_0 with root w^(N/R_0) (times psi^(N/R_0) for the
//                forward negacyclic), then the four-step twiddle
//                cyclic w^(k c) mod N | forward nega psi^(c(2k+1)) | inverse nega psi^(k(2c+1)),
//                with n^-1 folded in for inverses; transposed store.
//   pass 0<i<P-1: cyclic R_i-point, twiddle w_{M_i}^(k u), u = c >> log L_i.
//   pass P-1   : cyclic R_{P-1}-point, in-place store; the inverse negacyclic
//                ends with psi_{2 M_1}^(-(L_{P-1}/R_0) k - (c >> log R_0)).
// With P = 1 the inverse scales by n^-1 (cyclic) or psi^-k n^-1 (negacyclic).
// Checked against ntt_host in emulation of these exact index formulas.
// root: a primitive n-th root of unity to use instead of the generator rule's
// (cyclic transforms only; 0 = the generator rule).
static ntt_status build_plan(uint64_t q, uint32_t log_n, bool nega, bool inv, Arith arith, const Split& sp,
                             std::unique_ptr<Plan>& out, uint64_t root = 0) {
    ntt_status st = check_modulus(q, log_n, nega);
    if (st != NTT_OK) return st;
    if (arith == Arith::M32 && q >= (1ull << 31)) return fail(NTT_ERR_UNSUPPORTED, "32-bit arithmetic needs q < 2^31");
    if (log_n < 1 || log_n > 26) return fail(NTT_ERR_UNSUPPORTED, "log2 n must be in [1, 26]");
    uint32_t sum = 0;
    for (uint32_t i = 0; i < sp.npass; i++) sum += sp.log_r[i];
    if (sp.npass < 1 || sp.npass > 4 || sum != log_n) return fail(NTT_ERR_INVALID_ARG, "radix split does not cover n");
    auto P = std::make_unique<Plan>();
    P->q = q;
    P->log_n = log_n;
    P->nega = nega;
    P->inv = inv;
    P->arith = arith;
    P->split = sp;
    P->npass = (int)sp.npass;
    const uint64_t N = 1ull << log_n;
    if (root && (nega || root >= q || powmod(root, N, q) != 1 || powmod(root, N / 2, q) != q - 1))
        return fail(NTT_ERR_INVALID_ARG, "root is not a primitive n-th root of unity (or the transform is negacyclic)");
    const uint64_t g = generator(q);
    const uint64_t omega = root ? root : powmod(g, (q - 1) / N, q);
    const uint64_t psi = nega ? powmod(g, (q - 1) / (2 * N), q) : 1;
    const uint64_t w = inv ? invmod(omega, q) : omega;
    const uint64_t p = inv ? invmod(psi, q) : psi;
    const uint64_t ninv = invmod(N % q, q);
    const int sbits = arith == Arith::M32 ? 32 : 64;
    TableBuilder tb{arith, q, {}};
    // Post-multiply tables of pass s: the power tables, plus a row table when
    // the exponent is constant across a tile's columns (no u term, or u =
    // c >> e_shift with C <= 2^e_shift) and the table is small.
    // The first pass of a P-pass plan (row-fast store) with 32-bit arithmetic
    // steps its factors down each column instead (column table), when every
    // thread gets whole quads: log_c <= log2(threads) and R >= 4 * threads / 2^log_c.
    auto post = [&](PassSpec& s, uint64_t base, uint32_t mbits, uint64_t mult, bool first) {
        tb.pow_tables(s, base, mbits, mult);
        const uint32_t R = 1u << s.log_r;
        if (first && P->npass >= 2 && arith == Arith::M32 && s.log_c <= kLogNT) {
            const uint32_t kstep = 4u << (kLogNT - s.log_c);
            if (R >= kstep && (N >> s.log_r) <= kRowTableMax) tb.cols(s, base, (uint32_t)(N >> s.log_r), kstep);
            return;
        }
        const bool u_free = s.e_a == 0 && s.e_d == 0;
        if (!u_free && s.e_shift < s.log_c) return;
        const uint64_t U = u_free ? 1 : std::max<uint64_t>(1, (N >> s.log_r) >> s.e_shift);
        if (U * R <= kRowTableMax) tb.rows(s, base, mult, R, (uint32_t)U, u_free ? 31u : s.e_shift);
    };
    uint32_t log_l = 0;
    for (int i = 0; i < P->npass; i++) {
        PassSpec& s = P->pass[i];
        const uint32_t lr = sp.log_r[i];
        s.log_r = lr;
        s.log_l = log_l;
        s.log_c = std::min(sp.log_e > lr ? sp.log_e - lr : 0u, log_n - lr);
        const uint64_t fac = (i == 0 && nega && !inv) ? powmod(p, N >> lr, q) : 1;
        s.cyc = fac == 1;
        // Matrix-core rounds where the split asks for them and mfma_ok allows
        // (make_jobs also needs a 16-unit group, >= 2^8 elements, per tile).
        const std::vector<uint64_t> Z = ct_table(lr, powmod(w, N >> lr, q), fac, q);
        s.z_off = tb.add(Z);
        s.log_c = std::min(s.log_c, max_log_e(arith) > lr ? max_log_e(arith) - lr : 0u);  // tile fits LDS
        s.mfma = ((sp.mfma >> i) & 1u) && mfma_ok(lr, arith);
        if (s.mfma) {
            std::vector<unsigned char> mb;
            if (!build_mf16_block(lr, Z, q, mb)) return fail(NTT_ERR_DEVICE, "internal: radix-16 matrix factorisation failed");
            s.mz_off = tb.add_raw(mb);
        }
        if (i < P->npass - 1) {
            s.post = POST_FOURSTEP;
            if (i == 0) {
                if (!nega) { s.e_a = 1; post(s, w, log_n, inv ? ninv : 1, true); }
                else if (!inv) { s.e_a = 2; s.e_d = 1; post(s, p, log_n + 1, 1, true); }
                else { s.e_a = 2; s.e_b = 1; post(s, p, log_n + 1, ninv, true); }
            } else {
                s.e_a = 1;
                s.e_shift = log_l;
                post(s, powmod(w, 1ull << log_l, q), log_n - log_l, 1, false);
            }
        } else if (inv && nega) {
            s.post = POST_FOURSTEP;
            if (P->npass == 1) {
                s.e_b = 1;
                post(s, p, log_n + 1, ninv, false);
            } else {
                const uint32_t lr0 = sp.log_r[0];
                s.e_b = 1u << (log_l - lr0);
                s.e_d = 1;
                s.e_shift = lr0;
                post(s, powmod(p, 1ull << lr0, q), log_n - lr0 + 1, 1, false);
            }
        } else if (inv && P->npass == 1) {
            s.post = POST_SCALE;
            s.scale_w = ninv;
            s.scale_wp = shoup(ninv, q, sbits);
        }
        log_l += lr;
    }
    P->bytes = tb.host.size();
    hipError_t e = hipMalloc(&P->block, P->bytes);
    if (e != hipSuccess) return fail(NTT_ERR_OOM, std::string("hipMalloc(tables): ") + hipGetErrorString(e));
    e = hipMemcpy(P->block, tb.host.data(), P->bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(NTT_ERR_DEVICE, std::string("hipMemcpy(tables): ") + hipGetErrorString(e));
    out = std::move(P);
    return NTT_OK;
}

// Scratch elements (arithmetic width) one polynomial needs for this plan on
// data stored as st.  With 32-bit arithmetic on uint64 data, plans of three or
// more passes keep every intermediate in two 32-bit regions instead of
// writing 8-byte words back into the data (make_jobs).
static bool two_scratch(const Plan& P, Store st) {
    return P.arith == Arith::M32 && st == Store::U64 && P.npass >= 3;
}
static uint64_t plan_scratch_elems(const Plan& P, Store st) {
    return P.npass >= 2 ? (1ull << P.log_n) << (two_scratch(P, st) ? 1 : 0) : 0;
}

static bool aligned16(const void* p) { return ((uintptr_t)p & 15u) == 0; }

// Entries of the post-multiply tables (lo, mid, hi) of a POST_FOURSTEP job.
static uint32_t post_table_entries(const PassJob& J) {
    return (1u << J.lo_bits) + (J.mid_bits ? 1u << J.mid_bits : 0u) + (J.e_mask >> (J.lo_bits + J.mid_bits)) + 1u;
}

// Load/store order, vector width and LDS staging flags of a job.
static void finish_flags(PassJob& J, Arith ar, Store sin, Store sout, bool mfma) {
    const uint32_t R = 1u << J.log_r, C = 1u << J.log_c;
    uint32_t f = mfma ? (uint32_t)JF_MFMA : 0u;
    const bool lcf = J.s_c == 1;
    if (lcf) f |= JF_LOAD_CFAST;
    const bool scf = J.o_c == 1 && J.o_lbits >= 2;
    if (scf) f |= JF_STORE_CFAST;
    const int64_t ib = (int64_t)sbytes(sin), ob = (int64_t)sbytes(sout);
    bool vl = aligned16(J.in) && (!J.in2 || aligned16(J.in2)) && (J.in_pstride * ib) % 16 == 0 &&
              (J.ip_n <= 1 || (J.ip_stride * ib) % 16 == 0);
    if (lcf) vl = vl && C >= 4 && (J.s_r % 4) == 0 && (J.ncols % 4) == 0;
    else vl = vl && J.s_r == 1 && R >= 4 && (J.s_c % 4) == 0;
    if (vl) f |= JF_VEC4;
    bool vs = aligned16(J.out) && (J.out_pstride * ob) % 16 == 0;
    if (scf) vs = vs && C >= 4 && (J.o_r % 4) == 0 && (J.o_hs % 4) == 0 && (J.ncols % 4) == 0;
    else vs = vs && J.o_r == 1 && R >= 4 && (J.o_hs % 4) == 0 && (J.o_c % 4 == 0 || J.o_lbits == 0);
    if (vs) f |= JF_VEC4_ST;
    // LDS after the tile: the post tables first (their reads are scattered),
    // then the CT table (its reads are mostly broadcast) if it still fits.
    const uint32_t tile = (R << J.log_c) * vbytes(ar);
    uint32_t used = tile;
    if (J.post == POST_FOURSTEP) {
        f |= JF_POST4;
        const uint32_t rb = R * 2 * vbytes(ar), pb = post_table_entries(J) * 2 * vbytes(ar);
        if (J.tw_row && used + rb <= kLdsMax) {
            f |= JF_POST_LDS | JF_POST_ROW;
            used += rb;
        } else if (used + pb <= kLdsMax) {
            f |= JF_POST_LDS;
            used += pb;
        }
    }
    const uint32_t zn = mfma ? mf_tw_entries(J.log_r) : R;  // staged CT table or MFMA round twiddles
    if (used + zn * 2 * vbytes(ar) <= kLdsMax) f |= JF_Z_LDS;
    J.flags = f;
}

static uint32_t job_lds(const PassJob& J, Arith ar) {
    const uint32_t R = 1u << J.log_r;
    uint32_t b = (R << J.log_c) * vbytes(ar);
    if (J.flags & JF_Z_LDS) b += ((J.flags & JF_MFMA) ? mf_tw_entries(J.log_r) : R) * 2 * vbytes(ar);
    if (J.flags & JF_POST_ROW) b += R * 2 * vbytes(ar);
    else if (J.flags & JF_POST_LDS) b += post_table_entries(J) * 2 * vbytes(ar);
    return b;
}

// Launch kind of a job: which kernel instantiation runs it.
static uint32_t job_kind(const PassJob& J) { return J.flags & (JF_MFMA | JF_POST4); }
// Jobs of one batch launch share their round kind (matrix-core or VALU
// rounds: different instantiations), while jobs with and without a four-step
// post-multiply share one launch of the post-multiply instantiation, which
// runs every post kind (store_dispatch picks it per job).
static uint32_t launch_key(const PassJob& J) { return J.flags & JF_MFMA; }

// log2 of every stride of a job.  make_jobs only makes strides that are powers
// of two (N >> lr, 1, N, 2^ll, 2^(lr+ll)) or zero where the stride only
// multiplies a term that is always zero; anything else is a bug in this file.
static uint8_t stride_shift(int64_t s) {
    if (s <= 0) return 0;
    uint8_t k = 0;
    while (((int64_t)1 << k) < s) k++;
    if (((int64_t)1 << k) != s) {
        std::fprintf(stderr, "ntt: internal error: stride %lld is not a power of two\n", (long long)s);
        std::abort();
    }
    return k;
}
static void set_shifts(PassJob& J) {
    J.sh_sr = stride_shift(J.s_r);
    J.sh_sc = stride_shift(J.s_c);
    J.sh_or = stride_shift(J.o_r);
    J.sh_oc = stride_shift(J.o_c);
    J.sh_ohs = stride_shift(J.o_hs);
}

// Fills the PassJobs of `npoly` polynomials of plan P at data (stride n
// elements, or dstride with two or more passes), with workspace at `work`
// (stride n, arithmetic width).  With mul_a, the first pass reads
// mul_a[i]*mul_b[i] mod q (canonical residues, same layout and storage as
// data) instead of data, or with ip_n > 1 the inner product of the ip_n pairs
// mul_a[i + k*ip_stride]*mul_b[i + k*ip_stride].
static void make_jobs(const Plan& P, void* data, Store st, uint64_t npoly, void* work, PassJob out[4],
                      Store sin_out[4][2], const void* mul_a = nullptr, const void* mul_b = nullptr,
                      int64_t dstride = 0, uint32_t ip_n = 1, int64_t ip_stride = 0) {
    const uint32_t log_n = P.log_n;
    const int64_t N = (int64_t)1 << log_n;
    if (dstride == 0) dstride = N;
    const char* blk = (const char*)P.block;
    const Store mid = P.arith == Arith::M32 ? Store::U32 : Store::U64;
    // second 32-bit workspace region (after npoly polynomials of the first)
    void* const s2 = two_scratch(P, st) ? (char*)work + ((npoly << log_n) << 2) : nullptr;
    void* src = data;
    Store src_st = st;
    for (int i = 0; i < P.npass; i++) {
        const PassSpec& s = P.pass[i];
        PassJob& J = out[i];
        memset(&J, 0, sizeof(J));
        J.z = blk + s.z_off;
        J.tw_lo = s.lo_off >= 0 ? blk + s.lo_off : nullptr;
        J.tw_hi = s.hi_off >= 0 ? blk + s.hi_off : nullptr;
        J.tw_mid = s.mid_off >= 0 ? blk + s.mid_off : nullptr;
        J.tw_row = s.row_off >= 0 ? blk + s.row_off : nullptr;
        J.tw_col = s.col_off >= 0 ? blk + s.col_off : nullptr;
        J.mz = s.mz_off >= 0 ? blk + s.mz_off : nullptr;
        J.row_shift = s.row_shift;
        J.q = P.q;
        J.scale_w = s.scale_w;
        J.scale_wp = s.scale_wp;
        J.log_r = s.log_r;
        J.post = s.post;
        J.e_a = s.e_a; J.e_b = s.e_b; J.e_d = s.e_d;
        J.e_shift = s.e_shift; J.e_mask = s.e_mask; J.lo_bits = s.lo_bits; J.mid_bits = s.mid_bits;
        const bool last = i == P.npass - 1;
        void* dst;
        Store dst_st;
        if (last) dst = data, dst_st = st;
        else if (s2) dst = src == work ? s2 : work, dst_st = mid;
        else dst = src == data ? work : data, dst_st = src == data ? mid : st;
        J.in = src;
        J.out = dst;
        if (i == 0 && mul_a) {
            J.in = mul_a;
            J.in2 = mul_b;
            J.ip_n = ip_n;
            J.ip_stride = ip_stride;
            const u128 r = P.arith == Arith::M32 ? ((u128)1 << 32) % P.q : ((u128)1 << 64) % P.q;
            J.mul_r2 = (uint64_t)(r * r % P.q);
        }
        sin_out[i][0] = src_st;
        sin_out[i][1] = dst_st;
        if (P.npass == 1) {
            // columns are polynomials: element (r, p) at p*N + r
            J.s_r = 1; J.s_c = N;
            J.o_r = 1; J.o_c = 0; J.o_hs = N; J.o_lbits = 0;
            const uint32_t lc = std::min<uint32_t>(s.log_c, ilog2_floor(npoly));
            J.log_c = lc;
            J.ncols = (uint32_t)npoly;
            J.tiles_per_poly = (uint32_t)((npoly + (1ull << lc) - 1) >> lc);
            J.npoly = 1;
            J.in_pstride = J.out_pstride = 0;
        } else {
            const uint32_t lr = s.log_r, ll = s.log_l;
            J.log_c = s.log_c;
            J.s_r = N >> lr; J.s_c = 1;
            if (last) { J.o_lbits = log_n - lr; J.o_hs = 0; J.o_c = 1; J.o_r = (int64_t)1 << ll; }
            else { J.o_lbits = ll; J.o_hs = (int64_t)1 << (lr + ll); J.o_c = 1; J.o_r = (int64_t)1 << ll; }
            J.ncols = (uint32_t)(N >> lr);
            J.tiles_per_poly = J.ncols >> J.log_c;
            J.npoly = (uint32_t)npoly;
            J.in_pstride = src == data ? dstride : N;
            J.out_pstride = dst == data ? dstride : N;
        }
        set_shifts(J);
        // a radix-16 MFMA round needs one full group of 16 units: >= 2^8 elements per tile
        finish_flags(J, P.arith, src_st, dst_st, s.mfma && J.log_r + J.log_c >= 8);
        if (s.cyc) J.flags |= JF_CT_CYC;
        if (J.in2) J.flags |= JF_MUL_IN;
        // the stepped store needs staged power tables and row-fast vector stores
        if (J.tw_col && (J.flags & JF_POST_LDS) && !(J.flags & (JF_POST_ROW | JF_STORE_CFAST)) &&
            (J.flags & JF_VEC4_ST))
            J.flags |= JF_POST_STEP;
        src = dst;
        src_st = dst_st;
    }
    // A row-table or stepped post whose output the next pass reads first with
    // a matrix-core round (log_r mod 4 == 0) may leave values in [0, 2q).
    for (int i = 0; i + 1 < P.npass; i++) {
        const PassJob& Nx = out[i + 1];
        if ((out[i].flags & (JF_POST_ROW | JF_POST_STEP)) && (Nx.flags & JF_MFMA) && (Nx.log_r & 3u) == 0 &&
            P.arith == Arith::M32 && sin_out[i][1] == Store::U32)
            out[i].flags |= JF_LAZY_OUT;
    }
}

static uint32_t job_tiles(const PassJob& J) { return J.tiles_per_poly * (J.npoly ? J.npoly : 1); }

// One launch of the generic pass kernel for (arith, input storage, output
// storage); MF: the jobs run matrix-core rounds (32-bit arithmetic only).
template <bool P4, bool MF>
static ntt_status launch_generic(Arith ar, Store sin, Store sout, const PassJob* djobs, uint32_t njobs,
                                 const PassJob& inl, uint32_t nblocks, uint32_t lds, hipStream_t stream) {
    hipError_t e;
    if (ar == Arith::M32 && sin == Store::U32) {
        e = sout == Store::U32
                ? launch_pass_kernel<Mod32, uint32_t, uint32_t, P4, 0, 0, MF>(djobs, njobs, inl, nblocks, lds, stream)
                : launch_pass_kernel<Mod32, uint32_t, uint64_t, P4, 0, 0, MF>(djobs, njobs, inl, nblocks, lds, stream);
    } else if (ar == Arith::M32) {
        e = sout == Store::U32
                ? launch_pass_kernel<Mod32, uint64_t, uint32_t, P4, 0, 0, MF>(djobs, njobs, inl, nblocks, lds, stream)
                : launch_pass_kernel<Mod32, uint64_t, uint64_t, P4, 0, 0, MF>(djobs, njobs, inl, nblocks, lds, stream);
    } else {
        if (MF) return fail(NTT_ERR_DEVICE, "internal: matrix-core rounds need 32-bit arithmetic");
        e = ar == Arith::GOLD ? launch_pass_kernel<ModGold, uint64_t, uint64_t, P4>(djobs, njobs, inl, nblocks, lds, stream)
                              : launch_pass_kernel<Mod64, uint64_t, uint64_t, P4>(djobs, njobs, inl, nblocks, lds, stream);
    }
    HIP_TRY(e);
    return NTT_OK;
}

// Shape-specialised launch for uint64 data (first pass uint64 -> 32-bit with
// post, last pass 32-bit -> uint64), matrix-core rounds; false if none.
template <bool P4>
static bool launch_shape_u64(Store sin, uint32_t shape, const PassJob* djobs, uint32_t njobs, const PassJob& inl,
                             uint32_t nblocks, uint32_t lds, hipStream_t stream, hipError_t* e) {
    switch (shape) {
#define NTT_SHAPE_CASE(LR, LC)                                                                                    \
    case (LR << 8) | LC:                                                                                           \
        if (sin == Store::U64) {                                                                                   \
            if (!P4) return false;                                                                                 \
            *e = launch_pass_kernel<Mod32, uint64_t, uint32_t, true, LR, LC, true>(djobs, njobs, inl, nblocks, lds, \
                                                                                    stream);                       \
        } else {                                                                                                   \
            *e = launch_pass_kernel<Mod32, uint32_t, uint64_t, P4, LR, LC, true>(djobs, njobs, inl, nblocks, lds,   \
                                                                                  stream);                         \
        }                                                                                                          \
        return true;
        NTT_SHAPES_U64(NTT_SHAPE_CASE)
#undef NTT_SHAPE_CASE
        default: return false;
    }
}

// Shape-specialised launch; false if the shape has no instance.
template <bool P4, bool MF>
static bool launch_shape(uint32_t shape, const PassJob* djobs, uint32_t njobs, const PassJob& inl, uint32_t nblocks,
                         uint32_t lds, hipStream_t stream, hipError_t* e) {
    switch (shape) {
#define NTT_SHAPE_CASE(LR, LC)                                                                                   \
    case (LR << 8) | LC:                                                                                          \
        *e = launch_pass_kernel<Mod32, uint32_t, uint32_t, P4, LR, LC, MF>(djobs, njobs, inl, nblocks, lds, stream); \
        return true;
        NTT_SHAPES(NTT_SHAPE_CASE)
#undef NTT_SHAPE_CASE
        default: return false;
    }
}

// Shape value of the product / inner-product instances (pass_shape_of).
constexpr uint32_t kShapeIp = 1u << 16;

static bool launch_shape_ip(uint32_t shape, const PassJob* djobs, uint32_t njobs, const PassJob& inl,
                            uint32_t nblocks, uint32_t lds, hipStream_t stream, hipError_t* e) {
    switch (shape) {
#define NTT_SHAPE_CASE(LR, LC)                                                                                   \
    case (LR << 8) | LC:                                                                                          \
        *e = launch_pass_kernel<Mod32, uint32_t, uint32_t, true, LR, LC, true, true>(djobs, njobs, inl, nblocks,  \
                                                                                    lds, stream);                \
        return true;
        NTT_IP_SHAPES(NTT_SHAPE_CASE)
#undef NTT_SHAPE_CASE
        default: return false;
    }
}

// One launch of the pass kernel for (arith, input storage, output storage,
// kind = JF_MFMA / JF_POST4 bits of its jobs, all alike).
static ntt_status launch_pass(Arith ar, Store sin, Store sout, uint32_t kind, const PassJob* djobs, uint32_t njobs,
                              const PassJob& inl, uint32_t nblocks, uint32_t lds, hipStream_t stream,
                              uint32_t shape = 0) {
    if (nblocks == 0) return NTT_OK;
    const bool p4 = kind & JF_POST4, mf = kind & JF_MFMA;
    if (shape) {
        hipError_t e = hipSuccess;
        bool found;
        if (shape & kShapeIp)
            found = p4 && mf && sin == Store::U32 && sout == Store::U32 &&
                    launch_shape_ip(shape & ~kShapeIp, djobs, njobs, inl, nblocks, lds, stream, &e);
        else if (sin != Store::U32 || sout != Store::U32)
            found = mf && (p4 ? launch_shape_u64<true>(sin, shape, djobs, njobs, inl, nblocks, lds, stream, &e)
                              : launch_shape_u64<false>(sin, shape, djobs, njobs, inl, nblocks, lds, stream, &e));
        else
            found = p4 ? (mf ? launch_shape<true, true>(shape, djobs, njobs, inl, nblocks, lds, stream, &e)
                             : launch_shape<true, false>(shape, djobs, njobs, inl, nblocks, lds, stream, &e))
                       : (mf ? launch_shape<false, true>(shape, djobs, njobs, inl, nblocks, lds, stream, &e)
                             : launch_shape<false, false>(shape, djobs, njobs, inl, nblocks, lds, stream, &e));
        if (!found) return fail(NTT_ERR_DEVICE, "internal: no kernel for this tile shape");
        HIP_TRY(e);
        return NTT_OK;
    }
    if (p4) return mf ? launch_generic<true, true>(ar, sin, sout, djobs, njobs, inl, nblocks, lds, stream)
                      : launch_generic<true, false>(ar, sin, sout, djobs, njobs, inl, nblocks, lds, stream);
    return mf ? launch_generic<false, true>(ar, sin, sout, djobs, njobs, inl, nblocks, lds, stream)
              : launch_generic<false, false>(ar, sin, sout, djobs, njobs, inl, nblocks, lds, stream);
}

// Shape-specialised kernel for job J, as (log_r << 8) | log_c, or 0: vector
// loads and stores, full tiles, post tables in LDS, and a shape that has an
// instance for the job's storage (32-bit in and out, or a matrix-core first
// or last pass on uint64 data; the first may carry the fused product).  A
// 32-bit first pass whose load forms a product or inner product runs the
// IP instances (kShapeIp | key) where its shape has one.
static uint32_t pass_shape_of(const PassJob& J, Arith ar, Store sin, Store sout) {
    if (ar != Arith::M32) return 0;
    const bool mul = J.flags & JF_MUL_IN;
    // uint64 first passes take single products only
    if (mul && sin == Store::U64 && J.ip_n > 1) return 0;
    if (!(J.flags & JF_VEC4) || !(J.flags & JF_VEC4_ST)) return 0;
    if (J.post == POST_FOURSTEP && !(J.flags & JF_POST_LDS)) return 0;
    if (J.ncols % (1u << J.log_c) != 0) return 0;
    const uint32_t key = (J.log_r << 8) | J.log_c;
    if (mul && sin == Store::U32) {
        if (sout != Store::U32 || !(J.flags & JF_MFMA) || !(J.flags & JF_POST4)) return 0;
        switch (key) {
#define NTT_SHAPE_KEY(LR, LC) case (LR << 8) | LC:
            NTT_IP_SHAPES(NTT_SHAPE_KEY)
#undef NTT_SHAPE_KEY
            return kShapeIp | key;
            default: return 0;
        }
    }
    if (sin == Store::U32 && sout == Store::U32) {
        switch (key) {
#define NTT_SHAPE_KEY(LR, LC) case (LR << 8) | LC:
            NTT_SHAPES(NTT_SHAPE_KEY)
#undef NTT_SHAPE_KEY
            return key;
            default: return 0;
        }
    }
    // uint64 data: matrix-core first passes (with post) and last passes
    if (!(J.flags & JF_MFMA)) return 0;
    if (sin == Store::U64 && (sout != Store::U32 || !(J.flags & JF_POST4))) return 0;
    if (sin == Store::U32 && sout != Store::U64) return 0;
    if (sin == Store::U64 && sout == Store::U64) return 0;
    switch (key) {
#define NTT_SHAPE_KEY(LR, LC) case (LR << 8) | LC:
        NTT_SHAPES_U64(NTT_SHAPE_KEY)
#undef NTT_SHAPE_KEY
        return key;
        default: return 0;
    }
}

// The shape-specialised kernel every job of a launch can use, or 0.
static uint32_t group_shape(const PassJob* jobs, size_t n, Arith ar, Store sin, Store sout) {
    const uint32_t sh = n ? pass_shape_of(jobs[0], ar, sin, sout) : 0;
    for (size_t i = 1; i < n && sh; i++)
        if (pass_shape_of(jobs[i], ar, sin, sout) != sh) return 0;
    return sh;
}

// NTT_OPT_PIPELINE (ntt_c_set_option): 0 off, 1 (default) where every
// workgroup slot gets two tiles or more, 2 wherever a pipelined instance
// exists (16 tiles or more).  About 3 % on the 2^24 transform and 4 % on the
// 2^20 batch of 16, measured with alternating runs on one device.
static std::atomic<int64_t> g_opt_pipeline{1};

// CUs of the current device (cached per device).
static uint32_t device_cus() {
    static std::mutex mu;
    static std::map<int, uint32_t> known;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) return it->second;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
    known[dev] = (uint32_t)cus;
    return (uint32_t)cus;
}

// The pipelined kernel for one pass job, when enabled and the job has a
// pipelined instance: 32-bit storage and arithmetic, matrix-core rounds, a
// shape of NTT_PIPE_SHAPES and a tile count divisible by 8.  By default only
// where every workgroup slot of the device gets two tiles or more: with fewer
// tiles the pipelined grid would leave CUs idle that one tile per workgroup
// fills (option value 2 lifts this, down to 16 tiles).  Returns false to fall
// back to one tile per workgroup.
static bool launch_pipe(const PassJob& J, Arith ar, Store sin, Store sout, uint32_t shape, uint32_t lds,
                        hipStream_t stream, ntt_status* st) {
    const int64_t opt = g_opt_pipeline.load();
    if (opt == 0 || ar != Arith::M32 || sin != Store::U32 || sout != Store::U32 || !(J.flags & JF_MFMA) ||
        (J.flags & JF_MUL_IN) || shape == 0 || (shape & kShapeIp))
        return false;
    const uint32_t tiles = job_tiles(J), cus = device_cus();
    if (tiles % 8 != 0 || cus == 0) return false;
    const bool p4 = J.flags & JF_POST4;
    int per_cu = 0;
    switch (shape) {
#define NTT_PIPE_CASE(LR, LC)                                                                                    \
    case (LR << 8) | LC:                                                                                          \
        per_cu = p4 ? pipe_kernel_per_cu<Mod32, uint32_t, uint32_t, true, LR, LC, true>(lds)                     \
                    : pipe_kernel_per_cu<Mod32, uint32_t, uint32_t, false, LR, LC, true>(lds);                   \
        break;
        NTT_PIPE_SHAPES(NTT_PIPE_CASE)
#undef NTT_PIPE_CASE
        default: return false;
    }
    const uint32_t slots = (uint32_t)per_cu * cus;
    if (opt == 1 && tiles < 2 * slots) return false;
    uint32_t nb = std::min<uint32_t>(tiles / 2, slots) & ~7u;
    if (nb < 8) return false;
    hipError_t e = hipSuccess;
    switch (shape) {
#define NTT_PIPE_CASE(LR, LC)                                                                                    \
    case (LR << 8) | LC:                                                                                          \
        e = p4 ? launch_pipe_kernel<Mod32, uint32_t, uint32_t, true, LR, LC, true>(J, nb, lds, stream)           \
               : launch_pipe_kernel<Mod32, uint32_t, uint32_t, false, LR, LC, true>(J, nb, lds, stream);         \
        break;
        NTT_PIPE_SHAPES(NTT_PIPE_CASE)
#undef NTT_PIPE_CASE
        default: return false;
    }
    *st = e == hipSuccess ? NTT_OK : fail(NTT_ERR_DEVICE, std::string("HIP: ") + hipGetErrorString(e));
    return true;
}

extern "C" ntt_status ntt_c_set_option(uint32_t option, int64_t value, int64_t* old) {
    if (option != NTT_OPT_PIPELINE) return fail(NTT_ERR_INVALID_ARG, "set_option: unknown option");
    if (value < 0 || value > 2) return fail(NTT_ERR_INVALID_ARG, "set_option: NTT_OPT_PIPELINE takes 0, 1 or 2");
    const int64_t prev = g_opt_pipeline.exchange(value);
    if (old) *old = prev;
    return NTT_OK;
}

// Runs plan P on npoly polynomials at data (stride n, or dstride); with
// mul_a, on mul_a*mul_b mod q elementwise, the result going to data
// (make_jobs).
static ntt_status run_plan(const Plan& P, void* data, Store st, uint64_t npoly, void* work, hipStream_t stream,
                           const void* mul_a = nullptr, const void* mul_b = nullptr, int64_t dstride = 0) {
    PassJob jobs[4];
    Store io[4][2];
    make_jobs(P, data, st, npoly, work, jobs, io, mul_a, mul_b, dstride);
    for (int i = 0; i < P.npass; i++) {
        const uint32_t shape = pass_shape_of(jobs[i], P.arith, io[i][0], io[i][1]);
        const uint32_t lds = job_lds(jobs[i], P.arith);
        ntt_status s = NTT_OK;
        if (!launch_pipe(jobs[i], P.arith, io[i][0], io[i][1], shape, lds, stream, &s))
            s = launch_pass(P.arith, io[i][0], io[i][1], job_kind(jobs[i]), nullptr, 0, jobs[i], job_tiles(jobs[i]),
                            lds, stream, shape);
        if (s != NTT_OK) return s;
    }
    return NTT_OK;
}

// ------------------------------------------------------------------ config (include/ntt.cuh)
// Device workspace of one stream: allocated by a set-up call (ntt_c_init for
// the NULL stream, ntt_c_config_reserve / ntt_c_context_reserve for others)
// and never resized by a transform, so the transform path allocates nothing.
struct Workspace {
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

// Replaces the workspace of `stream` by one of at least `bytes` (a set-up
// call: it waits for the stream before freeing a smaller buffer).
static ntt_status reserve_workspace(std::mutex& mu, std::map<hipStream_t, Workspace>& work, hipStream_t stream,
                                    uint64_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    Workspace& w = work[stream];
    if (w.bytes >= bytes) return NTT_OK;
    if (w.ptr) {
        HIP_TRY(hipStreamSynchronize(stream));
        HIP_TRY(hipFree(w.ptr));
        w.ptr = nullptr;
        w.bytes = 0;
    }
    if (hipMalloc(&w.ptr, bytes) != hipSuccess) return fail(NTT_ERR_OOM, "workspace allocation failed");
    w.bytes = bytes;
    return NTT_OK;
}

// The workspace of `stream`; NTT_ERR_OOM if it holds fewer than `need` bytes
// (nothing was reserved for the stream, or less than this call needs).
static ntt_status find_workspace(std::mutex& mu, std::map<hipStream_t, Workspace>& work, hipStream_t stream,
                                 uint64_t need, Workspace* out) {
    *out = Workspace{};
    if (need == 0) return NTT_OK;
    std::lock_guard<std::mutex> lk(mu);
    auto it = work.find(stream);
    if (it == work.end() || it->second.bytes < need)
        return fail(NTT_ERR_OOM, "no workspace of " + std::to_string(need) +
                                     " bytes reserved for this stream (reserve it once at set-up)");
    *out = it->second;
    return NTT_OK;
}

static void free_workspaces(std::map<hipStream_t, Workspace>& work) {
    for (auto& kv : work)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    work.clear();
}

struct ConfigImpl {
    std::unique_ptr<Plan> plans[4];  // [nega*2 + inv]
    // Negacyclic forward plan of ntt_c_negacyclic_multiply where plans[2] is a
    // one-pass plan (a whole polynomial per workgroup): two passes batch the
    // two operands into one launch per pass.  Null: the product uses plans[2].
    std::unique_ptr<Plan> prod_fwd;
    ntt_status plan_status[4] = {NTT_OK, NTT_OK, NTT_OK, NTT_OK};
    std::string plan_error[4];
    Arith arith = Arith::M32;
    uint64_t ws_poly = 0;  // workspace bytes per polynomial of the largest plan
    uint64_t ws_call = 0;  // workspace bytes of the largest single-polynomial call
    std::mutex mu;
    std::map<hipStream_t, Workspace> work;  // per stream
    ~ConfigImpl() { free_workspaces(work); }
};

extern "C" uint64_t ntt_c_find_primitive_root(uint64_t q, uint64_t n) {
    if (q < 3 || n == 0) return 0;
    return powmod(generator(q), (q - 1) / n, q);
}

extern "C" ntt_status ntt_c_create_params(uint64_t q, uint64_t n, ntt_c_modular_params* p) {
    if (!p || q < 3 || n == 0 || (n & (n - 1))) return fail(NTT_ERR_INVALID_ARG, "create_params: bad q or n");
    // create_params, reference include/modular_arith.cuh:153-170
    p->q = q;
    p->q_twice = 2 * q;
    int k = 64;
    while ((1ull << (k - 64)) < q && k < 96) k++;
    p->barrett_k = (uint64_t)k;
    p->barrett_m = (uint64_t)(((u128)1 << k) / q);
    p->omega = ntt_c_find_primitive_root(q, n);
    p->omega_inv = invmod(p->omega, q);
    p->n_inv = invmod(n % q, q);
    return NTT_OK;
}

extern "C" ntt_status ntt_c_cleanup(ntt_c_config* cfg);
static ntt_status static_scratch(void** out);

// The config's tables and the NULL stream's workspace are made here; no
// transform call allocates.  On failure everything built so far is freed.
static ntt_status init_config(ntt_c_config* C, ConfigImpl* I) {
    const uint64_t n = C->n, q = C->q;
    const uint32_t log_n = (uint32_t)C->log_n;
    for (int nega = 0; nega < 2; nega++)
        for (int inv = 0; inv < 2; inv++) {
            const int k = nega * 2 + inv;
            I->plan_status[k] = build_plan(q, log_n, nega, inv, I->arith, choose_split(log_n, I->arith, n), I->plans[k]);
            if (I->plan_status[k] != NTT_OK) {
                I->plan_error[k] = g_last_error;
                if (!nega) return I->plan_status[k];  // cyclic must exist
            }
        }
    // 2^10..2^13: [log_n/2, rest] with 2^10-element tiles on the VALU.  At
    // 2^12 a product takes 22.2 us against 28.5 us with the one-pass plan's
    // three single-workgroup launches (BabyBear), 29.2 against 41.6 us with a
    // 62-bit prime.  If it cannot be built, the product keeps plans[2].
    if (I->plans[2] && I->plans[2]->npass == 1 && log_n >= 10 && log_n <= 13) {
        Split sp;
        sp.npass = 2;
        sp.log_r[0] = log_n / 2;
        sp.log_r[1] = log_n - log_n / 2;
        sp.log_e = 10;
        sp.mfma = 0;
        if (build_plan(q, log_n, true, false, I->arith, sp, I->prod_fwd) != NTT_OK) I->prod_fwd.reset();
    }
    // Workspace: per polynomial, the largest plan's; per call, also the
    // product's two-operand forward batch and the rowmaj layout copy.
    for (auto& P : I->plans)
        if (P) I->ws_poly = std::max(I->ws_poly, plan_scratch_elems(*P, Store::U64) * vbytes(P->arith));
    I->ws_call = I->ws_poly;
    if (I->prod_fwd)
        I->ws_call = std::max(I->ws_call, 2 * plan_scratch_elems(*I->prod_fwd, Store::U64) * vbytes(I->arith));
    I->ws_call = std::max(I->ws_call, n * sizeof(uint64_t) +
                                          plan_scratch_elems(*I->plans[0], Store::U64) * vbytes(I->arith));
    ntt_status st = reserve_workspace(I->mu, I->work, nullptr, I->ws_call);
    if (st != NTT_OK) return st;
    // d_twiddles[k] = omega^k, d_twiddles_inv[k] = omega^-k (reference src/ntt_basic.cu:233-246),
    // generated on the device from two sqrt(n)-entry tables.
    HIP_TRY(hipMalloc(&C->d_twiddles, n * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&C->d_twiddles_inv, n * sizeof(uint64_t)));
    for (int inv = 0; inv < 2; inv++) {
        const uint64_t w = inv ? C->params.omega_inv : C->params.omega;
        const uint32_t bits = (log_n + 1) / 2;
        TableBuilder tb{I->arith, q, {}};
        const int lo = tb.add(powers(w, 1ull << bits, q));
        const int hi = tb.add(powers(powmod(w, 1ull << bits, q), (n >> bits) + 1, q));
        std::unique_ptr<void, void (*)(void*)> dt(nullptr, [](void* p) { if (p) (void)hipFree(p); });
        void* raw = nullptr;
        HIP_TRY(hipMalloc(&raw, tb.host.size()));
        dt.reset(raw);
        HIP_TRY(hipMemcpy(raw, tb.host.data(), tb.host.size(), hipMemcpyHostToDevice));
        uint64_t* dst = inv ? C->d_twiddles_inv : C->d_twiddles;
        const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 4096);
        if (I->arith == Arith::M32)
            hipLaunchKernelGGL(powers_kernel<Mod32>, dim3(blocks), dim3(256), 0, 0, dst, n,
                               (const TwP<uint32_t>*)((char*)raw + lo), (const TwP<uint32_t>*)((char*)raw + hi), bits, q);
        else if (I->arith == Arith::GOLD)
            hipLaunchKernelGGL(powers_kernel<ModGold>, dim3(blocks), dim3(256), 0, 0, dst, n,
                               (const TwP<uint64_t>*)((char*)raw + lo), (const TwP<uint64_t>*)((char*)raw + hi), bits, q);
        else
            hipLaunchKernelGGL(powers_kernel<Mod64>, dim3(blocks), dim3(256), 0, 0, dst, n,
                               (const TwP<uint64_t>*)((char*)raw + lo), (const TwP<uint64_t>*)((char*)raw + hi), bits, q);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipDeviceSynchronize());
    }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_init(uint64_t n, uint64_t q, ntt_c_config** out) {
    if (!out) return fail(NTT_ERR_INVALID_ARG, "ntt_init: null output");
    *out = nullptr;
    if (n < 2 || (n & (n - 1))) return fail(NTT_ERR_INVALID_ARG, "ntt_init: n must be a power of two >= 2");
    const uint32_t log_n = (uint32_t)__builtin_ctzll(n);
    ntt_status st = check_modulus(q, log_n, false);
    if (st != NTT_OK) return st;
    if (log_n > 26) return fail(NTT_ERR_UNSUPPORTED, "ntt_init: n > 2^26");
    ntt_c_config* C = new ntt_c_config();
    memset(C, 0, sizeof(ntt_c_config));
    C->n = n;
    C->log_n = log_n;
    C->q = q;
    C->mu = (uint64_t)(((u128)1 << 64) / q);
    ntt_c_create_params(q, n, &C->params);
    ConfigImpl* I = new ConfigImpl();
    I->arith = arith_of(q);
    C->impl = I;
    st = init_config(C, I);
    if (st == NTT_OK) {
        void* p = nullptr;
        st = static_scratch(&p);  // the current device's copy of `scratch`, for the no-allocation entry point
    }
    if (st != NTT_OK) {
        const std::string msg = g_last_error;
        (void)ntt_c_cleanup(C);
        return fail(st, msg);
    }
    *out = C;
    return NTT_OK;
}

extern "C" ntt_status ntt_c_cleanup(ntt_c_config* cfg) {
    if (!cfg) return NTT_OK;
    (void)hipDeviceSynchronize();
    if (cfg->d_twiddles) (void)hipFree(cfg->d_twiddles);
    if (cfg->d_twiddles_inv) (void)hipFree(cfg->d_twiddles_inv);
    delete (ConfigImpl*)cfg->impl;
    delete cfg;
    return NTT_OK;
}

// ------------------------------------------------------------------ static scratch (no allocation at all)
// Workspace of ntt_c_config_transform_noalloc: a file-scope device array, so
// that entry point needs no hipMalloc, not even at set-up.  Each GPU the
// library runs on holds its own copy until the process ends.  In host code
// `scratch` names a host-side stand-in, not device memory: the device address
// of each GPU's copy comes from hipGetSymbolAddress, once per device, and
// only that address is ever passed to a kernel.
constexpr uint64_t kStaticScratchBytes = 32ull << 20;
static __device__ __align__(256) uint32_t scratch[kStaticScratchBytes / 4];

constexpr int kMaxDevices = 64;
static std::mutex g_static_mu;
static void* g_static_ptr[kMaxDevices];  // device address of `scratch` on each device, or null

// The current device's address of `scratch` (looked up on the first call
// for the device, with that device current).
static ntt_status static_scratch(void** out) {
    *out = nullptr;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(NTT_ERR_UNSUPPORTED, "static scratch: device index out of range");
    std::lock_guard<std::mutex> lk(g_static_mu);
    if (!g_static_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(NTT_ERR_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_static_ptr[dev] = p;
    }
    *out = g_static_ptr[dev];
    return NTT_OK;
}

extern "C" ntt_status ntt_c_config_transform_noalloc(uint64_t* d, const ntt_c_config* cfg, uint32_t flags,
                                                     void* stream) {
    if (!d || !cfg || !cfg->impl) return fail(NTT_ERR_INVALID_ARG, "transform_noalloc: null data or config");
    if (flags & ~(uint32_t)(NTT_FLAG_INVERSE | NTT_FLAG_NEGACYCLIC))
        return fail(NTT_ERR_INVALID_ARG, "transform_noalloc: flags other than INVERSE | NEGACYCLIC");
    ConfigImpl* I = (ConfigImpl*)cfg->impl;
    const int k = ((flags & NTT_FLAG_NEGACYCLIC) ? 2 : 0) + ((flags & NTT_FLAG_INVERSE) ? 1 : 0);
    if (!I->plans[k]) return fail(I->plan_status[k], I->plan_error[k]);
    const Plan& P = *I->plans[k];
    const uint64_t need = plan_scratch_elems(P, Store::U64) * vbytes(P.arith);
    if (need > kStaticScratchBytes)
        return fail(NTT_ERR_UNSUPPORTED, "transform_noalloc: this transform needs " + std::to_string(need) +
                                             " bytes of scratch, more than the " +
                                             std::to_string(kStaticScratchBytes) + " of the static array");
    void* ptr = nullptr;
    if (need) {
        ntt_status st = static_scratch(&ptr);
        if (st != NTT_OK) return st;
    }
    return run_plan(P, d, Store::U64, 1, ptr, (hipStream_t)stream);
}

extern "C" ntt_status ntt_c_config_reserve(const ntt_c_config* cfg, void* stream, uint64_t max_batch) {
    if (!cfg || !cfg->impl) return fail(NTT_ERR_INVALID_ARG, "config_reserve: null config");
    ConfigImpl* I = (ConfigImpl*)cfg->impl;
    const uint64_t b = std::max<uint64_t>(max_batch, 1);
    if (I->ws_poly && b > (~0ull >> 1) / I->ws_poly) return fail(NTT_ERR_INVALID_ARG, "config_reserve: batch too large");
    return reserve_workspace(I->mu, I->work, (hipStream_t)stream, std::max(I->ws_call, I->ws_poly * b));
}

// Runs a config plan on `batch` polynomials at d.  The batch runs in chunks
// of as many polynomials as the stream's workspace holds (at least one: the
// reserve calls size it for one call of every kind), one after another on
// the stream, so a batch larger than the reservation needs no allocation.
static ntt_status run_config(uint64_t* d, const ntt_c_config* cfg, bool nega, bool inv, uint64_t batch,
                             void* stream, const uint64_t* mul_a = nullptr, const uint64_t* mul_b = nullptr) {
    if (!d || !cfg || !cfg->impl) return fail(NTT_ERR_INVALID_ARG, "null data or config");
    if (batch == 0) return NTT_OK;
    ConfigImpl* I = (ConfigImpl*)cfg->impl;
    const int k = (nega ? 2 : 0) + (inv ? 1 : 0);
    if (!I->plans[k]) return fail(I->plan_status[k], I->plan_error[k]);
    const Plan& P = *I->plans[k];
    const uint64_t per = plan_scratch_elems(P, Store::U64) * vbytes(P.arith);
    Workspace ws;
    ntt_status st = find_workspace(I->mu, I->work, (hipStream_t)stream, per, &ws);
    if (st != NTT_OK) return st;
    const uint64_t chunk = per ? std::max<uint64_t>(1, ws.bytes / per) : batch;
    for (uint64_t b0 = 0; b0 < batch && st == NTT_OK; b0 += chunk)
        st = run_plan(P, d + b0 * cfg->n, Store::U64, std::min(chunk, batch - b0), ws.ptr, (hipStream_t)stream,
                      mul_a, mul_b);
    return st;
}

static bool valid_variant(int v) { return v >= NTT_VARIANT_BASIC && v <= NTT_VARIANT_FULL_OPTIMIZED; }

extern "C" ntt_status ntt_c_forward(uint64_t* d, const ntt_c_config* cfg, int variant, void* stream) {
    if (!valid_variant(variant)) return fail(NTT_ERR_INVALID_ARG, "bad variant");
    return run_config(d, cfg, false, false, 1, stream);
}
extern "C" ntt_status ntt_c_inverse(uint64_t* d, const ntt_c_config* cfg, int variant, void* stream) {
    if (!valid_variant(variant)) return fail(NTT_ERR_INVALID_ARG, "bad variant");
    return run_config(d, cfg, false, true, 1, stream);
}
extern "C" ntt_status ntt_c_forward_negacyclic(uint64_t* d, const ntt_c_config* cfg, void* stream) {
    return run_config(d, cfg, true, false, 1, stream);
}
extern "C" ntt_status ntt_c_inverse_negacyclic(uint64_t* d, const ntt_c_config* cfg, void* stream) {
    return run_config(d, cfg, true, true, 1, stream);
}
extern "C" ntt_status ntt_c_forward_batch(uint64_t* d, uint64_t b, const ntt_c_config* cfg, int variant, void* stream) {
    if (!valid_variant(variant)) return fail(NTT_ERR_INVALID_ARG, "bad variant");
    return run_config(d, cfg, false, false, b, stream);
}
extern "C" ntt_status ntt_c_inverse_batch(uint64_t* d, uint64_t b, const ntt_c_config* cfg, int variant, void* stream) {
    if (!valid_variant(variant)) return fail(NTT_ERR_INVALID_ARG, "bad variant");
    return run_config(d, cfg, false, true, b, stream);
}

static ntt_status pointwise(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t q,
                            hipStream_t stream) {
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
    if (q < (1ull << 31)) {
        hipLaunchKernelGGL(pointwise_mul_kernel<uint64_t>, dim3(blocks), dim3(256), 0, stream, r, a, b, n, q,
                           (uint64_t)(((u128)1 << 64) / q));
    } else if (q == kGoldilocks) {
        hipLaunchKernelGGL(pointwise_mulgold_kernel, dim3(blocks), dim3(256), 0, stream, r, a, b, n);
    } else {
        uint64_t inv = 1;  // q^-1 mod 2^64 by Newton iteration
        for (int i = 0; i < 6; i++) inv *= 2 - q * inv;
        const uint64_t r2 = (uint64_t)(((u128)1 << 64) % q * (((u128)1 << 64) % q) % q);
        hipLaunchKernelGGL(pointwise_mul62_kernel, dim3(blocks), dim3(256), 0, stream, r, a, b, n, q, 0 - inv, r2);
    }
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

extern "C" ntt_status ntt_c_poly_multiply(uint64_t* r, const uint64_t* a, const uint64_t* b, const ntt_c_config* cfg,
                                          void* stream) {
    if (!r || !a || !b || !cfg) return fail(NTT_ERR_INVALID_ARG, "poly_multiply: null argument");
    return pointwise(r, a, b, cfg->n, cfg->q, (hipStream_t)stream);
}

extern "C" ntt_status ntt_c_negacyclic_multiply(uint64_t* r, uint64_t* a, uint64_t* b, const ntt_c_config* cfg,
                                                void* stream) {
    if (!r || !a || !b || !cfg) return fail(NTT_ERR_INVALID_ARG, "negacyclic_multiply: null argument");
    if (!cfg->impl) return fail(NTT_ERR_INVALID_ARG, "negacyclic_multiply: null config");
    ConfigImpl* I = (ConfigImpl*)cfg->impl;
    if (!I->plans[2]) return fail(I->plan_status[2], I->plan_error[2]);
    const Plan& P = I->prod_fwd ? *I->prod_fwd : *I->plans[2];
    const int64_t d = b - a;
    ntt_status st = NTT_OK;
    if (d == 0) {  // a squaring: one forward transform
        st = ntt_c_forward_negacyclic(a, cfg, stream);
    } else if ((uint64_t)(d < 0 ? -d : d) < cfg->n) {
        return fail(NTT_ERR_INVALID_ARG, "negacyclic_multiply: a and b overlap");
    } else if (P.npass >= 2) {
        // both forward transforms in one launch per pass: a batch of two
        // polynomials whose stride is b - a
        Workspace ws;
        st = find_workspace(I->mu, I->work, (hipStream_t)stream, plan_scratch_elems(P, Store::U64) * 2 * vbytes(P.arith),
                            &ws);
        if (st == NTT_OK) st = run_plan(P, a, Store::U64, 2, ws.ptr, (hipStream_t)stream, nullptr, nullptr, d);
    } else {
        st = ntt_c_forward_negacyclic(a, cfg, stream);
        if (st == NTT_OK) st = ntt_c_forward_negacyclic(b, cfg, stream);
    }
    // The pointwise product is read by the inverse transform's first pass
    // (JF_MUL_IN), so a*b is never written out.
    if (st == NTT_OK) st = run_config(r, cfg, true, true, 1, stream, a, b);
    return st;
}

extern "C" ntt_status ntt_c_bit_reverse(uint64_t* d, uint64_t n, void* stream) {
    if (!d || n == 0 || (n & (n - 1)) || n > (1ull << 32)) return fail(NTT_ERR_INVALID_ARG, "bit_reverse: bad n");
    const uint32_t log_n = (uint32_t)__builtin_ctzll(n);
    const unsigned blocks = (unsigned)std::min<uint64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(bitrev_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, d, n, log_n);
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

// Grid of a 4-elements-per-thread elementwise pass over n elements.
static unsigned elementwise_blocks(uint64_t n) { return (unsigned)std::min<uint64_t>((n / 4 + 255) / 256 + 1, 8192); }

extern "C" ntt_status ntt_c_u64_to_u32(uint32_t* dst, const uint64_t* src, uint64_t n, uint64_t q, void* stream) {
    if (!dst || !src || q < 2 || q >= (1ull << 32)) return fail(NTT_ERR_INVALID_ARG, "u64_to_u32: bad arguments");
    if (((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u)) return fail(NTT_ERR_INVALID_ARG, "u64_to_u32: pointers must be 16-byte aligned");
    const uint64_t lo = (uint64_t)(uintptr_t)src, hi = (uint64_t)(uintptr_t)dst;
    if (n && lo < hi + n * 4 && hi < lo + n * 8) return fail(NTT_ERR_INVALID_ARG, "u64_to_u32: dst overlaps src");
    if (n == 0) return NTT_OK;
    hipLaunchKernelGGL(narrow_u64_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, (hipStream_t)stream, dst, src, n, q,
                       (uint64_t)(((u128)1 << 64) / q));
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

extern "C" ntt_status ntt_c_u32_to_u64(uint64_t* dst, const uint32_t* src, uint64_t n, void* stream) {
    if (!dst || !src) return fail(NTT_ERR_INVALID_ARG, "u32_to_u64: null argument");
    if (((uintptr_t)dst & 15u) || ((uintptr_t)src & 15u)) return fail(NTT_ERR_INVALID_ARG, "u32_to_u64: pointers must be 16-byte aligned");
    const uint64_t lo = (uint64_t)(uintptr_t)src, hi = (uint64_t)(uintptr_t)dst;
    if (n && lo < hi + n * 8 && hi < lo + n * 4) return fail(NTT_ERR_INVALID_ARG, "u32_to_u64: dst overlaps src");
    if (n == 0) return NTT_OK;
    hipLaunchKernelGGL(widen_u32_kernel, dim3(elementwise_blocks(n)), dim3(256), 0, (hipStream_t)stream, dst, src, n);
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

static ntt_status transpose(uint64_t* dst, const uint64_t* src, uint32_t n1, uint32_t n2, hipStream_t stream) {
    dim3 grid((n2 + 31) / 32, (n1 + 31) / 32);
    hipLaunchKernelGGL(transpose_u64_kernel, grid, dim3(32, 8), 0, stream, dst, src, n1, n2);
    HIP_TRY(hipGetLastError());
    return NTT_OK;
}

// Pre-transposed layout of the reference (src/ntt_optimized.cu:922-934):
// pre[b*N1 + a] = canon[a*N2 + b].
extern "C" ntt_status ntt_c_rowmaj_to_pretransposed(uint64_t* dst, const uint64_t* src, int n1, int n2, void* stream) {
    if (!dst || !src || n1 <= 0 || n2 <= 0 || dst == src)
        return fail(NTT_ERR_INVALID_ARG, "rowmaj_to_pretransposed: bad arguments (dst must differ from src)");
    return transpose(dst, src, (uint32_t)n1, (uint32_t)n2, (hipStream_t)stream);
}
extern "C" ntt_status ntt_c_rowmaj_from_pretransposed(uint64_t* dst, const uint64_t* src, int n1, int n2,
                                                      void* stream) {
    if (!dst || !src || n1 <= 0 || n2 <= 0 || dst == src)
        return fail(NTT_ERR_INVALID_ARG, "rowmaj_from_pretransposed: bad arguments (dst must differ from src)");
    return transpose(dst, src, (uint32_t)n2, (uint32_t)n1, (hipStream_t)stream);
}

// ntt_forward_rowmaj: pre-transposed in, pre-transposed out.  Implemented as
// layout change -> natural-order transform -> layout change through the
// config's workspace.
extern "C" ntt_status ntt_c_forward_rowmaj(uint64_t* d, const ntt_c_config* cfg, void* stream) {
    if (!d || !cfg || !cfg->impl) return fail(NTT_ERR_INVALID_ARG, "forward_rowmaj: null argument");
    ConfigImpl* I = (ConfigImpl*)cfg->impl;
    const uint32_t l1 = (uint32_t)(cfg->log_n + 1) / 2, l2 = (uint32_t)cfg->log_n / 2;
    const uint32_t n1 = 1u << l1, n2 = 1u << l2;
    const Plan& P = *I->plans[0];
    const uint64_t tbytes = cfg->n * sizeof(uint64_t);
    const uint64_t sbytes = plan_scratch_elems(P, Store::U64) * vbytes(P.arith);
    Workspace ws;
    ntt_status st = find_workspace(I->mu, I->work, (hipStream_t)stream, tbytes + sbytes, &ws);
    if (st != NTT_OK) return st;
    uint64_t* tmp = (uint64_t*)ws.ptr;
    st = transpose(tmp, d, n2, n1, (hipStream_t)stream);  // pre [N2][N1] -> canon
    if (st == NTT_OK) st = run_plan(P, tmp, Store::U64, 1, (char*)ws.ptr + tbytes, (hipStream_t)stream);
    if (st == NTT_OK) st = transpose(d, tmp, n1, n2, (hipStream_t)stream);  // canon -> pre
    return st;
}

// ntt_host, reference src/ntt_basic.cu:412-463 — the public host transform.
extern "C" ntt_status ntt_c_host(uint64_t* data, uint64_t n, const ntt_c_modular_params* prm, int inverse) {
    if (!data || !prm || n == 0 || (n & (n - 1))) return fail(NTT_ERR_INVALID_ARG, "ntt_host: bad arguments");
    const int log_n = __builtin_ctzll(n);
    const uint64_t q = prm->q;
    const uint64_t omega = inverse ? prm->omega_inv : prm->omega;
    for (uint64_t i = 0; i < n; i++) {
        const uint64_t r = log_n ? (uint64_t)bitrev((uint32_t)i, (uint32_t)log_n) : 0;
        if (i < r) std::swap(data[i], data[r]);
    }
    for (int s = 0; s < log_n; s++) {
        const uint64_t half = 1ull << s, full = half << 1;
        const uint64_t wb = powmod(omega, n / full, q);
        for (uint64_t g = 0; g < n; g += full) {
            uint64_t w = 1;
            for (uint64_t p = 0; p < half; p++) {
                const uint64_t u = data[g + p], v = mulmod(data[g + p + half], w, q);
                data[g + p] = (uint64_t)(((u128)u + v) % q);  // exact for q >= 2^63 too
                data[g + p + half] = u >= v ? u - v : u + q - v;
                w = mulmod(w, wb, q);
            }
        }
    }
    if (inverse)
        for (uint64_t i = 0; i < n; i++) data[i] = mulmod(data[i], prm->n_inv, q);
    return NTT_OK;
}

// ------------------------------------------------------------------ context + descriptor batches
struct ntt_c_context {
    std::mutex mu;
    std::map<std::tuple<uint64_t, uint32_t, int, int, int, Split>, std::unique_ptr<Plan>> plans;
    uint64_t table_bytes = 0;
    std::mutex smu;
    std::map<hipStream_t, Workspace> work;  // per stream, from ntt_c_context_reserve
    ~ntt_c_context() { free_workspaces(work); }
};

static Arith arith_for(uint64_t q, uint32_t flags) {
    (void)flags;
    return arith_of(q);
}

// Plan for a descriptor; `elems_in_flight` (elements transformed by the same
// launches) steers the radix split, or `forced` overrides it.
static ntt_status ctx_plan(ntt_c_context* ctx, const ntt_c_desc& d, uint64_t elems_in_flight, const Plan** out,
                           const Split* forced = nullptr) {
    const bool nega = d.flags & NTT_FLAG_NEGACYCLIC, inv = d.flags & NTT_FLAG_INVERSE;
    const Arith ar = arith_for(d.q, d.flags);
    const Split sp = forced ? *forced : choose_split(d.log_n, ar, elems_in_flight);
    auto key = std::make_tuple(d.q, d.log_n, (int)nega, (int)inv, (int)ar, sp);
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = ctx->plans.find(key);
    if (it == ctx->plans.end()) {
        std::unique_ptr<Plan> P;
        ntt_status st = build_plan(d.q, d.log_n, nega, inv, ar, sp, P);
        if (st != NTT_OK) return st;
        ctx->table_bytes += P->bytes;
        it = ctx->plans.emplace(key, std::move(P)).first;
    }
    *out = it->second.get();
    return NTT_OK;
}

static ntt_status check_desc(const ntt_c_desc& d) {
    if (!d.data) return fail(NTT_ERR_INVALID_ARG, "descriptor: null data");
    if (d.flags & ~7u) return fail(NTT_ERR_INVALID_ARG, "descriptor: unknown flags");
    if (!(d.flags & NTT_FLAG_U64) && d.q >= (1ull << 31))
        return fail(NTT_ERR_INVALID_ARG, "descriptor: q >= 2^31 needs NTT_FLAG_U64 storage");
    if (d.log_n < 1 || d.log_n > 26) return fail(NTT_ERR_UNSUPPORTED, "descriptor: log_n outside [1,26]");
    return NTT_OK;
}

extern "C" ntt_status ntt_c_context_create(ntt_c_context** out) {
    if (!out) return fail(NTT_ERR_INVALID_ARG, "null output");
    *out = new ntt_c_context();
    return NTT_OK;
}
extern "C" ntt_status ntt_c_context_destroy(ntt_c_context* ctx) {
    if (ctx) {
        (void)hipDeviceSynchronize();
        delete ctx;
    }
    return NTT_OK;
}
extern "C" uint64_t ntt_c_context_table_bytes(const ntt_c_context* ctx) { return ctx ? ctx->table_bytes : 0; }

struct ntt_c_batch {
    // Launch groups: key (arith, in store, out store), a device job array.
    struct Group {
        Arith ar;
        Store sin, sout;
        uint32_t kind = 0;   // JF_MFMA / JF_POST4 bits shared by the group's jobs
        uint32_t shape = 0;  // shape-specialised kernel shared by all of them (0: generic)
        uint32_t njobs = 0, nblocks = 0, lds = 0;
        PassJob* djobs = nullptr;
        PassJob inl{};
    };
    std::vector<Group> phases[4];  // phase k: pass k of every polynomial
    // The whole batch as one work-queue launch (queue_kernel), when every job
    // runs on one kernel instantiation (32-bit arithmetic and storage, one
    // round kind): device block = jobs, queue bounds, control words.
    struct Queue {
        bool ok = false;
        bool mf = false;
        void* dev = nullptr;
        PassJob* djobs = nullptr;
        uint32_t* tjob = nullptr;  // job of every global tile
        uint32_t* qbeg = nullptr;
        unsigned long long* ctl = nullptr;  // status word, completion counters (queue_kernel)
        uint32_t njobs = 0, nq = 0, nblocks = 0, lds = 0, tslot = 0, tiles = 0;
        uint32_t spin_limit = 1u << 22;
        uint32_t proto = QP_AGENT;      // protocol in use
        uint32_t proto_auto = QP_AGENT; // what the device probe allows
    } queue;
    bool use_queue = false;
    int graph_mode = -1;           // use_queue of the captured graph
    void* work = nullptr;
    uint64_t work_bytes = 0;
    uint64_t workgroups = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        graph_mode = -1;
    }
    ~ntt_c_batch() {
        drop_graph();
        for (auto& ph : phases)
            for (auto& g : ph)
                if (g.djobs) (void)hipFree(g.djobs);
        if (queue.dev) (void)hipFree(queue.dev);
        if (work) (void)hipFree(work);
    }
};

// Work-queue plan of a batch (queue_kernel in ntt_kernels.hpp): polynomials
// are spread over nq queues by longest-processing-time first (one queue per
// XCD, so a polynomial's passes share an L2); within a queue the jobs go
// phase by phase, each job after the one it depends on.  items[i] holds the
// jobs of polynomial i in pass order.  Returns false (the batch keeps its
// per-phase launches) when the jobs need more than one kernel instantiation.
struct QueueItem {
    PassJob job;
    Arith ar;
    Store sin, sout;
};

// Writes the XCC_ID of every workgroup (the queue protocol's probe).
__global__ void xcc_probe_kernel(uint32_t* out) {
    if (threadIdx.x == 0) out[blockIdx.x] = hw_xcc_id();
}

// Whether the XCD-local queue protocol can run on `dev`: a probe launch of
// 64 * nq workgroups must see XCC_ID values below nq on every workgroup and
// every one of the nq XCDs in use.  Cached per device.
static bool xcd_queues_ok(int dev, uint32_t nq) {
    static std::mutex mu;
    static std::map<int, bool> known;
    std::lock_guard<std::mutex> lk(mu);
    auto it = known.find(dev);
    if (it != known.end()) return it->second;
    bool ok = false;
    const uint32_t nb = 64u * nq;
    uint32_t* d = nullptr;
    if (nq == 8 && hipMalloc(&d, nb * sizeof(uint32_t)) == hipSuccess) {
        std::vector<uint32_t> h(nb, 0xFFu);
        hipLaunchKernelGGL(xcc_probe_kernel, dim3(nb), dim3(64), 0, 0, d);
        if (hipDeviceSynchronize() == hipSuccess &&
            hipMemcpy(h.data(), d, nb * sizeof(uint32_t), hipMemcpyDeviceToHost) == hipSuccess) {
            std::vector<bool> seen(nq, false);
            ok = true;
            for (uint32_t b = 0; b < nb; b++) {
                ok = ok && h[b] < nq;
                if (h[b] < nq) seen[h[b]] = true;
            }
            for (uint32_t x = 0; x < nq; x++) ok = ok && seen[x];
        }
        (void)hipFree(d);
    }
    (void)hipGetLastError();
    known[dev] = ok;
    return ok;
}
static ntt_status build_queue(ntt_c_batch::Queue& Q, const std::vector<std::vector<QueueItem>>& polys) {
    Q.ok = false;
    if (polys.empty()) return NTT_OK;
    const QueueItem& f = polys[0][0];
    for (auto& pv : polys)
        for (auto& it : pv)
            if (it.ar != Arith::M32 || it.sin != Store::U32 || it.sout != Store::U32 ||
                launch_key(it.job) != launch_key(f.job) || it.job.npoly > 1)
                return NTT_OK;
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const uint32_t nq = (cus >= 8 && cus % 8 == 0) ? 8u : 1u;
    // longest processing time first: elements moved over all passes
    std::vector<size_t> order(polys.size());
    for (size_t i = 0; i < order.size(); i++) order[i] = i;
    auto cost = [&](size_t i) {
        uint64_t c = 0;
        for (auto& it : polys[i]) c += (uint64_t)job_tiles(it.job) << (it.job.log_r + it.job.log_c);
        return c;
    };
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) { return cost(a) > cost(b); });
    std::vector<uint64_t> load(nq, 0);
    std::vector<std::vector<size_t>> members(nq);
    for (size_t i : order) {
        const size_t x = (size_t)(std::min_element(load.begin(), load.end()) - load.begin());
        members[x].push_back(i);
        load[x] += cost(i);
    }
    std::vector<PassJob> jobs;
    std::vector<uint32_t> qbeg(nq + 1, 0);
    std::vector<int64_t> last(polys.size(), -1);  // index of the poly's previous-pass job
    uint32_t tiles = 0, lds = 0;
    for (uint32_t x = 0; x < nq; x++) {
        qbeg[x] = tiles;
        for (int k = 0; k < 4; k++)
            for (size_t i : members[x]) {
                if ((size_t)k >= polys[i].size()) continue;
                PassJob J = polys[i][k].job;
                J.tile_begin = tiles;
                J.dep = last[i] < 0 ? 0u : (uint32_t)last[i];
                J.dep_tiles = last[i] < 0 ? 0u : job_tiles(jobs[(size_t)last[i]]);
                J.counted = last[i] >= 0 ? 1u : 0u;
                if (last[i] >= 0) jobs[(size_t)last[i]].counted = 1;
                last[i] = (int64_t)jobs.size();
                tiles += job_tiles(J);
                lds = std::max(lds, job_lds(J, Arith::M32));
                jobs.push_back(J);
            }
    }
    qbeg[nq] = tiles;
    if (tiles == 0) return NTT_OK;
    Q.mf = launch_key(f.job) != 0;
    Q.njobs = (uint32_t)jobs.size();
    Q.nq = nq;
    Q.tiles = tiles;
    Q.tslot = (lds + 15u) & ~15u;  // two words: the tile taken and its job
    Q.lds = Q.tslot + 16u;
    if (Q.lds > kPassLdsMax) return NTT_OK;
    const int per_cu = Q.mf ? queue_kernel_per_cu<Mod32, uint32_t, uint32_t, true, true>(Q.lds)
                            : queue_kernel_per_cu<Mod32, uint32_t, uint32_t, true, false>(Q.lds);
    Q.nblocks = std::min<uint32_t>(tiles, (uint32_t)(per_cu * cus));
    Q.proto_auto = xcd_queues_ok(dev, nq) ? QP_XCD : QP_AGENT;
    Q.proto = Q.proto_auto;
    std::vector<uint32_t> tjob(tiles);
    for (size_t j = 0; j < jobs.size(); j++)
        for (uint32_t t = 0; t < job_tiles(jobs[j]); t++) tjob[jobs[j].tile_begin + t] = (uint32_t)j;
    const size_t jb = jobs.size() * sizeof(PassJob), tb = (size_t)tiles * sizeof(uint32_t),
                 qb = (nq + 1) * sizeof(uint32_t), cb = q_done((uint32_t)jobs.size()) * sizeof(unsigned long long);
    const size_t toff = (jb + 255) & ~(size_t)255, qoff = (toff + tb + 255) & ~(size_t)255,
                 coff = (qoff + qb + 255) & ~(size_t)255;
    if (hipMalloc(&Q.dev, coff + cb) != hipSuccess) return fail(NTT_ERR_OOM, "batch_prepare: queue allocation failed");
    Q.djobs = (PassJob*)Q.dev;
    Q.tjob = (uint32_t*)((char*)Q.dev + toff);
    Q.qbeg = (uint32_t*)((char*)Q.dev + qoff);
    Q.ctl = (unsigned long long*)((char*)Q.dev + coff);
    HIP_TRY(hipMemcpy(Q.djobs, jobs.data(), jb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(Q.tjob, tjob.data(), tb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(Q.qbeg, qbeg.data(), qb, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(Q.ctl, 0, cb));
    Q.ok = true;
    return NTT_OK;
}

// Jobs of one batch group that differ only in where their polynomials sit
// run as one job over npoly polynomials when the polynomials are equally
// spaced on both sides (a multiple of the element size, and of 16 bytes on a
// side with vector accesses): a single-job group can then take the pipelined
// kernel, and fewer jobs leave a shorter tile-to-job search.  Multi-pass jobs
// only (s_c == 1): a one-pass job lays its polynomials out as tile columns.
static bool same_but_place(const PassJob& a, const PassJob& b) {
    PassJob x = a, y = b;
    x.in = y.in = nullptr;
    x.out = y.out = nullptr;
    x.tile_begin = y.tile_begin = 0;
    return memcmp(&x, &y, sizeof(PassJob)) == 0;
}

static std::vector<PassJob> merge_jobs(const std::vector<PassJob>& jobs, Store sin, Store sout) {
    std::vector<PassJob> out;
    const int64_t ib = sbytes(sin), ob = sbytes(sout);
    for (size_t i = 0; i < jobs.size();) {
        PassJob J = jobs[i];
        size_t j = i + 1;
        if (J.npoly == 1 && !J.in2 && J.s_c == 1 && j < jobs.size() && same_but_place(J, jobs[j])) {
            const int64_t din = (const char*)jobs[j].in - (const char*)J.in;
            const int64_t dout = (const char*)jobs[j].out - (const char*)J.out;
            const bool ok = din > 0 && dout > 0 && din % ib == 0 && dout % ob == 0 &&
                            (!(J.flags & JF_VEC4) || din % 16 == 0) && (!(J.flags & JF_VEC4_ST) || dout % 16 == 0);
            if (ok) {
                while (j < jobs.size() && same_but_place(J, jobs[j]) &&
                       (const char*)jobs[j].in - (const char*)J.in == (int64_t)(j - i) * din &&
                       (const char*)jobs[j].out - (const char*)J.out == (int64_t)(j - i) * dout)
                    j++;
                J.npoly = (uint32_t)(j - i);
                J.in_pstride = din / ib;
                J.out_pstride = dout / ob;
            } else {
                j = i + 1;
            }
        }
        out.push_back(J);
        i = j;
    }
    uint32_t begin = 0;
    for (PassJob& J : out) {
        J.tile_begin = begin;
        begin += job_tiles(J);
    }
    return out;
}

// An inner-product source of ntt_c_batch_prepare_ip, checked before any
// device work: an inverse transform, dnum <= 64 pairs of non-null slices
// n <= kstride <= 2^40 elements apart, none overlapping the descriptor's data.
static ntt_status check_ip(const ntt_c_desc& d, const ntt_c_ip& ip) {
    if (ip.dnum == 0) return NTT_OK;
    if (ip.dnum > 64 || ip.reserved) return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: need dnum <= 64, reserved = 0");
    if (!(d.flags & NTT_FLAG_INVERSE))
        return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: an inner product feeds an inverse transform only");
    if (!ip.a || !ip.b) return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: null inner-product input");
    const int64_t n = (int64_t)1 << d.log_n;
    if (ip.dnum > 1 && (ip.kstride < n || ip.kstride > ((int64_t)1 << 40)))
        return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: need n <= kstride <= 2^40 elements");
    const uint64_t sb = (d.flags & NTT_FLAG_U64) ? 8 : 4;
    const uint64_t span = ((uint64_t)(ip.dnum - 1) * (uint64_t)(ip.dnum > 1 ? ip.kstride : 0) + (uint64_t)n) * sb;
    const uint64_t o = (uint64_t)(uintptr_t)d.data, ob = (uint64_t)n * sb;
    for (const void* src : {ip.a, ip.b}) {
        const uint64_t s = (uint64_t)(uintptr_t)src;
        if (s < o + ob && o < s + span) return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: data overlaps an input");
    }
    return NTT_OK;
}

static ntt_status batch_prepare(ntt_c_context* ctx, const ntt_c_desc* descs, const ntt_c_ip* ips, size_t count,
                                ntt_c_batch** out) {
    if (!ctx || !out || (count && !descs)) return fail(NTT_ERR_INVALID_ARG, "batch_prepare: null argument");
    *out = nullptr;
    auto B = std::make_unique<ntt_c_batch>();
    struct Item { PassJob job; Arith ar; Store sin, sout; int phase; };
    std::vector<Item> items;
    std::vector<std::vector<QueueItem>> qpolys(count);
    std::vector<const Plan*> plans(count);
    uint64_t work_bytes = 0, elems = 0;
    for (size_t i = 0; i < count; i++) {
        ntt_status st = check_desc(descs[i]);
        if (st == NTT_OK && ips) st = check_ip(descs[i], ips[i]);
        if (st != NTT_OK) return st;
        elems += 1ull << descs[i].log_n;
    }
    bool two_phase = false;
    for (size_t i = 0; i < count; i++) {
        ntt_status st = ctx_plan(ctx, descs[i], elems, &plans[i]);
        if (st != NTT_OK) return st;
        two_phase |= plans[i]->npass >= 2;
    }
    // A batch of at least 2^24 elements that already runs a second phase
    // gives its one-pass 32-bit-arithmetic polynomials of 2^13 and 2^14 a
    // two-pass plan ([7, 6] / [7, 7], tiles of 2^13, matrix-core rounds): the
    // second launch is paid anyway, and the first phase's largest tile, which
    // sets the LDS of every workgroup of that launch, drops from 2^14 to 2^13
    // words (three workgroups per CU instead of two).  On 512 polynomials of
    // 2^10..2^18 the batch takes 153 us against 176 us with the one-pass
    // plans; a batch of only such polynomials keeps the one-pass plan
    // (2048 x 2^13: 61 us against 72 us for two passes).
    if (two_phase && elems >= (1ull << 24))
        for (size_t i = 0; i < count; i++) {
            const uint32_t ln = descs[i].log_n;
            if (plans[i]->npass != 1 || plans[i]->arith != Arith::M32 || ln < 13 || ln > 14) continue;
            Split sp;
            sp.npass = 2;
            sp.log_r[0] = 7;
            sp.log_r[1] = ln - 7;
            sp.log_e = 13;
            sp.mfma = 3;
            ntt_status st = ctx_plan(ctx, descs[i], elems, &plans[i], &sp);
            if (st != NTT_OK) return st;
        }
    for (size_t i = 0; i < count; i++) {
        const Store sd = (descs[i].flags & NTT_FLAG_U64) ? Store::U64 : Store::U32;
        work_bytes += ((plan_scratch_elems(*plans[i], sd) * vbytes(plans[i]->arith)) + 255) & ~255ull;
    }
    if (work_bytes) {
        hipError_t e = hipMalloc(&B->work, work_bytes);
        if (e != hipSuccess) return fail(NTT_ERR_OOM, "batch_prepare: workspace allocation failed");
        B->work_bytes = work_bytes;
    }
    uint64_t soff = 0;
    for (size_t i = 0; i < count; i++) {
        const Plan& P = *plans[i];
        const Store st = (descs[i].flags & NTT_FLAG_U64) ? Store::U64 : Store::U32;
        void* scr = nullptr;
        if (P.npass >= 2) {
            scr = (char*)B->work + soff;
            soff += ((plan_scratch_elems(P, st) * vbytes(P.arith)) + 255) & ~255ull;
        }
        PassJob jobs[4];
        Store io[4][2];
        const ntt_c_ip* ip = ips && ips[i].dnum ? &ips[i] : nullptr;
        make_jobs(P, descs[i].data, st, 1, scr, jobs, io, ip ? ip->a : nullptr, ip ? ip->b : nullptr, 0,
                  ip ? ip->dnum : 1u, ip ? ip->kstride : 0);
        for (int k = 0; k < P.npass; k++) {
            items.push_back({jobs[k], P.arith, io[k][0], io[k][1], k});
            qpolys[i].push_back({jobs[k], P.arith, io[k][0], io[k][1]});
        }
    }
    {
        // The work queue is built where the mix allows it but not used by
        // default: the per-phase launches measured faster on every batch
        // workload (ntt_c_batch_set_queue switches).
        ntt_status st = build_queue(B->queue, qpolys);
        if (st != NTT_OK) return st;
        B->use_queue = false;
    }
    // Group by (phase, arith, storage); within a group order by LDS size so
    // that big tiles start first.
    std::stable_sort(items.begin(), items.end(), [](const Item& a, const Item& b) {
        return std::make_tuple(a.phase, (int)a.ar, (int)a.sin, (int)a.sout, launch_key(a.job),
                               -(int)(a.job.log_r + a.job.log_c)) <
               std::make_tuple(b.phase, (int)b.ar, (int)b.sin, (int)b.sout, launch_key(b.job),
                               -(int)(b.job.log_r + b.job.log_c));
    });
    for (size_t i = 0; i < items.size();) {
        size_t j = i;
        ntt_c_batch::Group g;
        g.ar = items[i].ar;
        memset(&g.inl, 0, sizeof(g.inl));
        g.sin = items[i].sin;
        g.sout = items[i].sout;
        g.kind = 0;
        std::vector<PassJob> jobs;
        uint32_t blocks = 0, lds = 0;
        while (j < items.size() && items[j].phase == items[i].phase && items[j].ar == g.ar && items[j].sin == g.sin &&
               items[j].sout == g.sout && launch_key(items[j].job) == launch_key(items[i].job)) {
            PassJob J = items[j].job;
            g.kind |= job_kind(J);
            J.tile_begin = blocks;
            blocks += job_tiles(J);
            lds = std::max(lds, job_lds(J, g.ar));
            jobs.push_back(J);
            j++;
        }
        jobs = merge_jobs(jobs, g.sin, g.sout);
        g.shape = group_shape(jobs.data(), jobs.size(), g.ar, g.sin, g.sout);
        g.nblocks = blocks;
        g.lds = lds;
        if (jobs.size() == 1) {  // one job: passed by value, no job table
            g.inl = jobs[0];
            g.njobs = 0;
        } else {
            g.njobs = (uint32_t)jobs.size();
            hipError_t e = hipMalloc(&g.djobs, jobs.size() * sizeof(PassJob));
            if (e != hipSuccess) return fail(NTT_ERR_OOM, "batch_prepare: job table allocation failed");
            e = hipMemcpy(g.djobs, jobs.data(), jobs.size() * sizeof(PassJob), hipMemcpyHostToDevice);
            if (e != hipSuccess) return fail(NTT_ERR_DEVICE, "batch_prepare: job table upload failed");
        }
        B->workgroups += blocks;
        B->phases[items[i].phase].push_back(g);
        i = j;
    }
    *out = B.release();
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_prepare(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count,
                                          ntt_c_batch** out) {
    return batch_prepare(ctx, descs, nullptr, count, out);
}

extern "C" ntt_status ntt_c_batch_prepare_ip(ntt_c_context* ctx, const ntt_c_desc* descs, const ntt_c_ip* ips,
                                             size_t count, ntt_c_batch** out) {
    if (count && !ips) return fail(NTT_ERR_INVALID_ARG, "batch_prepare_ip: null inner-product array");
    return batch_prepare(ctx, descs, ips, count, out);
}

static ntt_status batch_enqueue(ntt_c_batch* B, hipStream_t stream) {
    if (B->use_queue) {
        const ntt_c_batch::Queue& Q = B->queue;
        const hipError_t e =
            Q.mf ? launch_queue_kernel<Mod32, uint32_t, uint32_t, true, true>(Q.djobs, Q.tjob, Q.qbeg, Q.nq, Q.ctl,
                                                                               Q.tslot, Q.spin_limit, Q.proto,
                                                                               Q.nblocks, Q.lds, stream)
                 : launch_queue_kernel<Mod32, uint32_t, uint32_t, true, false>(Q.djobs, Q.tjob, Q.qbeg, Q.nq, Q.ctl,
                                                                                Q.tslot, Q.spin_limit, Q.proto,
                                                                                Q.nblocks, Q.lds, stream);
        HIP_TRY(e);
        return NTT_OK;
    }
    for (auto& ph : B->phases)
        for (auto& g : ph) {
            ntt_status st = NTT_OK;
            if (g.njobs == 0 && launch_pipe(g.inl, g.ar, g.sin, g.sout, g.shape, g.lds, stream, &st)) {
                if (st != NTT_OK) return st;
                continue;
            }
            st = launch_pass(g.ar, g.sin, g.sout, g.kind, g.djobs, g.njobs, g.inl, g.nblocks, g.lds, stream, g.shape);
            if (st != NTT_OK) return st;
        }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_launch(ntt_c_batch* B, void* stream) {
    if (!B) return fail(NTT_ERR_INVALID_ARG, "batch_launch: null batch");
    return batch_enqueue(B, (hipStream_t)stream);
}

extern "C" ntt_status ntt_c_batch_launch_graph(ntt_c_batch* B, void* stream) {
    if (!B) return fail(NTT_ERR_INVALID_ARG, "batch_launch_graph: null batch");
    if (B->exec && B->graph_mode != (int)B->use_queue) B->drop_graph();
    if (!B->exec) {
        hipStream_t cap = nullptr;
        HIP_TRY(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) {
            (void)hipStreamDestroy(cap);
            return fail(NTT_ERR_DEVICE, "hipStreamBeginCapture failed");
        }
        ntt_status st = batch_enqueue(B, cap);
        hipGraph_t graph = nullptr;
        e = hipStreamEndCapture(cap, &graph);
        (void)hipStreamDestroy(cap);
        if (st != NTT_OK) return st;
        if (e != hipSuccess) return fail(NTT_ERR_DEVICE, "hipStreamEndCapture failed");
        B->graph = graph;
        HIP_TRY(hipGraphInstantiate(&B->exec, graph, nullptr, nullptr, 0));
        B->graph_mode = (int)B->use_queue;
    }
    HIP_TRY(hipGraphLaunch(B->exec, (hipStream_t)stream));
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_destroy(ntt_c_batch* B) {
    if (B) {
        (void)hipDeviceSynchronize();
        delete B;
    }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_set_queue(ntt_c_batch* B, int enable, uint32_t spin_limit) {
    if (!B) return fail(NTT_ERR_INVALID_ARG, "batch_set_queue: null batch");
    if (enable && !B->queue.ok)
        return fail(NTT_ERR_UNSUPPORTED, "batch_set_queue: this mix needs more than one kernel instantiation");
    if (enable < 0 || enable > 2) return fail(NTT_ERR_INVALID_ARG, "batch_set_queue: enable must be 0, 1 or 2");
    B->use_queue = enable != 0;
    B->queue.spin_limit = spin_limit ? spin_limit : (1u << 22);
    const uint32_t proto = enable == 2 ? (uint32_t)QP_AGENT : B->queue.proto_auto;
    if (proto != B->queue.proto) {
        B->queue.proto = proto;
        B->drop_graph();
    }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_check(ntt_c_batch* B, void* stream) {
    if (!B) return fail(NTT_ERR_INVALID_ARG, "batch_check: null batch");
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    if (!B->queue.ok) return NTT_OK;
    uint32_t st = 0;
    HIP_TRY(hipMemcpy(&st, B->queue.ctl + kQStatus, sizeof(st), hipMemcpyDeviceToHost));
    if (st) {
        HIP_TRY(hipMemset(B->queue.ctl + kQStatus, 0, sizeof(st)));
        return fail(NTT_ERR_DEVICE, (st & 2u) ? "batch: a work queue was left with tiles no workgroup took"
                                             : "batch: a work-queue wait gave up before its inputs were complete");
    }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_batch_info(const ntt_c_batch* B, uint32_t* launches, uint64_t* ws, uint64_t* wgs) {
    if (!B) return fail(NTT_ERR_INVALID_ARG, "batch_info: null batch");
    if (launches) {
        uint32_t l = 0;
        for (auto& ph : B->phases) l += (uint32_t)ph.size();
        *launches = B->use_queue ? 1u : l;
    }
    if (ws) *ws = B->work_bytes;
    if (wgs) *wgs = B->use_queue ? B->queue.nblocks : B->workgroups;
    return NTT_OK;
}

extern "C" ntt_status ntt_c_context_reserve(ntt_c_context* ctx, void* stream, uint64_t bytes) {
    if (!ctx) return fail(NTT_ERR_INVALID_ARG, "context_reserve: null context");
    if (bytes == 0) return NTT_OK;
    return reserve_workspace(ctx->smu, ctx->work, (hipStream_t)stream, bytes);
}

// Workspace bytes of one transform of the descriptor's kind (the split it
// would get now; 0 for one-pass plans).
static uint64_t desc_workspace(const ntt_c_desc& d, const Split& sp) {
    if (sp.npass < 2) return 0;
    const Arith ar = arith_for(d.q, d.flags);
    const bool u64 = d.flags & NTT_FLAG_U64;
    const bool two = ar == Arith::M32 && u64 && sp.npass >= 3;
    return ((1ull << d.log_n) << (two ? 1 : 0)) * vbytes(ar);
}

extern "C" ntt_status ntt_c_context_prepare(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count, void* stream) {
    if (!ctx || (count && !descs)) return fail(NTT_ERR_INVALID_ARG, "context_prepare: null argument");
    uint64_t need = 0;
    for (size_t i = 0; i < count; i++) {
        ntt_status st = check_desc(descs[i]);
        if (st != NTT_OK) return st;
        const Plan* P = nullptr;
        st = ctx_plan(ctx, descs[i], 1ull << descs[i].log_n, &P);
        if (st != NTT_OK) return st;
        need = std::max(need, desc_workspace(descs[i], P->split));
    }
    return ntt_c_context_reserve(ctx, stream, need);
}

extern "C" ntt_status ntt_c_workspace_bytes(const ntt_c_desc* d, uint64_t* bytes) {
    if (!d || !bytes) return fail(NTT_ERR_INVALID_ARG, "workspace_bytes: null argument");
    ntt_status st = check_desc(*d);
    if (st != NTT_OK) return st;
    *bytes = desc_workspace(*d, choose_split(d->log_n, arith_for(d->q, d->flags), 1ull << d->log_n));
    return NTT_OK;
}

static ntt_status transform_one(ntt_c_context* ctx, const ntt_c_desc* d, const Split* forced, void* stream) {
    if (!ctx || !d) return fail(NTT_ERR_INVALID_ARG, "transform: null argument");
    ntt_status st = check_desc(*d);
    if (st != NTT_OK) return st;
    const Plan* P = nullptr;
    st = ctx_plan(ctx, *d, 1ull << d->log_n, &P, forced);
    if (st != NTT_OK) return st;
    // The four-step workspace is the stream's reservation (never allocated here).
    const Store sd = (d->flags & NTT_FLAG_U64) ? Store::U64 : Store::U32;
    Workspace ws;
    st = find_workspace(ctx->smu, ctx->work, (hipStream_t)stream, plan_scratch_elems(*P, sd) * vbytes(P->arith), &ws);
    if (st != NTT_OK) return st;
    return run_plan(*P, d->data, sd, 1, ws.ptr, (hipStream_t)stream);
}

extern "C" ntt_status ntt_c_transform(ntt_c_context* ctx, const ntt_c_desc* d, void* stream) {
    return transform_one(ctx, d, nullptr, stream);
}

// One transform call per descriptor, in order, on one stream: the
// reference's per-polynomial batch loop (src/ntt_basic.cu:337-362) issued from
// native code, each polynomial its own launch chain.
extern "C" ntt_status ntt_c_transform_each(ntt_c_context* ctx, const ntt_c_desc* descs, size_t count, void* stream) {
    if (!ctx || (count && !descs)) return fail(NTT_ERR_INVALID_ARG, "transform_each: null argument");
    for (size_t i = 0; i < count; i++) {
        ntt_status st = transform_one(ctx, &descs[i], nullptr, stream);
        if (st != NTT_OK) return st;
    }
    return NTT_OK;
}

extern "C" ntt_status ntt_c_transform_split(ntt_c_context* ctx, const ntt_c_desc* d, uint32_t npass,
                                            const uint32_t* log_r, uint32_t log_e, uint32_t mfma_mask, void* stream) {
    if (!log_r || npass < 1 || npass > 4) return fail(NTT_ERR_INVALID_ARG, "transform_split: bad split");
    Split sp;
    sp.npass = npass;
    for (uint32_t i = 0; i < npass; i++) sp.log_r[i] = log_r[i];
    sp.log_e = log_e;
    sp.mfma = mfma_mask & ((1u << npass) - 1u);
    if (d) {
        const uint32_t emax = max_log_e(arith_for(d->q, d->flags));
        for (uint32_t i = 0; i < npass; i++)
            if (log_r[i] < 1 || log_r[i] > emax) return fail(NTT_ERR_INVALID_ARG, "transform_split: radix too large");
        if (log_e > emax + 1) return fail(NTT_ERR_INVALID_ARG, "transform_split: tile too large");
    }
    return transform_one(ctx, d, &sp, stream);
}

#include "ntt_serve.inc"
#include "ntt_tune.inc"
#include "ntt_arb.inc"
#include "ntt_rns.inc"

// ------------------------------------------------------------------ device helpers
extern "C" ntt_status ntt_c_device_count(int* c) {
    if (!c) return fail(NTT_ERR_INVALID_ARG, "null");
    *c = 0;
    HIP_TRY(hipGetDeviceCount(c));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_malloc(void** p, uint64_t bytes) {
    if (!p) return fail(NTT_ERR_INVALID_ARG, "null");
    if (hipMalloc(p, bytes ? bytes : 1) != hipSuccess) return fail(NTT_ERR_OOM, "hipMalloc failed");
    return NTT_OK;
}
extern "C" ntt_status ntt_c_free(void* p) {
    if (p) HIP_TRY(hipFree(p));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_memcpy_h2d(void* dst, const void* src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_memcpy_d2h(void* dst, const void* src, uint64_t bytes) {
    HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_memset(void* dst, int value, uint64_t bytes) {
    HIP_TRY(hipMemset(dst, value, bytes));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_stream_create(void** s) {
    if (!s) return fail(NTT_ERR_INVALID_ARG, "null");
    HIP_TRY(hipStreamCreateWithFlags((hipStream_t*)s, hipStreamNonBlocking));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_stream_destroy(void* s) {
    if (s) HIP_TRY(hipStreamDestroy((hipStream_t)s));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_stream_sync(void* s) {
    HIP_TRY(hipStreamSynchronize((hipStream_t)s));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_device_sync(void) {
    HIP_TRY(hipDeviceSynchronize());
    return NTT_OK;
}
extern "C" ntt_status ntt_c_event_create(void** e) {
    if (!e) return fail(NTT_ERR_INVALID_ARG, "null");
    HIP_TRY(hipEventCreate((hipEvent_t*)e));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_event_destroy(void* e) {
    if (e) HIP_TRY(hipEventDestroy((hipEvent_t)e));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_event_record(void* e, void* s) {
    HIP_TRY(hipEventRecord((hipEvent_t)e, (hipStream_t)s));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_graph_begin(void* s) {
    if (!s) return fail(NTT_ERR_INVALID_ARG, "graph_begin: capture needs a created stream, not NULL");
    HIP_TRY(hipStreamBeginCapture((hipStream_t)s, hipStreamCaptureModeGlobal));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_graph_end(void* s, void** exec) {
    if (!s || !exec) return fail(NTT_ERR_INVALID_ARG, "graph_end: null argument");
    *exec = nullptr;
    hipGraph_t g = nullptr;
    HIP_TRY(hipStreamEndCapture((hipStream_t)s, &g));
    hipGraphExec_t x = nullptr;
    const hipError_t e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(NTT_ERR_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    *exec = x;
    return NTT_OK;
}
extern "C" ntt_status ntt_c_graph_launch(void* exec, void* s) {
    if (!exec) return fail(NTT_ERR_INVALID_ARG, "graph_launch: null graph");
    HIP_TRY(hipGraphLaunch((hipGraphExec_t)exec, (hipStream_t)s));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_graph_destroy(void* exec) {
    if (exec) HIP_TRY(hipGraphExecDestroy((hipGraphExec_t)exec));
    return NTT_OK;
}
extern "C" ntt_status ntt_c_event_elapsed_ms(float* ms, void* a, void* b) {
    if (!ms) return fail(NTT_ERR_INVALID_ARG, "null");
    HIP_TRY(hipEventSynchronize((hipEvent_t)b));
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return NTT_OK;
}
