// ntt_kernels.hpp — gfx950 device code for the modular NTT.
//
// One kernel shape does every pass of every transform: a workgroup loads a
// tile of C columns x R rows of a matrix view of one polynomial into LDS,
// runs an R-point Cooley-Tukey transform down every column (natural order in,
// bit-reversed order in LDS, "merged" twiddle table so the same loop computes
// the cyclic and the negacyclic transform), then streams the tile out in
// natural order through an optional per-element multiply:
//   * four-step twiddle  w^(a*k*c + b*k + d*c)  from two small tables,
//   * per-row table      (inverse negacyclic twist psi^-k, and/or n^-1),
//   * a constant         (n^-1 of the inverse cyclic transform).
// A single-pass transform is R = N, C polynomials per tile.  A two-pass
// (four-step) transform is N = N1*N2: pass 1 does N1-point transforms down
// the N2 columns and writes the tile transposed (contiguous rows), pass 2 does
// N2-point transforms and writes natural-order output.  The reference's
// driver (src/ntt_optimized.cu:515-603) does 1 + log2N%6 + log2N/6 HBM passes
// for the same result; this does 1 or 2.
//
// Arithmetic: canonical values in [0,q).  Products by a constant use Shoup's
// precomputed quotient w' = floor(w*2^B/q) (B = 32 or 64), which returns a
// value in [0,2q) that one conditional subtract makes canonical.  Outputs are
// therefore bit-identical to the reference's `% q` arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nttk {

template <class V>
struct TwP {  // Shoup pair: w and floor(w * 2^bits(V) / q)
    V w, wp;
};

// q < 2^31: every intermediate fits 32 bits (2q < 2^32).
struct Mod32 {
    using V = uint32_t;
    uint32_t q;
    __device__ __forceinline__ V reduce2(V r) const { return min(r, r - q); }  // [0,2q) -> [0,q)
    __device__ __forceinline__ V mul(V x, TwP<V> t) const {
        V hi = __umulhi(x, t.wp);
        return reduce2(x * t.w - hi * q);
    }
    __device__ __forceinline__ V add(V a, V b) const { return reduce2(a + b); }
    __device__ __forceinline__ V sub(V a, V b) const { return reduce2(a - b + q); }
};

// q < 2^62: 2q < 2^63, Shoup with a 64-bit quotient.
struct Mod64 {
    using V = uint64_t;
    uint64_t q;
    __device__ __forceinline__ V reduce2(V r) const { return r >= q ? r - q : r; }
    __device__ __forceinline__ V mul(V x, TwP<V> t) const {
        V hi = __umul64hi(x, t.wp);
        return reduce2(x * t.w - hi * q);
    }
    __device__ __forceinline__ V add(V a, V b) const { return reduce2(a + b); }
    __device__ __forceinline__ V sub(V a, V b) const { return reduce2(a - b + q); }
};

enum PostMode : uint32_t { POST_NONE = 0, POST_ROW = 1, POST_FOURSTEP = 2, POST_SCALE = 3 };

// One pass over one polynomial, or over `npoly` polynomials that share a plan
// (poly p starts at in + p*in_pstride / out + p*out_pstride).  Element (r, c)
// of the input view is at in[r*s_r + c*s_c]; output row k of column c goes to
// out[k*o_r + c*o_c].  Strides are in elements.
struct PassJob {
    const void* in;
    void* out;
    const void* z;         // TwP<V>[R]: merged CT twiddles, Z[2^l + i]
    const void* post_lo;   // POST_ROW: TwP<V>[R];  POST_FOURSTEP: low table
    const void* post_hi;   // POST_FOURSTEP: high table
    uint64_t q;
    uint64_t scale_w, scale_wp;  // POST_SCALE constant (Shoup pair)
    int64_t s_r, s_c, o_r, o_c;
    int64_t in_pstride, out_pstride;
    uint32_t log_r, log_c;       // R = 2^log_r rows, C = 2^log_c columns per tile
    uint32_t ncols;              // columns in the view; the last tile may be partial
    uint32_t post;               // PostMode
    uint32_t e_a, e_b, e_d;      // four-step exponent a*k*c + b*k + d*c
    uint32_t e_mask, lo_bits;    // exponent mod 2^bits; low table has 2^lo_bits entries
    uint32_t tiles_per_poly;     // ceil(ncols / C)
    uint32_t tile_begin;         // first workgroup index of this job in the launch
    uint32_t npoly;
};

// XOR swizzle of the LDS word index: permutes the low 6 bits by higher bits so
// that both contiguous runs and the bit-reversed read-out hit distinct banks.
__device__ __forceinline__ uint32_t swz(uint32_t a) { return a ^ (((a >> 6) ^ (a >> 12)) & 63u); }

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t k, uint32_t bits) {
    return bits ? (__brev(k) >> (32u - bits)) : 0u;
}

// m register-resident CT levels [l0, l0+m) for one unit of 2^m elements.
template <int m, class M>
__device__ __forceinline__ void ct_unit(typename M::V* lds, const TwP<typename M::V>* __restrict__ z,
                                        const M& mod, uint32_t u, uint32_t l0, uint32_t log_r,
                                        uint32_t log_c) {
    using V = typename M::V;
    const uint32_t C = 1u << log_c;
    const uint32_t c = u & (C - 1u);
    const uint32_t rest = u >> log_c;
    const uint32_t log_t = log_r - l0 - m;  // log2 of the distance of the last level's pairs
    const uint32_t s = rest & ((1u << log_t) - 1u);
    const uint32_t i0 = rest >> log_t;
    const uint32_t base = (i0 << (log_r - l0)) + s;
    V x[1 << m];
#pragma unroll
    for (int v = 0; v < (1 << m); v++) x[v] = lds[swz(((base + ((uint32_t)v << log_t)) << log_c) + c)];
#pragma unroll
    for (int lam = 0; lam < m; lam++) {
        const int half = 1 << (m - 1 - lam);
#pragma unroll
        for (int g = 0; g < (1 << lam); g++) {
            const TwP<V> w = z[(1u << (l0 + lam)) + (i0 << lam) + g];
#pragma unroll
            for (int v0 = 0; v0 < half; v0++) {
                const int v = g * 2 * half + v0;
                const V a = x[v];
                const V b = mod.mul(x[v + half], w);
                x[v] = mod.add(a, b);
                x[v + half] = mod.sub(a, b);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < (1 << m); v++) lds[swz(((base + ((uint32_t)v << log_t)) << log_c) + c)] = x[v];
}

template <int MAXM, class M>
__device__ __forceinline__ void ct_levels(typename M::V* lds, const TwP<typename M::V>* z, const M& mod,
                                          uint32_t log_r, uint32_t log_c) {
    const uint32_t nthreads = blockDim.x;
    for (uint32_t l0 = 0; l0 < log_r;) {
        const uint32_t m = min((uint32_t)MAXM, log_r - l0);
        const uint32_t units = 1u << (log_r + log_c - m);
        for (uint32_t u = threadIdx.x; u < units; u += nthreads) {
            switch (m) {
                case 1: ct_unit<1>(lds, z, mod, u, l0, log_r, log_c); break;
                case 2: ct_unit<2>(lds, z, mod, u, l0, log_r, log_c); break;
                case 3: ct_unit<3>(lds, z, mod, u, l0, log_r, log_c); break;
                default: ct_unit<(MAXM >= 4 ? 4 : 3)>(lds, z, mod, u, l0, log_r, log_c); break;
            }
        }
        l0 += m;
        __syncthreads();
    }
}

template <class M, class Sin, class Sout, int MAXM>
__device__ __forceinline__ void run_pass(const PassJob& J, uint32_t tile_in_job, typename M::V* lds) {
    using V = typename M::V;
    const uint32_t poly = tile_in_job / J.tiles_per_poly;
    const uint32_t tile = tile_in_job - poly * J.tiles_per_poly;
    const uint32_t log_r = J.log_r, log_c = J.log_c;
    const uint32_t R = 1u << log_r, C = 1u << log_c;
    const uint32_t col0 = tile << log_c;
    const uint32_t total = R << log_c;
    const uint32_t ncols = J.ncols;
    const Sin* __restrict__ in = reinterpret_cast<const Sin*>(J.in) + (int64_t)poly * J.in_pstride;
    Sout* __restrict__ out = reinterpret_cast<Sout*>(J.out) + (int64_t)poly * J.out_pstride;
    const M mod{(decltype(M::q))J.q};

    // ---- load: walk the tile along whichever stride is unit ----
    if (J.s_c == 1) {
        for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const uint32_t r = idx >> log_c, c = idx & (C - 1u);
            V v = 0;
            if (col0 + c < ncols) v = (V)in[(int64_t)r * J.s_r + (int64_t)(col0 + c)];
            lds[swz(idx)] = v;
        }
    } else {
        for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const uint32_t c = idx >> log_r, r = idx & (R - 1u);
            V v = 0;
            if (col0 + c < ncols) v = (V)in[(int64_t)r * J.s_r + (int64_t)(col0 + c) * J.s_c];
            lds[swz((r << log_c) + c)] = v;
        }
    }
    __syncthreads();

    ct_levels<MAXM>(lds, reinterpret_cast<const TwP<V>*>(J.z), mod, log_r, log_c);

    // ---- store natural order: output row k lives at LDS row bitrev(k) ----
    const TwP<V>* plo = reinterpret_cast<const TwP<V>*>(J.post_lo);
    const TwP<V>* phi = reinterpret_cast<const TwP<V>*>(J.post_hi);
    const TwP<V> scale{(V)J.scale_w, (V)J.scale_wp};
    const uint32_t post_mode = J.post;
    auto post = [&](V v, uint32_t k, uint32_t cg) -> V {
        switch (post_mode) {
            case POST_ROW: return mod.mul(v, plo[k]);
            case POST_FOURSTEP: {
                const uint32_t e = (J.e_a * k * cg + J.e_b * k + J.e_d * cg) & J.e_mask;
                v = mod.mul(v, plo[e & ((1u << J.lo_bits) - 1u)]);
                return mod.mul(v, phi[e >> J.lo_bits]);
            }
            case POST_SCALE: return mod.mul(v, scale);
            default: return v;
        }
    };
    if (J.o_c == 1) {
        for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const uint32_t k = idx >> log_c, c = idx & (C - 1u);
            if (col0 + c >= ncols) continue;
            V v = lds[swz((bitrev_bits(k, log_r) << log_c) + c)];
            v = post(v, k, col0 + c);
            out[(int64_t)k * J.o_r + (int64_t)(col0 + c)] = (Sout)v;
        }
    } else {
        for (uint32_t idx = threadIdx.x; idx < total; idx += blockDim.x) {
            const uint32_t c = idx >> log_r, k = idx & (R - 1u);
            if (col0 + c >= ncols) continue;
            V v = lds[swz((bitrev_bits(k, log_r) << log_c) + c)];
            v = post(v, k, col0 + c);
            out[(int64_t)k * J.o_r + (int64_t)(col0 + c) * J.o_c] = (Sout)v;
        }
    }
}

// A launch runs either the single job passed by value (njobs == 0) or one
// job per workgroup range from a device array sorted by tile_begin.
template <class M, class Sin, class Sout, int MAXM>
__global__ void __launch_bounds__(256) pass_kernel(const PassJob* __restrict__ jobs, uint32_t njobs, PassJob inl) {
    using V = typename M::V;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V* lds = reinterpret_cast<V*>(lds_raw);
    const uint32_t b = blockIdx.x;
    if (njobs == 0) {
        run_pass<M, Sin, Sout, MAXM>(inl, b, lds);
        return;
    }
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile_begin <= b) lo = mid; else hi = mid;
    }
    const PassJob J = jobs[lo];
    run_pass<M, Sin, Sout, MAXM>(J, b - J.tile_begin, lds);
}

// ---- small elementwise kernels ----

// r = a*b mod q with q < 2^31 on uint64 storage (poly_mul_kernel semantics,
// reference src/ntt_basic.cu:392-398), Barrett with mu = floor(2^64/q).
template <class S>
__global__ void pointwise_mul_kernel(S* r, const S* a, const S* b, uint64_t n, uint64_t q, uint64_t mu) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (; i < n; i += stride) {
        const uint64_t p = (uint64_t)a[i] * (uint64_t)b[i];  // < 2^62
        uint64_t t = p - __umul64hi(p, mu) * q;             // in [0, 2q)
        r[i] = (S)(t >= q ? t - q : t);
    }
}

// 62-bit pointwise product: 128-bit product reduced by Barrett with a
// 128-bit reciprocal (mu_hi:mu_lo = floor(2^124/q)).
__global__ void pointwise_mul62_kernel(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t q,
                                       uint64_t mu_lo, uint64_t mu_hi);

// In-place bit reversal of n = 2^log_n uint64 (reference bitrev semantics).
__global__ void bitrev_kernel(uint64_t* d, uint64_t n, uint32_t log_n);

// out[b*N1 + a] = in[a*N2 + b]  (N1 x N2 row-major transpose), 32x32 LDS tiles.
__global__ void transpose_u64_kernel(uint64_t* out, const uint64_t* in, uint32_t n1, uint32_t n2);

}  // namespace nttk
