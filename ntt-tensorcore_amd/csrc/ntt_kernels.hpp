// ntt_kernels.hpp — gfx950 device code for the modular NTT.
//
// One kernel shape does every pass of every transform: a workgroup loads a
// tile of C columns x R rows of a matrix view of one polynomial into LDS,
// runs an R-point Cooley-Tukey transform down every column (natural order in,
// bit-reversed order in LDS, "merged" twiddle table so the same loop computes
// the cyclic and the negacyclic transform), then streams the tile out in
// natural order through an optional per-element multiply:
//   * a power w^e with e = a*k*u + b*k + d*u (u = column >> shift), from two
//     small tables (four-step twiddles, the inverse negacyclic twist, n^-1),
//   * a constant         (n^-1 of the inverse cyclic transform).
// A single-pass transform is R = N, C polynomials per tile.  A P-pass
// (four-step) transform splits N = R_0*...*R_{P-1}: pass 0 runs R_0-point
// transforms down the columns and writes each tile transposed (contiguous
// rows), middle passes regroup, the last pass writes natural-order output in
// place.  The reference's driver (src/ntt_optimized.cu:515-603) makes
// 1 + log2N%6 + log2N/6 HBM passes for the same result; this makes P <= 3.
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

enum PostMode : uint32_t { POST_NONE = 0, POST_FOURSTEP = 2, POST_SCALE = 3 };

// One pass of a P-pass plan over one polynomial, or over `npoly` polynomials
// that share the plan (polynomial p starts at in + p*in_pstride and
// out + p*out_pstride; strides in elements).
//   input  element (r, c)  at  in[r*s_r + c*s_c]
//   output row k, column c at  out[(c >> o_lbits)*o_hs + (c & (2^o_lbits-1))*o_c + k*o_r]
// The output form covers the in-place store of the last pass, the transposed
// store of the first pass and the grouped store of middle passes.
//   post-multiply exponent  e = (e_a*k*u + e_b*k + e_d*u) mod (e_mask+1),
//   u = c >> e_shift, factor = lo[e mod 2^lo_bits] * hi[e >> lo_bits].
enum JobFlags : uint32_t {
    JF_LOAD_CFAST = 1u,   // consecutive threads walk columns (s_c == 1), else rows (s_r == 1)
    JF_STORE_CFAST = 2u,  // consecutive threads walk columns (o_c == 1), else rows (o_r == 1)
    JF_VEC4 = 4u,         // 4 consecutive elements per thread on the load side (aligned)
    JF_Z_LDS = 8u,        // stage the CT table in LDS
    JF_VEC4_ST = 16u,     // 4 consecutive elements per thread on the store side
    JF_MFMA = 32u,        // pass runs on the matrix cores (ntt_mfma.hpp)
};

struct PassJob {
    const void* in;
    void* out;
    const void* z;        // TwP<V>[R]: merged CT twiddles, Z[2^l + i]
    const void* tw_lo;    // post-multiply tables (TwP<V>)
    const void* tw_hi;
    uint64_t q;
    uint64_t scale_w, scale_wp;  // POST_SCALE constant (Shoup pair)
    int64_t s_r, s_c;
    int64_t o_r, o_c, o_hs;
    int64_t in_pstride, out_pstride;
    uint32_t log_r, log_c;       // R = 2^log_r rows, C = 2^log_c columns per tile
    uint32_t ncols;              // valid columns per polynomial (last tile may be partial)
    uint32_t tiles_per_poly;
    uint32_t o_lbits, post;
    uint32_t e_a, e_b, e_d, e_shift, e_mask, lo_bits;
    uint32_t tile_begin;         // first workgroup of this job in the launch
    uint32_t npoly;
    uint32_t flags;
    uint32_t pad_;
};

// XOR swizzle of the LDS word index: permutes the low 6 bits by higher bits so
// that both contiguous runs and the bit-reversed read-out hit distinct banks.
__device__ __forceinline__ uint32_t swz(uint32_t a) { return a ^ (((a >> 6) ^ (a >> 12)) & 63u); }

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t k, uint32_t bits) {
    return bits ? (__brev(k) >> (32u - bits)) : 0u;
}

// m register-resident CT levels [l0, l0+m) for one unit of 2^m elements.
template <int m, class M, class ZP>
__device__ __forceinline__ void ct_unit(typename M::V* lds, ZP z, const M& mod, uint32_t u, uint32_t l0,
                                        uint32_t log_r, uint32_t log_c) {
    using V = typename M::V;
    const uint32_t C = 1u << log_c;
    const uint32_t c = u & (C - 1u);
    const uint32_t rest = u >> log_c;
    const uint32_t log_t = log_r - l0 - m;  // log2 of the distance of the last level's pairs
    const uint32_t s = rest & ((1u << log_t) - 1u);
    const uint32_t i0 = rest >> log_t;
    const uint32_t base = (i0 << (log_r - l0)) + s;
    TwP<V> w[(1 << m) - 1];
#pragma unroll
    for (int lam = 0; lam < m; lam++)
#pragma unroll
        for (int g = 0; g < (1 << lam); g++) w[(1 << lam) - 1 + g] = z[(1u << (l0 + lam)) + (i0 << lam) + g];
    V x[1 << m];
#pragma unroll
    for (int v = 0; v < (1 << m); v++) x[v] = lds[swz(((base + ((uint32_t)v << log_t)) << log_c) + c)];
#pragma unroll
    for (int lam = 0; lam < m; lam++) {
        const int half = 1 << (m - 1 - lam);
#pragma unroll
        for (int g = 0; g < (1 << lam); g++) {
            const TwP<V> wt = w[(1 << lam) - 1 + g];
#pragma unroll
            for (int v0 = 0; v0 < half; v0++) {
                const int v = g * 2 * half + v0;
                const V a = x[v];
                const V b = mod.mul(x[v + half], wt);
                x[v] = mod.add(a, b);
                x[v + half] = mod.sub(a, b);
            }
        }
    }
#pragma unroll
    for (int v = 0; v < (1 << m); v++) lds[swz(((base + ((uint32_t)v << log_t)) << log_c) + c)] = x[v];
}

template <int MAXM, class M, class ZP>
__device__ __forceinline__ void ct_levels(typename M::V* lds, ZP z, const M& mod, uint32_t log_r, uint32_t log_c) {
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

template <class S>
__device__ __forceinline__ void load4(const S* p, S v[4]);
template <>
__device__ __forceinline__ void load4<uint32_t>(const uint32_t* p, uint32_t v[4]) {
    const uint4 t = *reinterpret_cast<const uint4*>(p);
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
template <>
__device__ __forceinline__ void load4<uint64_t>(const uint64_t* p, uint64_t v[4]) {
    const ulonglong2 a = reinterpret_cast<const ulonglong2*>(p)[0], b = reinterpret_cast<const ulonglong2*>(p)[1];
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <class S>
__device__ __forceinline__ void store4(S* p, const S v[4]);
template <>
__device__ __forceinline__ void store4<uint32_t>(uint32_t* p, const uint32_t v[4]) {
    *reinterpret_cast<uint4*>(p) = make_uint4(v[0], v[1], v[2], v[3]);
}
template <>
__device__ __forceinline__ void store4<uint64_t>(uint64_t* p, const uint64_t v[4]) {
    reinterpret_cast<ulonglong2*>(p)[0] = make_ulonglong2(v[0], v[1]);
    reinterpret_cast<ulonglong2*>(p)[1] = make_ulonglong2(v[2], v[3]);
}

// Loads per thread kept in flight before their LDS stores.
constexpr int kUnroll = 8;

template <class M, class Sin, class Sout, int MAXM>
__device__ __forceinline__ void run_pass(const PassJob& J, uint32_t tile_in_job, unsigned char* lds_raw) {
    using V = typename M::V;
    V* lds = reinterpret_cast<V*>(lds_raw);
    const uint32_t poly = tile_in_job / J.tiles_per_poly;
    const uint32_t tile = tile_in_job - poly * J.tiles_per_poly;
    const uint32_t log_r = J.log_r, log_c = J.log_c;
    const uint32_t R = 1u << log_r, C = 1u << log_c;
    const uint32_t col0 = tile << log_c;
    const uint32_t total = R << log_c;
    const uint32_t ncols = J.ncols;
    const uint32_t T = blockDim.x, tid = threadIdx.x;
    const Sin* __restrict__ in = reinterpret_cast<const Sin*>(J.in) + (int64_t)poly * J.in_pstride;
    Sout* __restrict__ out = reinterpret_cast<Sout*>(J.out) + (int64_t)poly * J.out_pstride;
    const M mod{(decltype(M::q))J.q};
    const uint32_t flags = J.flags;

    // CT table: staged in LDS behind the tile when it fits, else read from global.
    TwP<V>* zl = reinterpret_cast<TwP<V>*>(lds_raw + (size_t)total * sizeof(V));
    const TwP<V>* zg = reinterpret_cast<const TwP<V>*>(J.z);
    if (flags & JF_Z_LDS)
        for (uint32_t i = tid; i < R; i += T) zl[i] = zg[i];

    // ---- load ----
    if (flags & JF_VEC4) {
        const uint32_t nq = total >> 2;
        for (uint32_t q0 = tid; q0 < nq; q0 += kUnroll * T) {
            Sin v[kUnroll][4];
            uint32_t rr[kUnroll], cc[kUnroll];
            bool ok[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t qi = q0 + u * T;
                const uint32_t idx = qi << 2;
                if (flags & JF_LOAD_CFAST) { rr[u] = idx >> log_c; cc[u] = idx & (C - 1u); }
                else { cc[u] = idx >> log_r; rr[u] = idx & (R - 1u); }
                ok[u] = qi < nq && col0 + cc[u] < ncols;
                if (ok[u]) load4<Sin>(in + (int64_t)rr[u] * J.s_r + (int64_t)(col0 + cc[u]) * J.s_c, v[u]);
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                if (q0 + u * T >= nq) continue;
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t r = rr[u] + ((flags & JF_LOAD_CFAST) ? 0u : (uint32_t)e);
                    const uint32_t c = cc[u] + ((flags & JF_LOAD_CFAST) ? (uint32_t)e : 0u);
                    lds[swz((r << log_c) + c)] = ok[u] ? (V)v[u][e] : (V)0;
                }
            }
        }
    } else {
        for (uint32_t i0 = tid; i0 < total; i0 += kUnroll * T) {
            Sin v[kUnroll];
            uint32_t li[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t idx = i0 + u * T;
                uint32_t r, c;
                if (flags & JF_LOAD_CFAST) { r = idx >> log_c; c = idx & (C - 1u); }
                else { c = idx >> log_r; r = idx & (R - 1u); }
                li[u] = (r << log_c) + c;
                v[u] = 0;
                if (idx < total && col0 + c < ncols) v[u] = in[(int64_t)r * J.s_r + (int64_t)(col0 + c) * J.s_c];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (i0 + u * T < total) lds[swz(li[u])] = (V)v[u];
        }
    }
    __syncthreads();

    if (flags & JF_Z_LDS) ct_levels<MAXM>(lds, (const TwP<V>*)zl, mod, log_r, log_c);
    else ct_levels<MAXM>(lds, zg, mod, log_r, log_c);

    // ---- store natural order: output row k lives at LDS row bitrev(k) ----
    const TwP<V>* tlo = reinterpret_cast<const TwP<V>*>(J.tw_lo);
    const TwP<V>* thi = reinterpret_cast<const TwP<V>*>(J.tw_hi);
    const uint32_t post_mode = J.post;
    const TwP<V> scale{(V)J.scale_w, (V)J.scale_wp};
    const uint32_t lomask = (1u << J.lo_bits) - 1u;
    const uint32_t omask = (1u << J.o_lbits) - 1u;
    auto post = [&](V v, uint32_t k, uint32_t cg) -> V {
        if (post_mode == POST_FOURSTEP) {
            const uint32_t u = cg >> J.e_shift;
            const uint32_t e = (J.e_a * k * u + J.e_b * k + J.e_d * u) & J.e_mask;
            v = mod.mul(v, tlo[e & lomask]);
            return mod.mul(v, thi[e >> J.lo_bits]);
        }
        if (post_mode == POST_SCALE) return mod.mul(v, scale);
        return v;
    };
    auto oaddr = [&](uint32_t k, uint32_t cg) -> int64_t {
        return (int64_t)(cg >> J.o_lbits) * J.o_hs + (int64_t)(cg & omask) * J.o_c + (int64_t)k * J.o_r;
    };
    if (flags & JF_VEC4_ST) {
        const uint32_t nq = total >> 2;
        for (uint32_t q0 = tid; q0 < nq; q0 += kUnroll * T) {
            // Phase 1: positions, LDS reads and post-multiply for kUnroll quads;
            // phase 2: the stores.  Keeps the table loads of all quads in flight.
            Sout v[kUnroll][4];
            int64_t a[kUnroll];
            bool ok[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t qi = q0 + u * T;
                const uint32_t idx = qi << 2;
                uint32_t k, c;
                if (flags & JF_STORE_CFAST) { k = idx >> log_c; c = idx & (C - 1u); }
                else { c = idx >> log_r; k = idx & (R - 1u); }
                ok[u] = qi < nq && col0 + c < ncols;
                a[u] = oaddr(k, col0 + c);
#pragma unroll
                for (int e = 0; e < 4; e++) {
                    const uint32_t kk = k + ((flags & JF_STORE_CFAST) ? 0u : (uint32_t)e);
                    const uint32_t ce = c + ((flags & JF_STORE_CFAST) ? (uint32_t)e : 0u);
                    v[u][e] = ok[u] ? (Sout)post(lds[swz((bitrev_bits(kk & (R - 1u), log_r) << log_c) + (ce & (C - 1u)))],
                                                 kk, col0 + ce)
                                    : (Sout)0;
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (ok[u]) store4<Sout>(out + a[u], v[u]);
        }
    } else {
        for (uint32_t i0 = tid; i0 < total; i0 += kUnroll * T) {
            V v[kUnroll];
            int64_t a[kUnroll];
            bool ok[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t idx = i0 + u * T;
                uint32_t k, c;
                if (flags & JF_STORE_CFAST) { k = idx >> log_c; c = idx & (C - 1u); }
                else { c = idx >> log_r; k = idx & (R - 1u); }
                ok[u] = idx < total && col0 + c < ncols;
                a[u] = oaddr(k, col0 + c);
                v[u] = ok[u] ? post(lds[swz((bitrev_bits(k, log_r) << log_c) + c)], k, col0 + c) : (V)0;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (ok[u]) out[a[u]] = (Sout)v[u];
        }
    }
}

// Workgroup index -> logical tile so that consecutive tiles run on the same
// XCD: hardware dispatches workgroup b to XCD b % 8, and neighbouring tiles
// share cache lines of the column segments they read.
__device__ __forceinline__ uint32_t xcd_remap(uint32_t b, uint32_t nb) {
    const uint32_t per = nb >> 3, rem = nb & 7u;
    const uint32_t x = b & 7u, i = b >> 3;
    return x < rem ? x * (per + 1) + i : rem * (per + 1) + (x - rem) * per + i;
}

// A launch runs either the single job passed by value (njobs == 0) or one
// job per workgroup range from a device array sorted by tile_begin.
template <class M, class Sin, class Sout, int MAXM, int NT>
__global__ void __launch_bounds__(NT) pass_kernel(const PassJob* __restrict__ jobs, uint32_t njobs, PassJob inl) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const uint32_t b = xcd_remap(blockIdx.x, gridDim.x);
    if (njobs == 0) {
        run_pass<M, Sin, Sout, MAXM>(inl, b, lds_raw);
        return;
    }
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile_begin <= b) lo = mid; else hi = mid;
    }
    const PassJob J = jobs[lo];
    run_pass<M, Sin, Sout, MAXM>(J, b - J.tile_begin, lds_raw);
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
