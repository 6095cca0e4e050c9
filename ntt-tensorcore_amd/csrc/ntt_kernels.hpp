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
// 1 + log2N%6 + log2N/6 HBM passes for the same result; this makes P <= 4
// (three at 2^24).  With 32-bit arithmetic the CT levels of a pass may run as
// radix-16 rounds on the INT8 matrix cores instead (ntt_mf16.hpp).
//
// Arithmetic: canonical values in [0,q).  Products by a constant use Shoup's
// precomputed quotient w' = floor(w*2^B/q) (B = 32 or 64), which returns a
// value in [0,2q) that one conditional subtract makes canonical.  Outputs are
// therefore bit-identical to the reference's `% q` arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace nttk {

template <class V>
struct TwP {  // Shoup pair: w and floor(w * 2^bits(V) / q)
    V w, wp;
};

// q < 2^31: every intermediate fits 32 bits (2q < 2^32).
struct Mod32 {
    using V = uint32_t;
    static constexpr bool kLazy = false;  // butterflies keep canonical values (4q does not fit 32 bits)
    static constexpr int kMinWaves = 2;   // waves per SIMD a 512-thread workgroup needs anyway
    uint32_t q;
    __device__ __forceinline__ V reduce2(V r) const { return min(r, r - q); }  // [0,2q) -> [0,q)
    __device__ __forceinline__ V mul(V x, TwP<V> t) const { return reduce2(mul_lazy(x, t)); }
    // x * w mod q in [0, 2q), for any x < 2^32
    __device__ __forceinline__ V mul_lazy(V x, TwP<V> t) const { return x * t.w - __umulhi(x, t.wp) * q; }
    __device__ __forceinline__ V add(V a, V b) const { return reduce2(a + b); }
    // a - b wraps above 2^32 - q exactly when a < b, and then t + q = a - b + q
    // is the smaller of the two (2q < 2^32)
    __device__ __forceinline__ V sub(V a, V b) const {
        const V t = a - b;
        return min(t, t + q);
    }
};

// Montgomery product x * F * 2^-32 mod q (q < 2^31 odd, x, F < q), canonical.
// m = (x*F mod 2^32) * qinv makes m*q agree with x*F in the low word, so
// x*F - m*q = (hi(x*F) - hi(m*q)) * 2^32 exactly, and that difference lies in
// (-q, q): one conditional add of q (the min trick of Mod32::sub).
__device__ __forceinline__ uint32_t mont32(uint32_t x, uint32_t F, uint32_t q, uint32_t qinv) {
    const uint32_t lo = x * F, hi = __umulhi(x, F);
    const uint32_t r = hi - __umulhi(lo * qinv, q);
    return min(r, r + q);
}
// The same product as a representative in (0, 2q): r lies in (-q, q).
__device__ __forceinline__ uint32_t mont32_lazy(uint32_t x, uint32_t F, uint32_t q, uint32_t qinv) {
    const uint32_t lo = x * F, hi = __umulhi(x, F);
    return hi - __umulhi(lo * qinv, q) + q;
}

// q^-1 mod 2^32 for odd q: Newton's iteration from q (right to 3 bits).
__device__ __forceinline__ uint32_t inv_mod_2_32(uint32_t q) {
    uint32_t x = q;
#pragma unroll
    for (int i = 0; i < 4; i++) x *= 2u - q * x;
    return x;
}

// Montgomery product x * y * 2^-64 mod q (q < 2^62 odd, x, y < q), canonical;
// qneg_inv = -q^-1 mod 2^64.  x*y + m*q is a multiple of 2^64 whose low
// words sum to zero, carrying exactly when lo != 0; the quotient is < 2q.
__device__ __forceinline__ uint64_t mont64(uint64_t x, uint64_t y, uint64_t q, uint64_t qneg_inv) {
    const uint64_t lo = x * y, hi = __umul64hi(x, y);
    const uint64_t m = lo * qneg_inv;
    const uint64_t t = hi + __umul64hi(m, q) + (lo != 0);
    return t >= q ? t - q : t;
}

// a*b mod q of canonical residues for the product fused into a pass's load
// (JF_MUL_IN), by arithmetic: a Montgomery product, then one by R^2 mod q to
// remove R^-1 (Mod32, Mod64); the field's own product (ModGold).  An inner
// product sums the canonical prod() terms with add() and applies fin() once.
template <class M>
struct LoadMul;
struct Mod32;
struct Mod64;
struct ModGold;
template <>
struct LoadMul<Mod32> {
    uint32_t q, qinv, r2;
    __device__ __forceinline__ explicit LoadMul(uint64_t q_, uint64_t r2_)
        : q((uint32_t)q_), qinv(inv_mod_2_32((uint32_t)q_)), r2((uint32_t)r2_) {}
    __device__ __forceinline__ uint32_t prod(uint32_t a, uint32_t b) const { return mont32(a, b, q, qinv); }
    __device__ __forceinline__ uint32_t add(uint32_t a, uint32_t b) const {  // q < 2^31: no wrap
        const uint32_t s = a + b;
        return min(s, s - q);
    }
    __device__ __forceinline__ uint32_t fin(uint32_t x) const { return mont32(x, r2, q, qinv); }
    __device__ __forceinline__ uint32_t operator()(uint32_t a, uint32_t b) const { return fin(prod(a, b)); }
};
template <>
struct LoadMul<Mod64> {
    uint64_t q, qneg_inv, r2;
    __device__ __forceinline__ explicit LoadMul(uint64_t q_, uint64_t r2_) : q(q_), qneg_inv(0), r2(r2_) {
        uint64_t x = q_;  // Newton from q (right to 3 bits): 5 steps reach 96 bits
#pragma unroll
        for (int i = 0; i < 5; i++) x *= 2u - q_ * x;
        qneg_inv = 0 - x;
    }
    __device__ __forceinline__ uint64_t prod(uint64_t a, uint64_t b) const { return mont64(a, b, q, qneg_inv); }
    __device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b) const {  // q < 2^63: no wrap
        const uint64_t s = a + b;
        return s >= q ? s - q : s;
    }
    __device__ __forceinline__ uint64_t fin(uint64_t x) const { return mont64(x, r2, q, qneg_inv); }
    __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return fin(prod(a, b)); }
};

// q < 2^62: 2q < 2^63, Shoup with a 64-bit quotient.
// The VALU rounds run Harvey's lazy butterflies (kLazy): values between
// levels are any representatives below 4q (4q < 2^64), so a butterfly is one
// lazy Shoup product (any input below 2^64 gives [0, 2q)), one conditional
// subtraction of 2q on the other input, an add and a subtract:
//   a' = a mod 2q in [0, 2q),  t = b w mod q in [0, 2q),
//   (a' + t, a' - t + 2q), both in [0, 4q).
// The store canonicalises (canon4), or the post-multiply does (mul takes any
// input below 2^64 and returns [0, q)).
struct Mod64 {
    using V = uint64_t;
    static constexpr bool kLazy = true;
    // 4 waves per SIMD (at most 128 VGPRs): the 64-bit rounds otherwise take
    // a few registers more and drop to 3 waves, which costs far more than the
    // lazy butterflies save.
    static constexpr int kMinWaves = 4;
    uint64_t q;
    __device__ __forceinline__ V reduce2(V r) const { return r >= q ? r - q : r; }
    __device__ __forceinline__ V mul(V x, TwP<V> t) const { return reduce2(mul_lazy(x, t)); }
    __device__ __forceinline__ V mul_lazy(V x, TwP<V> t) const { return x * t.w - __umul64hi(x, t.wp) * q; }
    __device__ __forceinline__ V add(V a, V b) const { return reduce2(a + b); }
    __device__ __forceinline__ V sub(V a, V b) const { return reduce2(a - b + q); }
    __device__ __forceinline__ V half4(V a) const {  // [0, 4q) -> [0, 2q)
        const V q2 = q << 1;
        return a >= q2 ? a - q2 : a;
    }
    __device__ __forceinline__ V canon4(V a) const { return reduce2(half4(a)); }  // [0, 4q) -> [0, q)
};

// The Goldilocks field, q = 2^64 - 2^32 + 1.  2^64 = 2^32 - 1 and 2^96 = -1
// mod q, so the 128-bit product h*2^64 + l (h = h1*2^32 + h0) is congruent to
// l - h1 + h0*(2^32 - 1): a subtraction and an addition of 64-bit words, each
// corrected by 2^32 - 1 when it borrows or carries (an exact wrap of 2^64 is
// one of 2^32 - 1 mod q), give a representative below 2^64.  No Shoup
// quotient is needed (the tables' second word is ignored), and values stay
// canonical between levels (4q does not fit 64 bits).
constexpr uint64_t kGoldilocks = 0xFFFFFFFF00000001ull;
__device__ __forceinline__ uint64_t gold_reduce128(uint64_t lo, uint64_t hi) {
    constexpr uint64_t eps = 0xFFFFFFFFull;  // 2^64 mod q
    const uint64_t h1 = hi >> 32, h0 = hi & eps;
    uint64_t t = lo - h1;
    if (lo < h1) t -= eps;  // t wrapped by 2^64 (t >= 2^64 - 2^32, so no second wrap)
    const uint64_t u = (h0 << 32) - h0;  // h0 * (2^32 - 1) < 2^64
    uint64_t r = t + u;
    if (r < u) r += eps;  // carried: r < u < 2^64 - 2^33, so no second carry
    return r;
}
__device__ __forceinline__ uint64_t gold_mul_any(uint64_t x, uint64_t y) {  // [0, 2^64), any x, y
    return gold_reduce128(x * y, __umul64hi(x, y));
}
struct ModGold {
    using V = uint64_t;
    static constexpr bool kLazy = false;
    static constexpr int kMinWaves = 4;  // as Mod64: at most 128 VGPRs
    uint64_t q;                          // kGoldilocks (the job's modulus)
    __device__ __forceinline__ V reduce2(V r) const { return r >= kGoldilocks ? r - kGoldilocks : r; }  // [0, 2^64) -> [0, q)
    __device__ __forceinline__ V mul_lazy(V x, TwP<V> t) const { return gold_mul_any(x, t.w); }
    __device__ __forceinline__ V mul(V x, TwP<V> t) const { return reduce2(gold_mul_any(x, t.w)); }
    // a + b < 2q: a carry out of 64 bits leaves a + b - 2^64 + (2^32 - 1) = a + b - q < q
    __device__ __forceinline__ V add(V a, V b) const {
        const V s = a + b;
        return s < a ? s + 0xFFFFFFFFull : reduce2(s);
    }
    // a < b: a - b + 2^64 - (2^32 - 1) = a - b + q
    __device__ __forceinline__ V sub(V a, V b) const {
        const V d = a - b;
        return a < b ? d - 0xFFFFFFFFull : d;
    }
    __device__ __forceinline__ V half4(V a) const { return reduce2(a); }
    __device__ __forceinline__ V canon4(V a) const { return reduce2(a); }
};

template <>
struct LoadMul<ModGold> {
    __device__ __forceinline__ explicit LoadMul(uint64_t, uint64_t) {}
    __device__ __forceinline__ uint64_t prod(uint64_t a, uint64_t b) const {
        const uint64_t r = gold_mul_any(a, b);
        return r >= kGoldilocks ? r - kGoldilocks : r;
    }
    // a + b < 2q: a carry out of 64 bits leaves a + b - 2^64 + (2^32 - 1) = a + b - q < q
    __device__ __forceinline__ uint64_t add(uint64_t a, uint64_t b) const {
        const uint64_t s = a + b;
        return s < a ? s + 0xFFFFFFFFull : (s >= kGoldilocks ? s - kGoldilocks : s);
    }
    __device__ __forceinline__ uint64_t fin(uint64_t x) const { return x; }
    __device__ __forceinline__ uint64_t operator()(uint64_t a, uint64_t b) const { return prod(a, b); }
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
//   u = c >> e_shift, factor = lo[e mod 2^lo_bits] * mid[(e >> lo_bits) mod
//   2^mid_bits] * hi[e >> (lo_bits + mid_bits)] (no mid factor if mid_bits = 0),
//   or, when e is constant across a tile's columns (JF_POST_ROW), the row
//   table entry of (u, k) staged for the tile.
enum JobFlags : uint32_t {
    JF_LOAD_CFAST = 1u,   // consecutive threads walk columns (s_c == 1), else rows (s_r == 1)
    JF_STORE_CFAST = 2u,  // consecutive threads walk columns (o_c == 1), else rows (o_r == 1)
    JF_VEC4 = 4u,         // 4 consecutive elements per thread on the load side (aligned)
    JF_Z_LDS = 8u,        // stage the CT table in LDS
    JF_VEC4_ST = 16u,     // 4 consecutive elements per thread on the store side
    JF_MFMA = 32u,        // radix-16 rounds on the matrix cores (ntt_mf16.hpp), q < 2^31
    JF_POST4 = 64u,       // post == POST_FOURSTEP (selects the kernel instantiation)
    JF_POST_LDS = 128u,   // stage the post-multiply tables in LDS
    JF_CT_CYC = 256u,     // CT table is cyclic (Z[2^l] = 1): round 0 skips products by 1
    JF_POST_ROW = 512u,   // post from the row table (one factor per row k), staged in LDS
    JF_POST_STEP = 1024u, // row-fast first-pass store with per-column stepped factors (tw_col)
    JF_LAZY_OUT = 2048u,  // row-table or stepped post may leave values in [0, 2q): the next pass's
                          // matrix-core round takes any 32-bit representative
    JF_MUL_IN = 4096u,    // the load multiplies in[i] by in2[i] mod q (pointwise product fused into
                          // the first pass of an inverse transform); inputs canonical
};

struct PassJob {
    const void* in;
    void* out;
    const void* z;        // TwP<V>[R]: merged CT twiddles, Z[2^l + i]
    const void* tw_lo;    // post-multiply tables (TwP<V>)
    const void* tw_hi;
    const void* tw_mid;   // middle table of a three-table split (mid_bits > 0)
    const void* tw_row;   // row table: factor of (u, k) at [u*R + k], u = first column >> row_shift
    const void* tw_col;   // column table of JF_POST_STEP: 4 Shoup pairs per column
    const void* mz;       // JF_MFMA: matrix block of the radix-16 rounds (ntt_mf16.hpp)
    const void* in2;      // JF_MUL_IN: second factor, same layout and storage as in (first pair)
    uint64_t q;
    uint64_t mul_r2;      // JF_MUL_IN: R^2 mod q, R = 2^(8*sizeof(V)) (undoes the Montgomery factor)
    uint64_t scale_w, scale_wp;  // POST_SCALE constant; with JF_POST_STEP, 2^32 mod q (Shoup pair)
    int64_t s_r, s_c;
    int64_t o_r, o_c, o_hs;
    int64_t in_pstride, out_pstride;
    uint32_t log_r, log_c;       // R = 2^log_r rows, C = 2^log_c columns per tile
    uint32_t ncols;              // valid columns per polynomial (last tile may be partial)
    uint32_t tiles_per_poly;
    uint32_t o_lbits, post;
    uint32_t e_a, e_b, e_d, e_shift, e_mask, lo_bits, mid_bits;
    uint32_t tile_begin;         // first workgroup of this job in the launch
    uint32_t npoly;
    uint32_t flags;
    uint32_t row_shift;
    // log2 of s_r, s_c, o_r, o_c, o_hs (all powers of two; a zero stride only
    // ever multiplies a zero term, so its shift is arbitrary)
    uint8_t sh_sr, sh_sc, sh_or, sh_oc;
    uint8_t sh_ohs, pad_[3];
    // Work-queue launches (queue_kernel): the tiles of this job start once
    // dep_tiles tiles of job `dep` (the previous pass of the same polynomial)
    // have completed; dep_tiles = 0: no dependency.  counted: the job's tiles
    // count their completions (a later job depends on it, or it has a
    // dependency itself).
    uint32_t dep, dep_tiles, counted;
    // JF_MUL_IN: the load sums ip_n >= 1 products in[i + k*ip_stride] *
    // in2[i + k*ip_stride], k < ip_n (an inner product; 0 counts as 1)
    uint32_t ip_n;
    int64_t ip_stride;
};

// Threads of every pass-kernel workgroup (a compile-time constant so that the
// uniform index arithmetic folds).
constexpr uint32_t kLogNT = 9;
constexpr uint32_t kNT = 1u << kLogNT;

// This thread's index, opaque to the optimiser: the per-lane offsets each
// phase of a pass derives from it are recomputed where the phase runs instead
// of being hoisted out of a tile loop (pipe_kernel) and held in registers
// across the whole loop.
__device__ __forceinline__ uint32_t ltid() {
    uint32_t t = threadIdx.x;
    __asm__ volatile("" : "+v"(t));
    return t;
}

// LDS addressing by byte offset.  The pass kernel declares no static LDS, so
// its dynamic LDS starts at LDS address 0 (launch_pass_kernel checks the
// kernel's static size once) and a byte offset is the address itself.  Built
// this way the compiler folds constant parts of an address into the
// instruction's offset field; through the extern array it adds the array's
// (link-time) address to every address, one VALU op per access.
template <class T>
using lds_p = __attribute__((address_space(3))) T*;
template <class T>
__device__ __forceinline__ T lds_ld(uint32_t off) {
    return *reinterpret_cast<lds_p<const T>>((size_t)off);
}
template <class T>
__device__ __forceinline__ void lds_st(uint32_t off, T v) {
    *reinterpret_cast<lds_p<T>>((size_t)off) = v;
}

// Load through a global-address-space pointer: the job's table pointers are
// generic (void*), and a generic load is a flat load, which the compiler
// must also order against LDS accesses.
// (The host pass of the compiler has no global address space.)
template <class T>
__device__ __forceinline__ T gld(const T* p) {
#if defined(__HIP_DEVICE_COMPILE__)
    return *(const __attribute__((address_space(1))) T*)p;
#else
    return *p;
#endif
}

// Store through a global-address-space pointer (the coefficient buffers are
// device memory): a generic store would be a flat store, which also counts in
// lgkmcnt, so every LDS wait (lds_barrier) would wait for it too.
template <class T>
__device__ __forceinline__ void gst(T* p, T v) {
#if defined(__HIP_DEVICE_COMPILE__)
    *(__attribute__((address_space(1))) T*)p = v;
#else
    *p = v;
#endif
}

// A table of Shoup pairs staged in LDS at byte offset `off` (a multiple of the
// pair's size): indexes and offsets like the global pointer it replaces.  A
// pair moves as one aligned 8- or 16-byte access.
template <class V>
struct LdsTw {
    typedef uint64_t U64x2 __attribute__((ext_vector_type(2)));
    using Raw = typename std::conditional<sizeof(V) == 4, uint64_t, U64x2>::type;  // TwP<V>'s bytes, aligned
    uint32_t off;
    __device__ __forceinline__ TwP<V> operator[](uint32_t i) const {
        const Raw r = lds_ld<Raw>(off + i * (uint32_t)sizeof(TwP<V>));
        if constexpr (sizeof(V) == 4) return TwP<V>{(V)r, (V)(r >> 32)};
        else return TwP<V>{(V)r.x, (V)r.y};
    }
    __device__ __forceinline__ LdsTw operator+(uint32_t i) const { return LdsTw{off + i * (uint32_t)sizeof(TwP<V>)}; }
    __device__ __forceinline__ void set(uint32_t i, TwP<V> t) const {
        Raw r;
        if constexpr (sizeof(V) == 4) r = (uint64_t)t.w | ((uint64_t)t.wp << 32);
        else r = Raw{(uint64_t)t.w, (uint64_t)t.wp};
        lds_st<Raw>(off + i * (uint32_t)sizeof(TwP<V>), r);
    }
};

// XOR swizzle of the LDS word index: permutes the low 6 bits by higher bits so
// that both contiguous runs and the bit-reversed read-out hit distinct banks.
__device__ __forceinline__ uint32_t swz(uint32_t a) { return a ^ (((a >> 6) ^ (a >> 12)) & 63u); }

// Workgroup barrier that orders LDS accesses only: waits for this wave's LDS
// operations (lgkmcnt 0, vmcnt and expcnt left at their maxima), then
// s_barrier.  __syncthreads() would also wait for every outstanding global
// load and store (vmcnt 0), which serialises tile loads in flight with the
// CT rounds.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_s_waitcnt(0xC07F);
    __builtin_amdgcn_s_barrier();
    __asm__ volatile("" ::: "memory");
}

// Position of entry i of a post-multiply table staged in LDS: XORs bits 5..9
// into bits 0..4 (a bijection on every power-of-two table size) so that the
// strided exponents of one wave spread over the banks.
__device__ __forceinline__ uint32_t tsw(uint32_t i) { return i ^ ((i >> 5) & 31u); }

__device__ __forceinline__ uint32_t bitrev_bits(uint32_t k, uint32_t bits) {
    return bits ? (__brev(k) >> (32u - bits)) : 0u;
}

// Address arithmetic.  Every map from an element index of a tile to an LDS
// word (row-major r*C + c, bit-reversed read-out, the XOR swizzle below) is
// linear over GF(2): map(a ^ b) = map(a) ^ map(b).  A thread's index is
// lane_bits | uniform_bits (disjoint fields), so an LDS address is
// map(lane part), computed once per thread, XOR map(uniform part), which the
// scalar unit computes: one vector XOR per access.  Global offsets are split
// the same way into a 32-bit per-lane offset and a wave-uniform base pointer.

// m register-resident CT levels [l0, l0+m) for one unit of 2^m elements.
// slot0 = LDS byte offset of the unit's first element, vsl[v] = the byte XOR
// offset of its v-th element; zb[lam] points at the level-(l0+lam) twiddles.
// ONE0: the first twiddle of every level, zb[lam][0], is 1 (round 0 of a
// cyclic table: Z[2^l] = fac^(R >> (l+1)) with fac = 1), and those butterflies
// skip the product: 2^m - 1 of the unit's m * 2^(m-1) products.
template <int m, bool ONE0, class M, class ZP>
__device__ __forceinline__ void ct_unit(const ZP* zb, const M& mod, uint32_t slot0, const uint32_t* vsl) {
    using V = typename M::V;
    TwP<V> w[(1 << m) - 1];
#pragma unroll
    for (int lam = 0; lam < m; lam++)
#pragma unroll
        for (int g = ONE0 ? 1 : 0; g < (1 << lam); g++) {
            // member-wise: zb may point into the constant address space
            w[(1 << lam) - 1 + g] = TwP<V>{zb[lam][g].w, zb[lam][g].wp};
        }
    V x[1 << m];
#pragma unroll
    for (int v = 0; v < (1 << m); v++) x[v] = lds_ld<V>(slot0 ^ vsl[v]);
#pragma unroll
    for (int lam = 0; lam < m; lam++) {
        const int half = 1 << (m - 1 - lam);
#pragma unroll
        for (int g = 0; g < (1 << lam); g++) {
            const TwP<V> wt = w[(1 << lam) - 1 + g];
#pragma unroll
            for (int v0 = 0; v0 < half; v0++) {
                const int v = g * 2 * half + v0;
                if constexpr (M::kLazy) {
                    const V a = mod.half4(x[v]);
                    const V t = (ONE0 && g == 0) ? mod.half4(x[v + half]) : mod.mul_lazy(x[v + half], wt);
                    x[v] = a + t;
                    x[v + half] = a - t + (mod.q << 1);
                } else {
                    const V a = x[v];
                    const V b = (ONE0 && g == 0) ? x[v + half] : mod.mul(x[v + half], wt);
                    x[v] = mod.add(a, b);
                    x[v + half] = mod.sub(a, b);
                }
            }
        }
    }
#pragma unroll
    for (int v = 0; v < (1 << m); v++) lds_st<V>(slot0 ^ vsl[v], x[v]);
}

// One round of m levels starting at level l0 over the whole tile.  A unit is
// one column c and 2^m rows base + v*2^t, t = log_r - l0 - m, base = i0*2^(t+m)
// + s (s < 2^t); unit index u = (i0 * 2^t + s) * C + c.
// ONE0 (l0 = 0 only): the table is cyclic, see ct_unit.
template <int m, bool ONE0, class M, class ZP>
__device__ __forceinline__ void ct_round(ZP z, const M& mod, uint32_t l0, uint32_t log_r, uint32_t log_c) {
    const uint32_t log_t = log_r - l0 - m;
    const uint32_t units = 1u << (log_r + log_c - m);
    constexpr uint32_t VB = sizeof(typename M::V);
    uint32_t vsl[1 << m];
#pragma unroll
    for (int v = 0; v < (1 << m); v++) vsl[v] = swz((uint32_t)v << (log_t + log_c)) * VB;
    const uint32_t lowmask = (1u << (log_t + log_c)) - 1u;
    // (threadIdx.x, not ltid(): the 62-bit instances spill with the opaque index)
    for (uint32_t u = threadIdx.x; u < units; u += kNT) {
        // row-major word of (base, c): u's low log_t + log_c bits stay, i0 moves up by m
        const uint32_t word = (u & lowmask) | ((u & ~lowmask) << m);
        // i0 < 2^l0: at l0 = 0 every unit reads the same table entries, and
        // the constant lets them come in as scalar loads
        const uint32_t i0 = l0 == 0 ? 0u : u >> (log_t + log_c);
        ZP zb[m];
#pragma unroll
        for (int lam = 0; lam < m; lam++) zb[lam] = z + ((1u << (l0 + lam)) + (i0 << lam));
        ct_unit<m, ONE0, M>(zb, mod, swz(word) * VB, vsl);
    }
}

// All log_r levels, in rounds of at most MAXM levels balanced so that no
// round is much shorter than the others.  cyc: the table is cyclic (fac = 1),
// so round 0 skips its products by 1.
template <int MAXM, class M, class ZP>
__device__ __forceinline__ void ct_levels(ZP z, const M& mod, uint32_t log_r, uint32_t log_c, bool cyc) {
    const uint32_t rounds = (log_r + MAXM - 1) / MAXM;
    uint32_t l0 = 0;
    for (uint32_t i = 0; i < rounds; i++) {
        const uint32_t left = log_r - l0, m = (left + (rounds - i) - 1) / (rounds - i);
        constexpr int M4 = MAXM >= 4 ? 4 : 3;
        if (i == 0 && cyc) {
            switch (m) {
                case 1: ct_round<1, true, M>(z, mod, l0, log_r, log_c); break;
                case 2: ct_round<2, true, M>(z, mod, l0, log_r, log_c); break;
                case 3: ct_round<3, true, M>(z, mod, l0, log_r, log_c); break;
                default: ct_round<M4, true, M>(z, mod, l0, log_r, log_c); break;
            }
        } else {
            switch (m) {
                case 1: ct_round<1, false, M>(z, mod, l0, log_r, log_c); break;
                case 2: ct_round<2, false, M>(z, mod, l0, log_r, log_c); break;
                case 3: ct_round<3, false, M>(z, mod, l0, log_r, log_c); break;
                default: ct_round<M4, false, M>(z, mod, l0, log_r, log_c); break;
            }
        }
        l0 += m;
        lds_barrier();
    }
}

// ct_levels for a tile shape fixed at compile time: the rounds unrolled with
// constant first levels, so every uniform offset is an immediate.
template <int LR, int LC, int MAXM, int L0, class M, class ZP>
__device__ __forceinline__ void ct_levels_c(ZP z, const M& mod, bool cyc) {
    if constexpr (L0 < LR) {
        constexpr int rounds = (LR - L0 + MAXM - 1) / MAXM;
        constexpr int m = (LR - L0 + rounds - 1) / rounds;
        if (L0 == 0 && cyc) ct_round<m, true, M>(z, mod, (uint32_t)L0, (uint32_t)LR, (uint32_t)LC);
        else ct_round<m, false, M>(z, mod, (uint32_t)L0, (uint32_t)LR, (uint32_t)LC);
        lds_barrier();
        ct_levels_c<LR, LC, MAXM, L0 + m, M, ZP>(z, mod, false);
    }
}

}  // namespace nttk

#include "ntt_mf16.hpp"

namespace nttk {

template <class S>
__device__ __forceinline__ void load4(const S* p, S v[4]);
template <>
__device__ __forceinline__ void load4<uint32_t>(const uint32_t* p, uint32_t v[4]) {
    const uint4 t = gld(reinterpret_cast<const uint4*>(p));
    v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
}
template <>
__device__ __forceinline__ void load4<uint64_t>(const uint64_t* p, uint64_t v[4]) {
    const ulonglong2 a = gld(reinterpret_cast<const ulonglong2*>(p)), b = gld(reinterpret_cast<const ulonglong2*>(p) + 1);
    v[0] = a.x; v[1] = a.y; v[2] = b.x; v[3] = b.y;
}
template <class S>
__device__ __forceinline__ void store4(S* p, const S v[4]);
template <>
__device__ __forceinline__ void store4<uint32_t>(uint32_t* p, const uint32_t v[4]) {
    gst(reinterpret_cast<uint4*>(p), make_uint4(v[0], v[1], v[2], v[3]));
}
template <>
__device__ __forceinline__ void store4<uint64_t>(uint64_t* p, const uint64_t v[4]) {
    gst(reinterpret_cast<ulonglong2*>(p), make_ulonglong2(v[0], v[1]));
    gst(reinterpret_cast<ulonglong2*>(p) + 1, make_ulonglong2(v[2], v[3]));
}

// Vector loads (or stores) per thread kept in flight: 8 of 32-bit words.  The
// load loop holds the same bytes, so fewer of 64-bit words or of two inputs
// (each extra load register costs occupancy).
constexpr int kUnroll = 8;

// Natural-order store of a tile whose CT output sits at LDS row bitrev(k),
// with the pass's post-multiply.  Template parameters fix everything that is
// uniform over a job so that the loop body is straight-line code:
//   PER  elements per thread and sweep (4: vector store, 1: scalar),
//   POST 0: none, 1: constant (POST_SCALE), 2: power table (POST_FOURSTEP),
//   NTAB tables of the power (2 or 3), QM exponent scheme within a quad
//   (0: per element, 1: one exponent per quad, 2: a fixed step per element,
//   3: row table, the factor of row k at tl[k]),
//   PRED per-element bounds predicates (partial tiles, tiles smaller than a sweep),
//   SWT  tables are staged in LDS at tsw(i),
//   LR, LC the tile shape log2 R, log2 C when fixed at compile time (LR = 0:
//   taken from the job),
//   LZ   (QM 3) outputs in [0, 2q) instead of canonical (JF_LAZY_OUT),
//   CF   the store direction (JF_STORE_CFAST) when fixed at compile time (1 or
//        0; -1: taken from the job), so that shape-specialised kernels fold
//        the uniform part of every address into constants and scalars.
template <int PER, int POST, int NTAB, int QM, bool PRED, bool SWT, int LR, int LC, class M, class Sout,
          bool LZ = false, int CF = -1, class TP>
__device__ __forceinline__ void store_loop(const PassJob& J, const M& mod, uint32_t poly, uint32_t col0, TP tl, TP tm,
                                           TP th) {
    using V = typename M::V;
    constexpr uint32_t VB = sizeof(V);
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r, log_c = LR ? (uint32_t)LC : J.log_c;
    const uint32_t R = 1u << log_r, C = 1u << log_c;
    const uint32_t total = R << log_c;
    const uint32_t ncols = J.ncols;
    const uint32_t T = kNT, tid = ltid();
    const bool cf = CF < 0 ? (J.flags & JF_STORE_CFAST) != 0 : CF != 0;
    const uint32_t sweep = PER * T;
    const TwP<V> scale{(V)J.scale_w, (V)J.scale_wp};
    const uint32_t o_lbits = J.o_lbits, omask = (1u << o_lbits) - 1u;
    const uint32_t sh_or = J.sh_or, sh_oc = J.sh_oc, sh_ohs = J.sh_ohs;
    const uint32_t e_a = J.e_a, e_b = J.e_b, e_d = J.e_d, e_shift = J.e_shift, e_mask = J.e_mask;
    const uint32_t b0 = J.lo_bits, b1 = J.mid_bits, b01 = b0 + b1;
    const uint32_t m0 = (1u << b0) - 1u, m1 = (1u << b1) - 1u;
    auto expo = [&](uint32_t k, uint32_t cg) -> uint32_t {
        const uint32_t u = cg >> e_shift;
        return (e_a * k * u + e_b * k + e_d * u) & e_mask;
    };
    auto ti = [&](uint32_t i) -> uint32_t { return SWT ? tsw(i) : i; };
    auto twmul = [&](V x, uint32_t e) -> V {
        x = mod.mul(x, tl[ti(e & m0)]);
        if (NTAB == 3) x = mod.mul(x, tm[ti((e >> b0) & m1)]);
        return mod.mul(x, th[ti(e >> b01)]);
    };
    auto kc = [&](uint32_t i, uint32_t& k, uint32_t& c) {
        if (cf) { k = i >> log_c; c = i & (C - 1u); }
        else { c = i >> log_r; k = i & (R - 1u); }
    };
    // Output element offset of (k, column c of the tile) from the tile base;
    // additive over disjoint bit fields of k and c (col0, a multiple of C, is
    // in the base pointer).
    auto ooff = [&](uint32_t k, uint32_t c) -> uint32_t {
        return ((c >> o_lbits) << sh_ohs) + ((c & omask) << sh_oc) + (k << sh_or);
    };
    auto slot = [&](uint32_t k, uint32_t c) -> uint32_t { return swz((bitrev_bits(k, log_r) << log_c) | c) * VB; };
    const uint32_t li = tid * PER;
    uint32_t lk, lc;
    kc(li, lk, lc);
    const bool lane_ok = li < total;
    const uint32_t l_off = ooff(lk, lc), l_slot = slot(lk, lc);
    uint32_t e_x[PER];  // byte XOR offsets of quad element e (consecutive c, or consecutive k)
#pragma unroll
    for (int e = 0; e < PER; e++)
        e_x[e] = cf ? (uint32_t)e * VB : swz(bitrev_bits((uint32_t)e, log_r) << log_c) * VB;
    Sout* __restrict__ out = reinterpret_cast<Sout*>(J.out) + (int64_t)poly * J.out_pstride +
                             ((int64_t)(col0 >> o_lbits) << sh_ohs) + ((int64_t)(col0 & omask) << sh_oc);
    for (uint32_t bb = 0; bb < total; bb += kUnroll * sweep) {
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            const uint32_t ui = bb + (uint32_t)u * sweep;
            if (ui >= total) break;
            uint32_t uk, uc;
            kc(ui, uk, uc);
            const uint32_t c = uc | lc, k = uk | lk, cg = col0 + c;
            const uint32_t s = l_slot ^ slot(uk, uc);
            V x[PER];
#pragma unroll
            for (int e = 0; e < PER; e++) x[e] = (!PRED || lane_ok) ? lds_ld<V>(s ^ e_x[e]) : (V)0;
            if (POST == 2) {
                if (QM == 3) {  // row table staged for this tile: entry k
                    if (cf) {
                        const TwP<V> f = tl[k];
#pragma unroll
                        for (int e = 0; e < PER; e++) x[e] = LZ ? mod.mul_lazy(x[e], f) : mod.mul(x[e], f);
                    } else {
#pragma unroll
                        for (int e = 0; e < PER; e++)
                            x[e] = LZ ? mod.mul_lazy(x[e], tl[k + (uint32_t)e]) : mod.mul(x[e], tl[k + (uint32_t)e]);
                    }
                } else if (QM == 1) {
                    const uint32_t ex = expo(k, cg);
                    const TwP<V> f0 = tl[ti(ex & m0)], f2 = th[ti(ex >> b01)];
                    TwP<V> f1{};
                    if (NTAB == 3) f1 = tm[ti((ex >> b0) & m1)];
#pragma unroll
                    for (int e = 0; e < PER; e++) {
                        V y = mod.mul(x[e], f0);
                        if (NTAB == 3) y = mod.mul(y, f1);
                        x[e] = mod.mul(y, f2);
                    }
                } else if (QM == 2) {
                    uint32_t ex = expo(k, cg);
                    const uint32_t stp = e_a * (cg >> e_shift) + e_b;
#pragma unroll
                    for (int e = 0; e < PER; e++) {
                        x[e] = twmul(x[e], ex);
                        ex = (ex + stp) & e_mask;
                    }
                } else {
#pragma unroll
                    for (int e = 0; e < PER; e++)
                        x[e] = twmul(x[e], expo(cf ? k : k + (uint32_t)e, cf ? cg + (uint32_t)e : cg));
                }
            } else if (POST == 1) {
#pragma unroll
                for (int e = 0; e < PER; e++) x[e] = mod.mul(x[e], scale);
            } else if constexpr (M::kLazy) {  // lazy butterfly outputs below 4q
#pragma unroll
                for (int e = 0; e < PER; e++) x[e] = mod.canon4(x[e]);
            }
            if (!PRED || (lane_ok && cg < ncols)) {
                Sout* __restrict__ p = out + (ooff(uk, uc) + l_off);
                if (PER == 4) {
                    Sout v[4] = {(Sout)x[0], (Sout)x[PER > 1 ? 1 : 0], (Sout)x[PER > 2 ? 2 : 0], (Sout)x[PER > 3 ? 3 : 0]};
                    store4<Sout>(p, v);
                } else {
                    gst(p, (Sout)x[0]);
                }
            }
        }
    }
}

// Store of a first pass (row-fast output, 4 consecutive rows k per vector
// store) whose post exponent is linear in k for each column:
// e(k, c) = s(c)*k + e_d*c with s(c) = e_a*c + e_b.  Each thread keeps one
// column and walks down its rows, so the factor steps by constants of that
// column instead of being looked up per element: thread t takes column
// t >> log_tpc and rows 4m + kstep*u + j (m = t mod 2^log_tpc,
// kstep = 4 * 2^log_tpc, j < 4).  G, the factor of row 4m + kstep*u in
// Montgomery form, starts from the power tables and steps by
// Wb = base^(s*kstep); rows +1..+3 first take base^(s*j) (Shoup), then every
// element one Montgomery product by G.  Column table tw_col[c] = {base^s,
// base^2s, base^3s, base^(s*kstep)}; 32-bit arithmetic, full tiles.
// LZ (JF_LAZY_OUT): outputs in (0, 2q) instead of canonical, for a next pass
// whose first round runs on the matrix cores (any 32-bit representative).
template <int NTAB, bool SWT, int LR, int LC, class Sout, bool LZ, class TP>
__device__ __forceinline__ void store_loop_step(const PassJob& J, const Mod32& mod, uint32_t poly, uint32_t col0,
                                                TP tl, TP tm, TP th) {
    using V = uint32_t;
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r, log_c = LR ? (uint32_t)LC : J.log_c;
    const uint32_t tid = ltid();
    const uint32_t log_tpc = kLogNT - log_c;  // 2^log_tpc threads per column
    const uint32_t cl = tid >> log_tpc, m = tid & ((1u << log_tpc) - 1u);
    const uint32_t cg = col0 + cl;
    const uint32_t kstep = 4u << log_tpc;
    const uint32_t q = mod.q, qinv = inv_mod_2_32(q);
    const TwP<V>* ct = reinterpret_cast<const TwP<V>*>(J.tw_col) + (size_t)cg * 4u;
    const TwP<V> B1 = gld(ct), B2 = gld(ct + 1), B3 = gld(ct + 2), Wb = gld(ct + 3);
    // G = base^(e_d*c + s*4m) * mult * 2^32 mod q
    const uint32_t b0 = J.lo_bits, b1 = J.mid_bits, b01 = b0 + b1;
    const uint32_t e0 = (J.e_d * cg + (J.e_a * cg + J.e_b) * (4u * m)) & J.e_mask;
    auto ti = [&](uint32_t i) -> uint32_t { return SWT ? tsw(i) : i; };
    V G = tl[ti(e0 & ((1u << b0) - 1u))].w;
    if (NTAB == 3) G = mod.mul(G, tm[ti((e0 >> b0) & ((1u << b1) - 1u))]);
    G = mod.mul(G, th[ti(e0 >> b01)]);
    G = mod.mul(G, TwP<V>{(V)J.scale_w, (V)J.scale_wp});
    // LDS byte offset of (row k, column c) is linear in the bit fields of k
    auto slot = [&](uint32_t k, uint32_t c) -> uint32_t { return swz((bitrev_bits(k, log_r) << log_c) | c) * 4u; };
    const uint32_t l_slot = slot(4u * m, cl);
    uint32_t jx[4];
#pragma unroll
    for (int j = 0; j < 4; j++) jx[j] = slot((uint32_t)j, 0);
    const uint32_t o_lbits = J.o_lbits, omask = (1u << o_lbits) - 1u;
    Sout* __restrict__ out = reinterpret_cast<Sout*>(J.out) + (int64_t)poly * J.out_pstride +
                             ((int64_t)(col0 >> o_lbits) << J.sh_ohs) + ((int64_t)(col0 & omask) << J.sh_oc) +
                             (((cl >> o_lbits) << J.sh_ohs) + ((cl & omask) << J.sh_oc) + ((4u * m) << J.sh_or));
    const uint32_t sweeps = 1u << (log_r - 2u - log_tpc);
    // (unrolled where the shape is fixed: straight-line stores let the
    // compiler count them when it waits for older loads)
#pragma unroll
    for (uint32_t u = 0; u < sweeps; u++) {
        const uint32_t su = l_slot ^ slot(u * kstep, 0);
        V x[4];
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = lds_ld<V>(su ^ jx[j]);
        x[1] = mod.mul_lazy(x[1], B1);  // mont32 takes any representative below 2^32
        x[2] = mod.mul_lazy(x[2], B2);
        x[3] = mod.mul_lazy(x[3], B3);
#pragma unroll
        for (int j = 0; j < 4; j++) x[j] = LZ ? mont32_lazy(x[j], G, q, qinv) : mont32(x[j], G, q, qinv);
        G = mod.mul(G, Wb);
        Sout v[4] = {(Sout)x[0], (Sout)x[1], (Sout)x[2], (Sout)x[3]};
        store4<Sout>(out + ((u * kstep) << J.sh_or), v);
    }
}

template <int PER, bool P4, bool SWT, int LR, int LC, class M, class Sout, int CF = -1, class TP>
__device__ __forceinline__ void store_dispatch(const PassJob& J, const M& mod, uint32_t poly, uint32_t col0, bool pred,
                                               TP tl, TP tm, TP th) {
    const uint32_t post = J.post;
    if (P4 && post == POST_FOURSTEP && SWT && (J.flags & JF_POST_ROW)) {
        if (pred) store_loop<PER, 2, 2, 3, true, SWT, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
        else if (J.flags & JF_LAZY_OUT) store_loop<PER, 2, 2, 3, false, SWT, LR, LC, M, Sout, true, CF>(J, mod, poly, col0, tl, tm, th);
        else store_loop<PER, 2, 2, 3, false, SWT, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
    } else if (std::is_same<M, Mod32>::value && PER == 4 && P4 && post == POST_FOURSTEP && SWT && !pred && CF != 1 &&
               (J.flags & JF_POST_STEP)) {
        if constexpr (std::is_same<M, Mod32>::value && PER == 4) {
            const bool lz = J.flags & JF_LAZY_OUT;
            if (J.mid_bits) {
                if (lz) store_loop_step<3, SWT, LR, LC, Sout, true>(J, mod, poly, col0, tl, tm, th);
                else store_loop_step<3, SWT, LR, LC, Sout, false>(J, mod, poly, col0, tl, tm, th);
            } else {
                if (lz) store_loop_step<2, SWT, LR, LC, Sout, true>(J, mod, poly, col0, tl, tm, th);
                else store_loop_step<2, SWT, LR, LC, Sout, false>(J, mod, poly, col0, tl, tm, th);
            }
        }
    } else if (P4 && post == POST_FOURSTEP) {
        const bool cf = CF < 0 ? (J.flags & JF_STORE_CFAST) != 0 : CF != 0;
        const int qm = PER == 4 ? (cf ? (J.e_shift >= 2 ? 1 : 0) : 2) : 0;
        const bool three = J.mid_bits != 0;
#define NTT_STORE4(NT, QM)                                                                                      \
    (pred ? store_loop<PER, 2, NT, QM, true, SWT, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th) \
          : store_loop<PER, 2, NT, QM, false, SWT, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th))
        if (three) {
            if (qm == 1) NTT_STORE4(3, 1);
            else if (qm == 2) NTT_STORE4(3, 2);
            else NTT_STORE4(3, 0);
        } else {
            if (qm == 1) NTT_STORE4(2, 1);
            else if (qm == 2) NTT_STORE4(2, 2);
            else NTT_STORE4(2, 0);
        }
#undef NTT_STORE4
    } else if (post == POST_SCALE) {
        if (pred) store_loop<PER, 1, 2, 0, true, false, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
        else store_loop<PER, 1, 2, 0, false, false, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
    } else {
        if (pred) store_loop<PER, 0, 2, 0, true, false, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
        else store_loop<PER, 0, 2, 0, false, false, LR, LC, M, Sout, false, CF>(J, mod, poly, col0, tl, tm, th);
    }
}

// Tile load into LDS (row-major, swizzled): PER elements per thread and
// sweep, PRED per-element bounds predicates, LR/LC as in store_loop; MUL
// stores in[i]*in2[i] mod q, or the inner product of J.ip_n such pairs
// (JF_MUL_IN).
//   pre() runs once, after the first batch's global loads are issued and before
//   their LDS stores, so that loads it issues share their latency.
template <int PER, bool PRED, int LR, int LC, class Sin, class M, bool MUL = false, int CF = -1, class Pre>
__device__ __forceinline__ void load_loop(const PassJob& J, uint32_t poly, uint32_t col0, Pre pre) {
    using V = typename M::V;
    constexpr uint32_t VB = sizeof(V);
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r, log_c = LR ? (uint32_t)LC : J.log_c;
    const uint32_t R = 1u << log_r, C = 1u << log_c;
    const uint32_t total = R << log_c;
    const uint32_t ncols = J.ncols;
    const uint32_t T = kNT, tid = ltid();
    const bool cf = CF < 0 ? (J.flags & JF_LOAD_CFAST) != 0 : CF != 0;
    const uint32_t sweep = PER * T;
    const uint32_t sh_sr = J.sh_sr, sh_sc = J.sh_sc;
    auto rc = [&](uint32_t i, uint32_t& r, uint32_t& c) {
        if (cf) { r = i >> log_c; c = i & (C - 1u); }
        else { c = i >> log_r; r = i & (R - 1u); }
    };
    // element offset of (r, c) from the tile's first column: < 2^29
    auto goff = [&](uint32_t r, uint32_t c) -> uint32_t { return (r << sh_sr) + (c << sh_sc); };
    auto slot = [&](uint32_t r, uint32_t c) -> uint32_t { return swz((r << log_c) | c) * VB; };
    const uint32_t li = tid * PER;
    uint32_t lr, lc;
    rc(li, lr, lc);
    const bool lane_ok = li < total;
    const uint32_t l_off = goff(lr, lc), l_slot = slot(lr, lc);
    uint32_t e_x[PER];  // byte XOR offsets of the elements of a quad
#pragma unroll
    for (int e = 0; e < PER; e++) e_x[e] = cf ? (uint32_t)e * VB : swz((uint32_t)e << log_c) * VB;
    const int64_t base = (int64_t)poly * J.in_pstride + ((int64_t)col0 << sh_sc);
    const Sin* __restrict__ in = reinterpret_cast<const Sin*>(J.in) + base;
    const Sin* __restrict__ in2 = MUL ? reinterpret_cast<const Sin*>(J.in2) + base : nullptr;
    auto fetch = [&](const Sin* __restrict__ p, bool ok, Sin* v) {
        if (PER == 4) {
            if (ok) load4<Sin>(p, v);
            else {
#pragma unroll
                for (int e = 0; e < PER; e++) v[e] = 0;
            }
        } else {
            v[0] = ok ? gld(p) : (Sin)0;
        }
    };
    const LoadMul<M> mul(MUL ? J.q : 3u, J.mul_r2);
    constexpr int UN = kUnroll * 4 / (int)(sizeof(Sin) * (MUL ? 2 : 1));
    for (uint32_t bb = 0; bb < total; bb += UN * sweep) {
        Sin v[UN][PER];
        Sin w[MUL ? UN : 1][PER];
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const uint32_t ui = bb + (uint32_t)u * sweep;
            if (ui >= total) break;
            uint32_t ur, uc;
            rc(ui, ur, uc);
            const uint32_t off = goff(ur, uc) + l_off;
            const bool ok = !PRED || (lane_ok && col0 + (uc | lc) < ncols);
            fetch(in + off, ok, v[u]);
            if (MUL) fetch(in2 + off, ok, w[MUL ? u : 0]);
        }
        if (bb == 0) pre();
#pragma unroll
        for (int u = 0; u < UN; u++) {
            const uint32_t ui = bb + (uint32_t)u * sweep;
            if (ui >= total) break;
            if (PRED && !lane_ok) continue;
            uint32_t ur, uc;
            rc(ui, ur, uc);
            const uint32_t s = l_slot ^ slot(ur, uc);
#pragma unroll
            for (int e = 0; e < PER; e++) {
                V x = (V)v[u][e];
                if (MUL) x = mul(x, (V)w[MUL ? u : 0][e]);
                lds_st<V>(s ^ e_x[e], x);
            }
        }
    }
    // An inner product's further pairs (k >= 1) add to the slots this thread
    // has just written (no barrier: every slot has one writer): two sweeps and
    // up to KP pairs per step, all their loads issued before the first is
    // used, few enough registers that the loop above stays the kernel's peak.
    // Inner products run in the generic instances (LR = 0) and in the IP
    // instances of run_pass; the other shape-specialised loads (uint64 first
    // passes) take single products only.
    if (MUL && (LR == 0 || sizeof(Sin) == 4) && J.ip_n > 1) {
        constexpr uint32_t KP = sizeof(Sin) == 4 ? 2u : 1u;
        for (uint32_t k = 1; k < J.ip_n; k += KP) {
            const uint32_t kn = J.ip_n - k < KP ? J.ip_n - k : KP;
#pragma unroll 1
            for (uint32_t bb = 0; bb < total; bb += 2u * sweep) {
                Sin av[KP][2][PER], bv[KP][2][PER];
#pragma unroll
                for (uint32_t p = 0; p < KP; p++) {
                    if (p >= kn) break;
                    const int64_t ko = (int64_t)(k + p) * J.ip_stride;
#pragma unroll
                    for (int u = 0; u < 2; u++) {
                        const uint32_t ui = bb + (uint32_t)u * sweep;
                        if (ui >= total) break;
                        uint32_t ur, uc;
                        rc(ui, ur, uc);
                        const uint32_t off = goff(ur, uc) + l_off;
                        const bool ok = !PRED || (lane_ok && col0 + (uc | lc) < ncols);
                        fetch(in + ko + off, ok, av[p][u]);
                        fetch(in2 + ko + off, ok, bv[p][u]);
                    }
                }
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const uint32_t ui = bb + (uint32_t)u * sweep;
                    if (ui >= total) break;
                    if (PRED && !lane_ok) continue;
                    uint32_t ur, uc;
                    rc(ui, ur, uc);
                    const uint32_t s = l_slot ^ slot(ur, uc);
#pragma unroll
                    for (int e = 0; e < PER; e++) {
                        V x = lds_ld<V>(s ^ e_x[e]);
#pragma unroll
                        for (uint32_t p = 0; p < KP; p++)
                            if (p < kn) x = mul.add(x, mul((V)av[p][u][e], (V)bv[p][u][e]));
                        lds_st<V>(s ^ e_x[e], x);
                    }
                }
            }
        }
    }
}

// LDS behind the tile: the CT table (R entries) or an MFMA pass's round
// twiddles, then the post-multiply tables (lo, mid, hi) or the tile's row
// table.  Offsets depend on the job only.
template <class V>
struct Staged {
    LdsTw<V> zl, pl, pm, ph;
};

template <class M, bool P4, int LR, bool MF>
__device__ __forceinline__ Staged<typename M::V> staged_of(const PassJob& J) {
    using V = typename M::V;
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r;
    constexpr bool mf = MF && std::is_same<M, Mod32>::value;
    uint32_t after = (1u << (log_r + J.log_c)) * (uint32_t)sizeof(V);
    Staged<V> t;
    t.zl = LdsTw<V>{after};
    if (J.flags & JF_Z_LDS) after += (mf ? mf_tw_entries(log_r) : (1u << log_r)) * (uint32_t)sizeof(TwP<V>);
    t.pl = LdsTw<V>{after};
    t.pm = t.pl + (1u << J.lo_bits);
    t.ph = t.pm + (J.mid_bits ? 1u << J.mid_bits : 0u);
    return t;
}

// Copies the job's tables into LDS: everything (job tables and the row
// table of column col0), or (row_only) just the row table.
template <class M, bool P4, int LR, bool MF>
__device__ __forceinline__ void stage_tables(const PassJob& J, const Staged<typename M::V>& t, uint32_t col0,
                                             bool row_only) {
    using V = typename M::V;
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r;
    const uint32_t R = 1u << log_r;
    const uint32_t T = kNT, tid = ltid();
    const uint32_t flags = J.flags;
    if (!row_only && (flags & JF_Z_LDS)) {
        constexpr bool mf = MF && std::is_same<M, Mod32>::value;
        const TwP<V>* zsrc = mf ? reinterpret_cast<const TwP<V>*>(reinterpret_cast<const char*>(J.mz) + kMfTwOff)
                                : reinterpret_cast<const TwP<V>*>(J.z);
        const uint32_t nz = mf ? mf_tw_entries(log_r) : R;
        for (uint32_t i = tid; i < nz; i += T) t.zl.set(i, gld(zsrc + i));
    }
    if (P4 && (flags & JF_POST_ROW)) {
        const TwP<V>* gr = reinterpret_cast<const TwP<V>*>(J.tw_row) + ((size_t)(col0 >> J.row_shift) << log_r);
        for (uint32_t i = tid; i < R; i += T) t.pl.set(i, gld(gr + i));
    } else if (P4 && !row_only && (flags & JF_POST_LDS)) {
        const uint32_t n_lo = 1u << J.lo_bits, n_mid = J.mid_bits ? 1u << J.mid_bits : 0u;
        const uint32_t n_hi = (J.e_mask >> (J.lo_bits + J.mid_bits)) + 1u;
        const TwP<V>* gl = reinterpret_cast<const TwP<V>*>(J.tw_lo);
        const TwP<V>* gm = reinterpret_cast<const TwP<V>*>(J.tw_mid);
        const TwP<V>* gh = reinterpret_cast<const TwP<V>*>(J.tw_hi);
        for (uint32_t i = tid; i < n_lo; i += T) t.pl.set(tsw(i), gld(gl + i));
        for (uint32_t i = tid; i < n_mid; i += T) t.pm.set(tsw(i), gld(gm + i));
        for (uint32_t i = tid; i < n_hi; i += T) t.ph.set(tsw(i), gld(gh + i));
    }
}

// Store kinds a pipelined instance fixes at compile time (pipe_store_kind): a
// tile loop with a single store loop lets the compiler count the stores that
// follow the next tile's loads, so the wait for those loads leaves the stores
// in flight.  With several store loops under branches it takes the path with
// the fewest stores and waits for all of them.
enum PipeStore : int { PS_ANY = 0, PS_NONE = 1, PS_ROW = 2, PS_ROW_LZ = 3, PS_STEP3_LZ = 4 };

// The Cooley-Tukey levels of the tile in LDS (VALU rounds, or with MF radix-16
// rounds on the matrix cores), then the store with the pass's post-multiply.
template <class M, class Sout, int MAXM, bool P4, int LR, int LC, bool MF, int PS = PS_ANY, bool PK = false>
__device__ __forceinline__ void ct_and_store(const PassJob& J, const Staged<typename M::V>& t, uint32_t poly,
                                             uint32_t col0, bool full, const MfMat& mat,
                                             const MfMat* mat0 = nullptr) {
    using V = typename M::V;
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r, log_c = LR ? (uint32_t)LC : J.log_c;
    const uint32_t total = 1u << (log_r + log_c);
    const uint32_t T = kNT;
    const M mod{(decltype(M::q))J.q};
    const uint32_t flags = J.flags;
    const TwP<V>* zg = reinterpret_cast<const TwP<V>*>(J.z);
    const bool cyc = flags & JF_CT_CYC;
    if constexpr (MF && std::is_same<M, Mod32>::value) {
        const TwP<V>* twg = reinterpret_cast<const TwP<V>*>(reinterpret_cast<const char*>(J.mz) + kMfTwOff);
        const bool canon = J.post == POST_NONE;
        if (flags & JF_Z_LDS) mf_levels<PK>(mod, zg, mat, t.zl, log_r, log_c, cyc, canon, mat0);
        else mf_levels<PK>(mod, zg, mat, twg, log_r, log_c, cyc, canon, mat0);
    } else if (LR) {
        if (flags & JF_Z_LDS) ct_levels_c<LR, LC, MAXM, 0, M>(t.zl, mod, cyc);
        else ct_levels_c<LR, LC, MAXM, 0, M>(zg, mod, cyc);
    } else {
        if (flags & JF_Z_LDS) ct_levels<MAXM, M>(t.zl, mod, log_r, log_c, cyc);
        else ct_levels<MAXM, M>(zg, mod, log_r, log_c, cyc);
    }

    const bool pred_s = !full || total < ((flags & JF_VEC4_ST) ? 4u : 1u) * T;
    const TwP<V>* tl = reinterpret_cast<const TwP<V>*>(J.tw_lo);
    const TwP<V>* tm = reinterpret_cast<const TwP<V>*>(J.tw_mid);
    const TwP<V>* th = reinterpret_cast<const TwP<V>*>(J.tw_hi);
    if (LR) {  // shape-specialised: vector stores, full tiles, LDS post tables, fixed direction
        // (the branches of store_dispatch these kinds take; every kind but
        // PS_ANY has one store direction: column-fast, row-fast for the step)
        if constexpr (PS == PS_ROW || PS == PS_ROW_LZ) {
            store_loop<4, 2, 2, 3, false, true, LR, LC, M, Sout, PS == PS_ROW_LZ, 1>(J, mod, poly, col0, t.pl, t.pm, t.ph);
        } else if constexpr (PS == PS_STEP3_LZ && std::is_same<M, Mod32>::value) {
            store_loop_step<3, true, LR, LC, Sout, true>(J, mod, poly, col0, t.pl, t.pm, t.ph);
        } else if constexpr (PS == PS_NONE) {
            store_loop<4, 0, 2, 0, false, false, LR, LC, M, Sout, false, 1>(J, mod, poly, col0, t.pl, t.pm, t.ph);
        } else {
            if (flags & JF_STORE_CFAST) store_dispatch<4, P4, true, LR, LC, M, Sout, 1>(J, mod, poly, col0, false, t.pl, t.pm, t.ph);
            else store_dispatch<4, P4, true, LR, LC, M, Sout, 0>(J, mod, poly, col0, false, t.pl, t.pm, t.ph);
        }
    } else if (P4 && (flags & JF_POST_LDS)) {
        if (flags & JF_VEC4_ST) store_dispatch<4, P4, true, LR, LC, M, Sout>(J, mod, poly, col0, pred_s, t.pl, t.pm, t.ph);
        else store_dispatch<1, P4, true, LR, LC, M, Sout>(J, mod, poly, col0, pred_s, t.pl, t.pm, t.ph);
    } else {
        if (flags & JF_VEC4_ST) store_dispatch<4, P4, false, LR, LC, M, Sout>(J, mod, poly, col0, pred_s, tl, tm, th);
        else store_dispatch<1, P4, false, LR, LC, M, Sout>(J, mod, poly, col0, pred_s, tl, tm, th);
    }
}

// IP: an instance for jobs whose load forms a product or inner product
// (JF_MUL_IN) on 32-bit shape-specialised tiles; kept apart from the plain
// instances, whose registers the inner-product load would raise.
template <class M, class Sin, class Sout, int MAXM, bool P4, int LR, int LC, bool MF, bool IP = false>
__device__ __forceinline__ void run_pass(const PassJob& J, uint32_t tile_in_job) {
    using V = typename M::V;
    const uint32_t poly = tile_in_job / J.tiles_per_poly;
    const uint32_t tile = tile_in_job - poly * J.tiles_per_poly;
    const uint32_t log_r = LR ? (uint32_t)LR : J.log_r, log_c = LR ? (uint32_t)LC : J.log_c;
    const uint32_t C = 1u << log_c;
    const uint32_t col0 = tile << log_c;
    const uint32_t total = 1u << (log_r + log_c);
    const bool full = col0 + C <= J.ncols;  // every column of the tile exists
    const uint32_t flags = J.flags;
    const Staged<V> t = staged_of<M, P4, LR, MF>(J);
    // The job's tables and matrix are fetched while the tile's first global
    // loads are in flight (load_loop's pre), so a tile waits for one memory
    // round trip before its LDS writes instead of two, and the matrix-core
    // rounds none after the barrier.
    MfMat mat{};
    auto pre = [&] {
        stage_tables<M, P4, LR, MF>(J, t, col0, false);
        if constexpr (MF && std::is_same<M, Mod32>::value) mf_load_mat(J.mz, mat);
    };

    // ---- load ----
    // Predicates are needed for partial tiles and for tiles smaller than one sweep.
    const bool pred_l = !full || total < ((flags & JF_VEC4) ? 4u : 1u) * kNT;
    if (LR) {  // shape-specialised launches are full tiles with vector loads
        // (the fused product's first pass exists only for uint64 data)
        if constexpr (std::is_same<Sin, uint64_t>::value) {
            if (flags & JF_MUL_IN) load_loop<4, false, LR, LC, Sin, M, true>(J, poly, col0, pre);
            else load_loop<4, false, LR, LC, Sin, M>(J, poly, col0, pre);
        } else if constexpr (IP) {
            load_loop<4, false, LR, LC, Sin, M, true>(J, poly, col0, pre);
        } else {
            if (flags & JF_LOAD_CFAST) load_loop<4, false, LR, LC, Sin, M, false, 1>(J, poly, col0, pre);
            else load_loop<4, false, LR, LC, Sin, M, false, 0>(J, poly, col0, pre);
        }
    } else if (flags & JF_MUL_IN) {  // pointwise product fused into the load (predicated: rare path)
        if (flags & JF_VEC4) load_loop<4, true, LR, LC, Sin, M, true>(J, poly, col0, pre);
        else load_loop<1, true, LR, LC, Sin, M, true>(J, poly, col0, pre);
    } else if (flags & JF_VEC4) {
        if (pred_l) load_loop<4, true, LR, LC, Sin, M>(J, poly, col0, pre);
        else load_loop<4, false, LR, LC, Sin, M>(J, poly, col0, pre);
    } else {
        if (pred_l) load_loop<1, true, LR, LC, Sin, M>(J, poly, col0, pre);
        else load_loop<1, false, LR, LC, Sin, M>(J, poly, col0, pre);
    }
    lds_barrier();
    ct_and_store<M, Sout, MAXM, P4, LR, LC, MF>(J, t, poly, col0, full, mat);
}

// ---- one pass over many tiles per workgroup, the next tile's loads in flight ----
//
// pipe_kernel runs a single job's full, shape-specialised tiles (2^LR rows x
// 2^LC columns, vector loads and stores, tables in LDS).  Each workgroup walks
// a range of tiles and issues the global loads of its next tile into
// registers (tile_issue: 2^(LR+LC) / (4 * threads) 16-byte loads per thread)
// before it transforms and stores the current one, then writes them to LDS
// (tile_commit) once the current tile's store has read the buffer.  The
// rounds and the store touch only LDS (JF_Z_LDS, post tables in LDS) and
// issue no global loads the compiler would have to wait behind the prefetch
// (the first pass's stepped store reads its column table once per column).
template <int LR, int LC, class Sin>
struct TileLoad {
    static constexpr uint32_t kN = (1u << (LR + LC)) / (4u * kNT);
    static_assert(kN >= 1, "tile smaller than one vector load per thread");
    Sin v[kN][4];
};

template <int LR, int LC, class Sin, bool CF>
__device__ __forceinline__ void tile_issue(const PassJob& J, uint32_t poly, uint32_t col0, TileLoad<LR, LC, Sin>& L) {
    constexpr uint32_t R = 1u << LR, C = 1u << LC, sweep = 4u * kNT;
    constexpr bool cf = CF;
    const uint32_t sh_sr = J.sh_sr, sh_sc = J.sh_sc;
    auto rc = [&](uint32_t i, uint32_t& r, uint32_t& c) {
        if (cf) { r = i >> LC; c = i & (C - 1u); }
        else { c = i >> LR; r = i & (R - 1u); }
    };
    uint32_t lr, lc;
    rc(ltid() * 4u, lr, lc);
    const uint32_t l_off = (lr << sh_sr) + (lc << sh_sc);
    const Sin* __restrict__ in =
        reinterpret_cast<const Sin*>(J.in) + (int64_t)poly * J.in_pstride + ((int64_t)col0 << sh_sc);
#pragma unroll
    for (uint32_t u = 0; u < TileLoad<LR, LC, Sin>::kN; u++) {
        uint32_t ur, uc;
        rc(u * sweep, ur, uc);
        load4<Sin>(in + (((ur << sh_sr) + (uc << sh_sc)) + l_off), L.v[u]);
    }
}

template <int LR, int LC, class Sin, class M, bool CF>
__device__ __forceinline__ void tile_commit(const PassJob& J, const TileLoad<LR, LC, Sin>& L) {
    using V = typename M::V;
    constexpr uint32_t R = 1u << LR, C = 1u << LC, sweep = 4u * kNT, VB = sizeof(V);
    constexpr bool cf = CF;
    auto rc = [&](uint32_t i, uint32_t& r, uint32_t& c) {
        if (cf) { r = i >> LC; c = i & (C - 1u); }
        else { c = i >> LR; r = i & (R - 1u); }
    };
    auto slot = [&](uint32_t r, uint32_t c) -> uint32_t { return swz((r << LC) | c) * VB; };
    uint32_t lr, lc;
    rc(ltid() * 4u, lr, lc);
    const uint32_t l_slot = slot(lr, lc);
    uint32_t e_x[4];
#pragma unroll
    for (int e = 0; e < 4; e++) e_x[e] = cf ? (uint32_t)e * VB : swz((uint32_t)e << LC) * VB;
#pragma unroll
    for (uint32_t u = 0; u < TileLoad<LR, LC, Sin>::kN; u++) {
        uint32_t ur, uc;
        rc(u * sweep, ur, uc);
        const uint32_t s = l_slot ^ slot(ur, uc);
#pragma unroll
        for (int e = 0; e < 4; e++) lds_st<V>(s ^ e_x[e], (V)L.v[u][e]);
    }
}

// Tiles are split into 8 contiguous ranges, range x walked by the workgroups
// b = x mod 8 (so that a range's tiles mostly share one XCD's L2, as in
// xcd_remap); the host launches a multiple of 8 workgroups on a tile count
// divisible by 8.  At most 6 waves per SIMD: the prefetch registers must not
// cost the third workgroup per CU.
template <class M, class Sin, class Sout, int MAXM, bool P4, int LR, int LC, bool MF, bool LCF, int PS>
__device__ __forceinline__ void pipe_run(const PassJob& J) {
    using V = typename M::V;
    const uint32_t T = J.tiles_per_poly * (J.npoly ? J.npoly : 1u);
    const uint32_t x = blockIdx.x & 7u, per = T >> 3, gs = gridDim.x >> 3;
    uint32_t i = blockIdx.x >> 3;
    if (i >= per) return;
    const uint32_t tpp = J.tiles_per_poly;
    auto tile_of = [&](uint32_t k, uint32_t& poly, uint32_t& col0) {
        const uint32_t g = x * per + k;
        poly = g / tpp;
        col0 = (g - poly * tpp) << LC;
    };
    const Staged<V> t = staged_of<M, P4, LR, MF>(J);
    MfMat mat{};
    // A remainder round of 2 or 3 levels runs packed on the matrix cores
    // where the registers allow a second matrix (2^14-word tiles, 4 waves
    // per SIMD); the one-tile kernels keep it on the VALU.
    constexpr bool kPack = MF && std::is_same<M, Mod32>::value && (LR & 3) >= 2 && LR + LC >= 14;
    MfMat mat0{};
    TileLoad<LR, LC, Sin> L;
    uint32_t poly, col0;
    tile_of(i, poly, col0);
    tile_issue<LR, LC, Sin, LCF>(J, poly, col0, L);
    stage_tables<M, P4, LR, MF>(J, t, col0, false);
    if constexpr (MF && std::is_same<M, Mod32>::value) mf_load_mat(J.mz, mat);
    if constexpr (kPack) mf_load_mat(J.mz, mat0, mf_pack_off(LR));
    tile_commit<LR, LC, Sin, M, LCF>(J, L);
    // Every load before the loop (matrix fragments, tables) has landed.  The
    // loop's first use of the fragments would otherwise carry a wait of its
    // own, counted past the next tile's loads issued just before it, and the
    // first round of every tile would stall until the prefetch arrived.
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
    lds_barrier();
    const bool row_post = P4 && (J.flags & JF_POST_ROW);
    // A row table of at most one entry per thread comes in with the next
    // tile's loads (one 8-byte load, before the current tile's stores), so no
    // load issued after those stores has to be waited for before the commit.
    constexpr bool kRowPre = (1u << LR) <= kNT;
    const TwP<V>* rows = reinterpret_cast<const TwP<V>*>(J.tw_row);
    TwP<V> rowp{};
    for (;;) {
        const uint32_t ni = i + gs;
        const bool more = ni < per;
        uint32_t npl = 0, ncol0 = 0;
        bool restage = false;
        if (more) {
            tile_of(ni, npl, ncol0);
            tile_issue<LR, LC, Sin, LCF>(J, npl, ncol0, L);
            restage = row_post && (ncol0 >> J.row_shift) != (col0 >> J.row_shift);
            if (kRowPre && restage && ltid() < (1u << LR))
                rowp = gld(rows + ((size_t)(ncol0 >> J.row_shift) << LR) + ltid());
        }
        ct_and_store<M, Sout, MAXM, P4, LR, LC, MF, PS, kPack>(J, t, poly, col0, true, mat, &mat0);
        if (!more) break;
        lds_barrier();  // the store has read the tile (and the row table)
        if (restage) {
            if (kRowPre) {
                if (ltid() < (1u << LR)) t.pl.set(ltid(), rowp);
            } else {
                stage_tables<M, P4, LR, MF>(J, t, ncol0, true);
            }
        }
        tile_commit<LR, LC, Sin, M, LCF>(J, L);
        lds_barrier();
        i = ni;
        poly = npl;
        col0 = ncol0;
    }
}

// Tiles of 2^14 words (64 KiB) leave room for two workgroups per CU (4 waves
// per SIMD), so their instances may hold the 32 registers of the next tile's
// loads within 128 VGPRs; smaller tiles run three per CU (6 waves, 80 VGPRs).
template <int LR, int LC>
constexpr int pipe_waves() { return LR + LC >= 14 ? 4 : 6; }

// The store kind of a pipelined job (PipeStore), as store_dispatch would
// choose it for a full, shape-specialised tile; PS_ANY keeps the branches.
template <bool P4>
__device__ __forceinline__ int pipe_store_kind(const PassJob& J) {
    const bool cf = J.flags & JF_STORE_CFAST;
    if (!P4) return J.post == POST_NONE && cf ? PS_NONE : PS_ANY;
    if (J.post != POST_FOURSTEP) return PS_ANY;
    if (J.flags & JF_POST_ROW) return !cf ? PS_ANY : (J.flags & JF_LAZY_OUT) ? PS_ROW_LZ : PS_ROW;
    if ((J.flags & JF_POST_STEP) && J.mid_bits && (J.flags & JF_LAZY_OUT) && !(J.flags & JF_STORE_CFAST))
        return PS_STEP3_LZ;
    return PS_ANY;
}

template <class M, class Sin, class Sout, int MAXM, int NT, bool P4, int LR, int LC, bool MF>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(pipe_waves<LR, LC>()))) pipe_kernel(PassJob J) {
    if (!(J.flags & JF_LOAD_CFAST)) {
        pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, false, PS_ANY>(J);
        return;
    }
    switch (pipe_store_kind<P4>(J)) {
        case PS_NONE:
            if constexpr (!P4) pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, true, PS_NONE>(J);
            break;
        case PS_ROW:
            if constexpr (P4) pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, true, PS_ROW>(J);
            break;
        case PS_ROW_LZ:
            if constexpr (P4) pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, true, PS_ROW_LZ>(J);
            break;
        case PS_STEP3_LZ:
            if constexpr (P4) pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, true, PS_STEP3_LZ>(J);
            break;
        default: pipe_run<M, Sin, Sout, MAXM, P4, LR, LC, MF, true, PS_ANY>(J); break;
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
__device__ __forceinline__ uint32_t find_job(const PassJob* __restrict__ jobs, uint32_t njobs, uint32_t tile) {
    uint32_t lo = 0, hi = njobs;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (jobs[mid].tile_begin <= tile) lo = mid; else hi = mid;
    }
    return lo;
}

// A launch runs either the single job passed by value (njobs == 0) or one
// job per tile range from a device array sorted by tile_begin; one tile per
// workgroup.
// LR > 0: every job of the launch has tiles of 2^LR x 2^LC, vector loads and
// stores, full tiles and (with P4) LDS post tables (pass_shape_of in ntt_lib.hip).
// MF: every job has JF_MFMA (32-bit arithmetic); without it none has.  Each
// instance holds only one kind of round, so its registers are those of that
// kind.  (Forcing 8 waves per SIMD on the matrix-core shapes, 4 workgroups
// per CU, fits without scratch but measured ~2% slower at 2^24.)
template <class M, class Sin, class Sout, int MAXM, int NT, bool P4, int LR, int LC, bool MF, bool IP = false>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(M::kMinWaves))) pass_kernel(const PassJob* __restrict__ jobs, uint32_t njobs, PassJob inl) {
    // The dynamic LDS (no static LDS: it starts at address 0) is addressed by
    // byte offset through lds_ld / lds_st.
    const uint32_t tile = xcd_remap(blockIdx.x, gridDim.x);
    if (njobs == 0) {
        run_pass<M, Sin, Sout, MAXM, P4, LR, LC, MF, IP>(inl, tile);
        return;
    }
    const PassJob J = jobs[find_job(jobs, njobs, tile)];
    run_pass<M, Sin, Sout, MAXM, P4, LR, LC, MF, IP>(J, tile - J.tile_begin);
}

// ---- a whole multi-pass batch in one persistent launch ----
//
// The jobs (one per polynomial and pass) are numbered in one tile space and
// split into nq queues, queue x holding the tiles [qbeg[x], qbeg[x+1]) in an
// order where every job comes after the job it depends on (tjob[g]: the job
// of tile g).  A workgroup takes tiles of its queue by atomic tickets; before
// a tile of a dependent job its thread 0 waits until the completion counter of
// the job it depends on shows all of that job's tiles done in this launch.
// Tickets go out in queue order to running workgroups only, so every wait is
// for tiles that running workgroups hold and the launch cannot deadlock,
// whatever the grid size or the dispatch order; each wait is also bounded
// (spin_limit polls; status bit 0 records one that gave up, whose tile then
// ran without its inputs complete).
//
// Completion counters are 64-bit and never reset: job j's counter has gained
// tiles(j) in every earlier launch, so a tile of a counted job computes the
// launch's generation G = counter(j) / tiles(j) (its own tile is not counted
// yet) and waits for counter(dep) >= (G + 1) * tiles(dep).  Tickets are
// reset by the last workgroup to leave (a relaxed exit counter; every
// workgroup's ticket atomics have returned before it counts itself out).
// Launches of one batch must not overlap: they share the counters.
//
// Two hand-over protocols (QueueProto):
//   QP_AGENT: queue = blockIdx mod nq, and a workgroup whose queue is empty
//     takes tiles of the next one.  The grid-barrier pattern at agent scope:
//     the producer's waves finish their stores, the workgroup barrier orders
//     them before thread 0's agent-scope fence and counter increment; the
//     consumer's thread 0 sees the count and fences, and the barrier orders
//     the workgroup's loads after it.  Correct on any dispatch, but each
//     fence writes back and invalidates the XCD's L2.
//   QP_XCD: queue = the XCD the workgroup runs on (hardware register
//     XCC_ID; nq = 8), and a workgroup only takes tiles of its own queue, so
//     all passes of a polynomial run on one XCD and its data moves through
//     that XCD's L2, which all its CUs share.  The producer's waves wait until
//     their stores are acknowledged (s_waitcnt vmcnt(0): the data is in the
//     L2) before the barrier and the counter increment; the consumer
//     invalidates its CU's L1 after the wait (buffer_inv sc0), so its loads
//     read the L2.  The host uses it after a probe launch has seen workgroups
//     on all 8 XCDs; a queue no workgroup drained is reported (status bit 1).
//
// ctl (64-bit words, each on its own 256-byte line so that polling and
// counting on one line do not serialise): status, exit count, tickets[nq],
// then the completion counters.
// tslot: LDS byte offset of two words past every job's tile and tables.
enum QueueProto : uint32_t { QP_AGENT = 0, QP_XCD = 1 };
constexpr uint32_t kQLine = 32;  // 64-bit words per 256-byte line
constexpr uint32_t kQStatus = 0, kQExits = kQLine;
__host__ __device__ constexpr uint32_t q_ticket(uint32_t x) { return (2u + x) * kQLine; }
__host__ __device__ constexpr uint32_t q_done(uint32_t j) { return (10u + j) * kQLine; }

__device__ __forceinline__ uint32_t hw_xcc_id() { return __builtin_amdgcn_s_getreg(6164) & 15u; }  // hwreg(HW_REG_XCC_ID, 0, 4)

template <class M, class Sin, class Sout, int MAXM, int NT, bool P4, bool MF>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(M::kMinWaves > 6 ? M::kMinWaves : 6)))
queue_kernel(const PassJob* __restrict__ jobs, const uint32_t* __restrict__ tjob, const uint32_t* __restrict__ qbeg,
             uint32_t nq, unsigned long long* ctl, uint32_t tslot, uint32_t spin_limit, uint32_t proto) {
    const uint32_t tid = threadIdx.x;
    const bool xcd = proto == QP_XCD;
    unsigned int* status = reinterpret_cast<unsigned int*>(ctl + kQStatus);
    // thread 0's queue cursor; QP_XCD never leaves its own queue
    uint32_t x = (xcd ? hw_xcc_id() : blockIdx.x) % nq, tried = xcd ? nq - 1u : 0u;
    for (;;) {
        if (tid == 0) {
            uint32_t g = 0xFFFFFFFFu, j = 0;
            while (tried < nq) {
                const uint32_t k = (uint32_t)__hip_atomic_fetch_add(ctl + q_ticket(x), 1ull, __ATOMIC_RELAXED,
                                                                    __HIP_MEMORY_SCOPE_AGENT);
                if (k < qbeg[x + 1] - qbeg[x]) {
                    g = qbeg[x] + k;
                    break;
                }
                x = x + 1 == nq ? 0 : x + 1;
                tried++;
            }
            if (g != 0xFFFFFFFFu) {
                j = tjob[g];
                const PassJob& J = jobs[j];
                if (J.dep_tiles) {
                    const unsigned long long gen =
                        __hip_atomic_load(ctl + q_done(j), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) /
                        (J.tiles_per_poly * (J.npoly ? J.npoly : 1u));
                    const unsigned long long need = (gen + 1) * J.dep_tiles;
                    const unsigned long long* c = ctl + q_done(J.dep);
                    uint32_t polls = 0;
                    while (__hip_atomic_load(c, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < need) {
                        __builtin_amdgcn_s_sleep(4);
                        if (++polls >= spin_limit) {
                            __hip_atomic_fetch_or(status, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                            break;
                        }
                    }
                    if (xcd) __asm__ volatile("buffer_inv sc0" ::: "memory");
                    else __threadfence();
                }
            }
            lds_st<uint32_t>(tslot, g);
            lds_st<uint32_t>(tslot + 4u, j);
        }
        __syncthreads();
        const uint32_t g = lds_ld<uint32_t>(tslot), j = lds_ld<uint32_t>(tslot + 4u);
        if (g == 0xFFFFFFFFu) break;
        const PassJob J = jobs[j];
        run_pass<M, Sin, Sout, MAXM, P4, 0, 0, MF>(J, g - J.tile_begin);
        if (J.counted) {
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): this wave's stores are acknowledged
            __syncthreads();
            if (tid == 0) {
                if (!xcd) __threadfence();
                __hip_atomic_fetch_add(ctl + q_done(j), 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        } else {
            __syncthreads();  // the tile's LDS is free again
        }
    }
    if (tid == 0) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // this wave's counter increments have landed
        const unsigned long long e =
            __hip_atomic_fetch_add(ctl + kQExits, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e + 1 == gridDim.x) {  // every other workgroup has taken its last ticket
            for (uint32_t q = 0; q < nq; q++) {
                const unsigned long long t =
                    __hip_atomic_load(ctl + q_ticket(q), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (t < qbeg[q + 1] - qbeg[q]) __hip_atomic_fetch_or(status, 2u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_store(ctl + q_ticket(q), 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __hip_atomic_store(ctl + kQExits, 0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
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

// 62-bit pointwise product r = a*b mod q: mont64(mont64(a, b), 2^128 mod q).
__global__ void pointwise_mul62_kernel(uint64_t* r, const uint64_t* a, const uint64_t* b, uint64_t n, uint64_t q,
                                       uint64_t qneg_inv, uint64_t r2);

// In-place bit reversal of n = 2^log_n uint64 (reference bitrev semantics).
__global__ void bitrev_kernel(uint64_t* d, uint64_t n, uint32_t log_n);

// out[b*N1 + a] = in[a*N2 + b]  (N1 x N2 row-major transpose), 32x32 LDS tiles.
__global__ void transpose_u64_kernel(uint64_t* out, const uint64_t* in, uint32_t n1, uint32_t n2);

}  // namespace nttk
