// ntt_mf16.hpp — radix-16 rounds of the pass kernel on the gfx950 INT8 matrix
// cores (v_mfma_i32_16x16x64_i8), for q < 2^31.
//
// The pass kernel's Cooley-Tukey phase runs log2 R levels over the LDS tile in
// rounds; a round of 4 levels maps every 16-element column unit (rows base +
// v*2^t, v < 16) to
//     y = F * diag(d_i0) * x  (mod q),
// where F[p][v] is the same 16x16 matrix in every round (the 4-level
// butterfly network of block 0 of the first round, with its twiddles) and
// d_i0 is a per-unit twiddle vector (i0 = the unit's block index).  The host
// derives F and d from the very butterfly network ct_unit<4> runs, so the
// output is the VALU round's, slot for slot.  A round here is therefore:
//   1. each lane reads 4 words of one unit (lane (g, n): unit n, v = 4g..4g+3),
//   2. multiplies them by d (lazy Shoup, any 32-bit representative; skipped
//      in a round at level 0, whose d is all ones),
//   3. flips every byte to a signed digit (XOR 0x80: byte - 128) — the four
//      words are the MFMA's B operand as they are: K slot (v, byte),
//   4. four MFMAs, one per balanced signed byte plane F_c of
//      G[p][(v, b)] = F[p][v] * 2^(8b) mod q (centred, G = sum_c 2^(8c) F_c),
//   5. recombines A_0 + 2^8 A_1 + 2^16 (A_2 + 2^8 A_3): the high half by one
//      signed Montgomery product with 2^16 * 2^32 (mod q), plus a constant
//      L0 that makes the sum positive (plane 0's initial values); the
//      byte-flip correction 128 * rowsum(G) is folded into plane 2's,
//   6. writes the 4 results back to the 4 words it read (the accumulator
//      layout puts output rows 4g..4g+3 of unit n in lane (g, n)).
// Bounds (q < 2^31): |F_c|, |digit| <= 128 over K = 64 give |A_c| <= 2^20,
// so lo = A_0 + 2^8 A_1 is within +-2^28.01, hi = A_2 + 2^8 A_3 + T_p within
// +-(2^28.01 + q/2), the Montgomery result r within +-0.66 q, and
// y = lo + r + L0 with L0 = 2^28 + 2^21 + 0.67 q lies in [0, 2^32): every
// intermediate is exact, and y is a representative of the output below 2^32
// (below 2q for q >= 2^29.6).
#pragma once

namespace nttk {

typedef int mf_v4i __attribute__((ext_vector_type(4)));

// Matrix block of one pass (host: build_mf16_block):
//   [plane c][lane] 16-byte A fragments   4 KiB
//   [lane] int4 plane-2 initial values    1 KiB   (rows 4g..4g+3)
//   header: int32 L0, cm = 2^48 mod q (centred), cm * q^-1 mod 2^32
//   then the round twiddle tables, TwP<uint32_t>: the round at level l0
//   holds d_i0[v] at [i0*16 + v], 2^(l0+4) entries, rounds in order.
constexpr uint32_t kMfFragBytes = 4u * 64u * 16u;
constexpr uint32_t kMfHdrOff = kMfFragBytes + 64u * 16u;
constexpr uint32_t kMfTwOff = kMfHdrOff + 256u;

struct MfMat {
    mf_v4i a[4];
    mf_v4i h;
    uint32_t l0, cm, cmq;
};

// The matrix at byte offset off of the block (0: F; mf_pack_off: the packed
// first round), with the block's shared header.
__device__ __forceinline__ void mf_load_mat(const void* blk, MfMat& M, uint32_t off = 0) {
    const uint32_t lane = threadIdx.x & 63u;
    const mf_v4i* f = reinterpret_cast<const mf_v4i*>(reinterpret_cast<const char*>(blk) + off);
#pragma unroll
    for (int c = 0; c < 4; c++) M.a[c] = gld(f + c * 64 + lane);
    M.h = gld(f + 256 + lane);
    const uint32_t* hd = reinterpret_cast<const uint32_t*>(reinterpret_cast<const char*>(blk) + kMfHdrOff);
    M.l0 = gld(hd);
    M.cm = gld(hd + 1);
    M.cmq = gld(hd + 2);
}

// Entries of the round twiddle tables of an R = 2^log_r pass (first MFMA
// round at level m0 = log_r mod 4).
__host__ __device__ inline uint32_t mf_tw_entries(uint32_t log_r) {
    uint32_t n = 0;
    for (uint32_t l0 = log_r & 3u; l0 + 4u <= log_r; l0 += 4u) n += 1u << (l0 + 4u);
    return n;
}

// A pass whose first m0 = 2 or 3 levels are a VALU remainder round can run
// that round on the matrix cores too: 16 >> m0 units of 2^m0 slots share an
// MFMA column under a block-diagonal matrix, stored after the round twiddle
// tables (fragments and plane-2 initial values, the header shared).  Slot w =
// unit h * 2^m0 + v of a column sits where a radix-16 unit at level 0 has its
// slot rot(w) = v * 2^(4 - m0) + h: the v stride is the round's, and the
// packed units are the neighbours just below it.
__host__ __device__ inline bool mf_has_pack(uint32_t log_r) { return (log_r & 3u) >= 2u && log_r >= 4u + (log_r & 3u); }
__host__ __device__ inline uint32_t mf_pack_off(uint32_t log_r) {
    return (kMfTwOff + mf_tw_entries(log_r) * 8u + 255u) & ~255u;
}
constexpr uint32_t kMfPackBytes = kMfFragBytes + 64u * 16u;

// One radix-16 round at level l0 over the whole tile (2^log_r rows, 2^log_c
// columns, row-major and swizzled at LDS byte offset 0).  TW: multiply the
// inputs by the round's twiddles tw[i0*16 + v].  CANON: 0 leaves any
// representative below 2^32 (the next round, or a post-multiply, takes it),
// 1 reduces by one conditional subtract (q >= 2^30), 2 by a Shoup reduction.
// PK (2 or 3): the packed first round of that many levels (l0 = 0, no
// twiddles, M the block-diagonal matrix); 0: a radix-16 round.
template <bool TW, int CANON, class TP, int PK = 0>
__device__ __forceinline__ void mf_round(const MfMat& M, TP tw, uint32_t q, uint32_t l0, uint32_t log_r,
                                         uint32_t log_c) {
    const uint32_t sh = log_r - l0 - 4u + log_c;  // log2 of the word stride of v within a unit
    const uint32_t ngroups = 1u << (log_r + log_c - 8u);
    const uint32_t tid = ltid();
    const uint32_t lane = tid & 63u, g = lane >> 4, n = lane & 15u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint32_t lowmask = (1u << sh) - 1u;
    // row-major word of unit u's element 0: u's low sh bits stay, the rest moves up by 4
    auto word = [&](uint32_t u) -> uint32_t { return (u & lowmask) | ((u & ~lowmask) << 4); };
    // Lane (g, n) of group ub reads unit ub | n, elements v = 4g + e.  Every map
    // here is GF(2)-linear and ub, n, g, e occupy disjoint bits, so an address
    // is (lane part la, one register for the round) ^ (element part se[e], a
    // scalar) ^ (group part U, a scalar): one vector XOR per access.
    // (The empty asm keeps the compiler from refactoring la ^ (se ^ U) back
    // into (word ^ word) * 4, more vector ops per access.)
    // slot w -> position in the 4-bit slot field (identity for radix 16)
    auto rot = [](uint32_t w) -> uint32_t {
        return PK ? (((w & ((1u << PK) - 1u)) << (4 - PK)) | (w >> PK)) : w;
    };
    uint32_t la = (swz(word(n)) ^ swz(rot(4u * g) << sh)) * 4u;
    __asm__("" : "+v"(la));
    uint32_t se[4];
#pragma unroll
    for (int e = 0; e < 4; e++) se[e] = __builtin_amdgcn_readfirstlane(swz(rot((uint32_t)e) << sh) * 4u);
    // twiddle index (((ub | n) >> sh) << 4) | 4g = lane part | (ub >> sh) << 4
    const uint32_t lti = ((n >> sh) << 4) | (4u * g);
    // with the table in LDS: byte address of the lane part
    uint32_t lta = 0;
    if constexpr (std::is_same<TP, LdsTw<uint32_t>>::value) {
        lta = tw.off + lti * (uint32_t)sizeof(TwP<uint32_t>);
        __asm__("" : "+v"(lta));
    }
    const uint32_t qr = CANON == 2 ? (uint32_t)(0xFFFFFFFFu / q) : 0u;  // floor((2^32 - 1) / q)
    const mf_v4i zero = {0, 0, 0, 0};
    // L0 enters through plane 0's accumulator: |sum| <= 2^20 and L0 < 2^31 - 2^20
    const mf_v4i l0v = {(int)M.l0, (int)M.l0, (int)M.l0, (int)M.l0};
    for (uint32_t grp = wave; grp < ngroups; grp += kNT / 64u) {
        const uint32_t ub = grp << 4;
        const uint32_t U = __builtin_amdgcn_readfirstlane(swz(word(ub)) * 4u);
        uint32_t x[4];
#pragma unroll
        for (int e = 0; e < 4; e++) x[e] = lds_ld<uint32_t>(la ^ (se[e] ^ U));
        if (TW) {
            const uint32_t tu = __builtin_amdgcn_readfirstlane((ub >> sh) << 4);
            TwP<uint32_t> t[4];
            if constexpr (std::is_same<TP, LdsTw<uint32_t>>::value) {
                const LdsTw<uint32_t> tl{lta + tu * (uint32_t)sizeof(TwP<uint32_t>)};
#pragma unroll
                for (int e = 0; e < 4; e++) t[e] = tl[(uint32_t)e];
            } else {
#pragma unroll
                for (int e = 0; e < 4; e++) t[e] = tw[lti + tu + (uint32_t)e];
            }
#pragma unroll
            for (int e = 0; e < 4; e++) x[e] = x[e] * t[e].w - __umulhi(x[e], t[e].wp) * q;  // [0, 2q)
        }
        mf_v4i b;
#pragma unroll
        for (int e = 0; e < 4; e++) b[e] = (int)(x[e] ^ 0x80808080u);
        const mf_v4i a0 = __builtin_amdgcn_mfma_i32_16x16x64_i8(M.a[0], b, l0v, 0, 0, 0);
        const mf_v4i a1 = __builtin_amdgcn_mfma_i32_16x16x64_i8(M.a[1], b, zero, 0, 0, 0);
        const mf_v4i a2 = __builtin_amdgcn_mfma_i32_16x16x64_i8(M.a[2], b, M.h, 0, 0, 0);
        const mf_v4i a3 = __builtin_amdgcn_mfma_i32_16x16x64_i8(M.a[3], b, zero, 0, 0, 0);
#pragma unroll
        for (int e = 0; e < 4; e++) {
            const uint32_t lo = (uint32_t)a0[e] + ((uint32_t)a1[e] << 8);  // includes L0
            const int32_t hi = (int32_t)((uint32_t)a2[e] + ((uint32_t)a3[e] << 8));
            const int32_t m = hi * (int32_t)M.cmq;
            const int32_t r = __mulhi(hi, (int32_t)M.cm) - __mulhi(m, (int32_t)q);  // hi * 2^16 mod q
            uint32_t y = lo + (uint32_t)r;
            if (CANON == 1) y = min(y, y - q);
            if (CANON == 2) {
                y = y - __umulhi(y, qr) * q;  // [0, 2q)
                y = min(y, y - q);
            }
            lds_st<uint32_t>(la ^ (se[e] ^ U), y);
        }
    }
}

// All log_r levels of the tile: a VALU round of log_r mod 4 levels first
// (merged CT table zg, as ct_levels), then radix-16 rounds on the matrix
// cores.  M: the pass's matrix (mf_load_mat); twl: its round twiddle tables (LDS or
// global); cyc: cyclic table, so the VALU round skips its products by 1; canon_out:
// the tile is stored without a post-multiply, so the outputs must be canonical.
// PK: the remainder round runs packed on the matrix cores with M0 (the
// block's mf_pack_off matrix; log_r mod 4 = 2 or 3).
template <bool PK = false, class ZP, class TP>
__device__ __forceinline__ void mf_levels(const Mod32& mod, ZP zg0, const MfMat& M, TP twl, uint32_t log_r,
                                          uint32_t log_c, bool cyc, bool canon_out, const MfMat* M0 = nullptr) {
    const uint32_t m0 = log_r & 3u;
    if (PK && m0 >= 2) {
        if (m0 == 3) mf_round<false, 0, TP, 3>(*M0, twl, mod.q, 0, log_r, log_c);
        else mf_round<false, 0, TP, 2>(*M0, twl, mod.q, 0, log_r, log_c);
        lds_barrier();
    } else if (m0) {
        // The VALU round reads at most 7 entries, the same in every thread: read
        // through the constant address space they are scalar loads, which no
        // wait for the vector loads in flight (a pipelined tile's) holds up.
        typedef const __attribute__((address_space(4))) TwP<uint32_t>* CZP;
        const CZP zg = (CZP)zg0;
        switch (m0) {
            case 1: cyc ? ct_round<1, true, Mod32>(zg, mod, 0, log_r, log_c) : ct_round<1, false, Mod32>(zg, mod, 0, log_r, log_c); break;
            case 2: cyc ? ct_round<2, true, Mod32>(zg, mod, 0, log_r, log_c) : ct_round<2, false, Mod32>(zg, mod, 0, log_r, log_c); break;
            default: cyc ? ct_round<3, true, Mod32>(zg, mod, 0, log_r, log_c) : ct_round<3, false, Mod32>(zg, mod, 0, log_r, log_c); break;
        }
        lds_barrier();
    }
    const uint32_t q = mod.q;
    // canonical output only where no post-multiply follows (which takes any
    // representative below 2^32)
    const int canon = !canon_out ? 0 : (q >= (1u << 30) ? 1 : 2);
    uint32_t off = 0;
    for (uint32_t l0 = m0; l0 + 4u <= log_r; l0 += 4u) {
        const bool last = l0 + 8u > log_r;
        // a first round at level 0 has all-one twiddles (the block's F is its
        // own matrix, negacyclic weights included)
        const bool tw = l0 > 0;
        const int cn = last ? canon : 0;
#define NTT_MF_ROUND(T, C) mf_round<T, C>(M, twl + off, q, l0, log_r, log_c)
        if (tw) {
            if (cn == 0) NTT_MF_ROUND(true, 0);
            else if (cn == 1) NTT_MF_ROUND(true, 1);
            else NTT_MF_ROUND(true, 2);
        } else {
            if (cn == 0) NTT_MF_ROUND(false, 0);
            else if (cn == 1) NTT_MF_ROUND(false, 1);
            else NTT_MF_ROUND(false, 2);
        }
#undef NTT_MF_ROUND
        off += 1u << (l0 + 4u);
        lds_barrier();
    }
}

}  // namespace nttk
