// ntt_mfma.hpp — R-point column transforms (R = 64 or 4096) as exact dense
// contractions on the gfx950 INT8 matrix cores (v_mfma_i32_16x16x64_i8).
//
// A 64-point transform Y = F X (mod q) of 16 column vectors is one 16x16x64
// MFMA per output tile and byte plane:
//   * B operand = the data's own uint32 words: the 4 bytes of element j sit in
//     K slots (j, byte); XOR 0x80808080 turns each byte x into the signed
//     digit x - 128.
//   * A operand = F' split into 4 balanced signed byte planes, where
//     F'[r][(j, b)] = F[r][j] * 2^(8b) mod q (centred), so that
//     sum_i 2^(8i) A_i == F X - 128*0x01010101*rowsum(F) (mod q); the
//     rowsum correction is preloaded, digit by digit, into the accumulators.
//   * |A_i| <= 256*128*128 = 2^22, so Y = (A0 + 2^8 A1) + 2^16 (A2 + 2^8 A3)
//     is exact in FP64 (|Y| < 2^47); Y mod q = Y - rint(Y/q) q by one FMA.
// R = 4096 = 64 x 64 is stage A (64-point over j1 of j = 64 j1 + j2), a twist
// w_R^(k1 j2) (times psi^(j2) for a negacyclic transform, whose psi^(64 j1)
// part is folded into stage A's matrix), and stage B (64-point over j2); the
// output k = k1 + 64 k2 is in natural order.  Which of this and the VALU
// Cooley-Tukey pass runs is chosen per plan by measured time.
#pragma once
#include "ntt_kernels.hpp"

namespace nttk {

typedef int v4i __attribute__((ext_vector_type(4)));

// Table block of one MFMA pass (built by the host, see build_mfma_tables):
constexpr uint32_t kMfmaFragBytes = 2u * 4u * 4u * 4u * 64u * 16u;  // [stage][plane][mtile][kstep][lane] int4
constexpr uint32_t kMfmaCorrOff = kMfmaFragBytes;                    // int32 [stage][row][plane]
constexpr uint32_t kMfmaTwistOff = kMfmaCorrOff + 2u * 64u * 4u * 4u;  // TwP<uint32_t>[k1 * 64 + j2]
constexpr uint32_t kMfmaTableBytes = kMfmaTwistOff + 4096u * 8u;
constexpr uint32_t kMfmaRow = 68;  // LDS words per 64-element row (pad keeps b128 reads conflict-free)
constexpr int kMfmaUnroll = 8;     // 16-byte global accesses in flight per thread

__device__ __forceinline__ uint32_t mf_col_words(uint32_t log_r) { return log_r == 12 ? 64u * kMfmaRow + 8u : 72u; }

// Y = (a0 + 2^8 a1) + 2^16 (a2 + 2^8 a3), then Y - rint(Y/q) q in (-0.52q, 0.52q).
__device__ __forceinline__ int mf_reduce(int a0, int a1, int a2, int a3, double qd, double qinv) {
    const int t = a0 + (a1 << 8), u = a2 + (a3 << 8);
    const double y = fma((double)u, 65536.0, (double)t);
    const double qe = rint(y * qinv);
    return (int)fma(-qe, qd, y);
}

// One job's MFMA pass over one tile (R rows x C columns), LDS words `lds`.
template <class Sin, class Sout>
__device__ __forceinline__ void mfma_tile(const PassJob& J, uint32_t tile_in_job, uint32_t* lds, const v4i (&fa)[4][4],
                                          const v4i (&ca)[4]) {
    const uint32_t tid = threadIdx.x, T = blockDim.x;
    const uint32_t lane = tid & 63u, wave = tid >> 6, mt = wave & 3u, nh = wave >> 2, nwh = T >> 8;  // wave halves
    const uint32_t n = lane & 15u, g = lane >> 4;
    const uint32_t poly = tile_in_job / J.tiles_per_poly;
    const uint32_t tile = tile_in_job - poly * J.tiles_per_poly;
    const uint32_t log_r = J.log_r, log_c = J.log_c;
    const uint32_t R = 1u << log_r, C = 1u << log_c;
    const bool two = log_r == 12;
    const uint32_t lg = two ? 6u : 0u;
    const uint32_t SC = mf_col_words(log_r);
    const uint32_t col0 = tile << log_c;
    const uint32_t total = R << log_c;
    const uint32_t ncols = J.ncols;
    const Sin* __restrict__ in = reinterpret_cast<const Sin*>(J.in) + (int64_t)poly * J.in_pstride;
    Sout* __restrict__ out = reinterpret_cast<Sout*>(J.out) + (int64_t)poly * J.out_pstride;
    const uint32_t q = (uint32_t)J.q;
    const double qd = (double)q, qinv = 1.0 / (double)q;
    const uint32_t flags = J.flags;

    // ---- load: element (r, c) -> lds[c*SC + (r & (2^lg-1))*68 + (r >> lg)] ----
    auto xaddr = [&](uint32_t r, uint32_t c) -> uint32_t {
        return c * SC + (r & ((1u << lg) - 1u)) * kMfmaRow + (r >> lg);
    };
    if (flags & JF_VEC4) {
        // 16-byte row segments, kMfmaUnroll in flight per thread.  With one
        // segment per row (C == 4) consecutive lanes take rows 64 apart, so
        // their LDS words (consecutive j1) fall in distinct banks.
        const uint32_t qpr = C >> 2;  // quads per row
        const uint32_t nq = total >> 2;
        for (uint32_t q0 = tid; q0 < nq; q0 += kMfmaUnroll * T) {
            Sin v[kMfmaUnroll][4];
            uint32_t rr[kMfmaUnroll], cc[kMfmaUnroll];
#pragma unroll
            for (int u = 0; u < kMfmaUnroll; u++) {
                const uint32_t qi = q0 + u * T;
                uint32_t r = qi / qpr;
                cc[u] = (qi - r * qpr) << 2;
                if (two && qpr == 1) r = ((r & 63u) << 6) | (r >> 6);
                rr[u] = r;
                if (qi < nq && col0 + cc[u] < ncols) load4<Sin>(in + (int64_t)r * J.s_r + (int64_t)(col0 + cc[u]), v[u]);
            }
#pragma unroll
            for (int u = 0; u < kMfmaUnroll; u++) {
                const uint32_t qi = q0 + u * T;
                if (qi >= nq) continue;
                const bool ok = col0 + cc[u] < ncols;
#pragma unroll
                for (int e = 0; e < 4; e++) lds[xaddr(rr[u], cc[u] + e)] = ok ? (uint32_t)v[u][e] : 0u;
            }
        }
    } else {
        for (uint32_t i0 = tid; i0 < total; i0 += kUnroll * T) {
            Sin v[kUnroll];
            uint32_t la[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const uint32_t idx = i0 + u * T;
                uint32_t r, c;
                if (flags & JF_LOAD_CFAST) { r = idx >> log_c; c = idx & (C - 1u); }
                else { c = idx >> log_r; r = idx & (R - 1u); }
                la[u] = xaddr(r & (R - 1u), c & (C - 1u));
                v[u] = 0;
                if (idx < total && col0 + c < ncols) v[u] = in[(int64_t)r * J.s_r + (int64_t)(col0 + c) * J.s_c];
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (i0 + u * T < total) lds[la[u]] = (uint32_t)v[u];
        }
    }
    __syncthreads();

    const TwP<uint32_t>* tw = reinterpret_cast<const TwP<uint32_t>*>(reinterpret_cast<const char*>(J.z) + kMfmaTwistOff);
    if (!two) {
        // R = 64: 16-column tiles; every wave reads its tiles, then all write.
        constexpr int kMaxNt = 8;  // C <= 256 columns
        int res[kMaxNt][4];
        const uint32_t ntiles = C >> 4;
        int it = 0;
        for (uint32_t nt = nh; nt < ntiles && it < kMaxNt; nt += nwh, it++) {
            const uint32_t c = 16u * nt + n;
            v4i b[4];
#pragma unroll
            for (int ks = 0; ks < 4; ks++) {
                b[ks] = *reinterpret_cast<const v4i*>(lds + c * SC + 16u * ks + 4u * g);
                b[ks] ^= (v4i)(int)0x80808080;
            }
            v4i acc[4];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                acc[i] = ca[i];
#pragma unroll
                for (int ks = 0; ks < 4; ks++) acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i][ks], b[ks], acc[i], 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; e++) {
                const int r = mf_reduce(acc[0][e], acc[1][e], acc[2][e], acc[3][e], qd, qinv);
                res[it][e] = r < 0 ? r + (int)q : r;
            }
        }
        __syncthreads();
        it = 0;
        for (uint32_t nt = nh; nt < ntiles && it < kMaxNt; nt += nwh, it++) {
            const uint32_t c = 16u * nt + n;
#pragma unroll
            for (int e = 0; e < 4; e++) lds[c * SC + 16u * mt + 4u * g + e] = (uint32_t)res[it][e];
        }
        __syncthreads();
    } else {
        // Column pipeline, in place: at step s waves 0-3 run stage A of column s
        // (X[s] -> Y[s]) while waves 4-7 run stage B of column s-1 (Y -> Z); each
        // half reads, holds its results in registers over a barrier, then
        // writes them into the same column region.
        const bool stage_a = wave < 4;
        for (uint32_t st = 0; st <= C; st++) {
            const bool active = stage_a ? st < C : st >= 1;
            const uint32_t c = stage_a ? st : st - 1;
            uint32_t res[4][4];
            if (active) {
                uint32_t* col = lds + c * SC;
#pragma unroll
                for (uint32_t nt = 0; nt < 4; nt++) {
                    const uint32_t u = 16u * nt + n;  // j2 (stage A) or k1 (stage B)
                    TwP<uint32_t> tv[4];
                    if (stage_a) {
#pragma unroll
                        for (int e = 0; e < 4; e++) tv[e] = tw[(16u * mt + 4u * g + e) * 64u + u];
                    }
                    v4i b[4];
#pragma unroll
                    for (int ks = 0; ks < 4; ks++) {
                        b[ks] = *reinterpret_cast<const v4i*>(col + u * kMfmaRow + 16u * ks + 4u * g);
                        if (stage_a) b[ks] ^= (v4i)(int)0x80808080;
                    }
                    v4i acc[4];
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        acc[i] = ca[i];
#pragma unroll
                        for (int ks = 0; ks < 4; ks++)
                            acc[i] = __builtin_amdgcn_mfma_i32_16x16x64_i8(fa[i][ks], b[ks], acc[i], 0, 0, 0);
                    }
#pragma unroll
                    for (int e = 0; e < 4; e++) {
                        const int r = mf_reduce(acc[0][e], acc[1][e], acc[2][e], acc[3][e], qd, qinv);
                        if (stage_a) {
                            const uint32_t x = (uint32_t)r + q;  // [0, 2q)
                            const uint32_t y = x * tv[e].w - __umulhi(x, tv[e].wp) * q;  // [0, 2q)
                            res[nt][e] = y ^ 0x80808080u;
                        } else {
                            res[nt][e] = (uint32_t)(r < 0 ? r + (int)q : r);
                        }
                    }
                }
            }
            __syncthreads();
            if (active) {
                uint32_t* col = lds + c * SC;
                // stage A writes Y[k1][j2]; stage B writes Z[k2][k1]: row 16mt+4g+e, column 16nt+n
#pragma unroll
                for (uint32_t nt = 0; nt < 4; nt++)
#pragma unroll
                    for (int e = 0; e < 4; e++) col[(16u * mt + 4u * g + e) * kMfmaRow + 16u * nt + n] = res[nt][e];
            }
            __syncthreads();
        }
    }

    // ---- store: output row k of column c at lds[c*SC + (k >> 6)*68 + (k & 63)] ----
    const Mod32 mod{q};
    const TwP<uint32_t>* tlo = reinterpret_cast<const TwP<uint32_t>*>(J.tw_lo);
    const TwP<uint32_t>* thi = reinterpret_cast<const TwP<uint32_t>*>(J.tw_hi);
    const TwP<uint32_t>* tmid = reinterpret_cast<const TwP<uint32_t>*>(J.tw_mid);
    const uint32_t post_mode = J.post;
    const TwP<uint32_t> scale{(uint32_t)J.scale_w, (uint32_t)J.scale_wp};
    const uint32_t lomask = (1u << J.lo_bits) - 1u;
    const uint32_t omask = (1u << J.o_lbits) - 1u;
    auto postop = [&](uint32_t x, uint32_t k, uint32_t cg) -> uint32_t {
        if (post_mode == POST_FOURSTEP) {
            const uint32_t uu = cg >> J.e_shift;
            const uint32_t e = (J.e_a * k * uu + J.e_b * k + J.e_d * uu) & J.e_mask;
            x = mod.mul(x, tlo[e & lomask]);
            if (J.mid_bits) x = mod.mul(x, tmid[(e >> J.lo_bits) & ((1u << J.mid_bits) - 1u)]);
            return mod.mul(x, thi[e >> (J.lo_bits + J.mid_bits)]);
        }
        if (post_mode == POST_SCALE) return mod.mul(x, scale);
        return x;
    };
    auto oaddr = [&](uint32_t k, uint32_t cg) -> int64_t {
        return (int64_t)(cg >> J.o_lbits) * J.o_hs + (int64_t)(cg & omask) * J.o_c + (int64_t)k * J.o_r;
    };
    auto zaddr = [&](uint32_t k, uint32_t c) -> uint32_t { return c * SC + (k >> 6) * kMfmaRow + (k & 63u); };
    if (flags & JF_VEC4_ST) {
        const uint32_t nq = total >> 2;
        for (uint32_t q0 = tid; q0 < nq; q0 += kMfmaUnroll * T) {
            Sout v[kMfmaUnroll][4];
            int64_t a[kMfmaUnroll];
            bool ok[kMfmaUnroll];
#pragma unroll
            for (int u = 0; u < kMfmaUnroll; u++) {
                const uint32_t qi = q0 + u * T;
                const uint32_t idx = qi << 2;
                uint32_t k, c;
                if (flags & JF_STORE_CFAST) { k = idx >> log_c; c = idx & (C - 1u); }
                else { c = idx >> log_r; k = idx & (R - 1u); }
                ok[u] = qi < nq && col0 + c < ncols;
                a[u] = oaddr(k, col0 + c);
                if (ok[u]) {
                    if (flags & JF_STORE_CFAST) {
#pragma unroll
                        for (int e = 0; e < 4; e++) v[u][e] = (Sout)postop(lds[zaddr(k, c + e)], k, col0 + c + e);
                    } else {
                        const v4i z = *reinterpret_cast<const v4i*>(lds + zaddr(k, c));
#pragma unroll
                        for (int e = 0; e < 4; e++) v[u][e] = (Sout)postop((uint32_t)z[e], k + e, col0 + c);
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < kMfmaUnroll; u++)
                if (ok[u]) store4<Sout>(out + a[u], v[u]);
        }
    } else {
        for (uint32_t i0 = tid; i0 < total; i0 += kUnroll * T) {
            uint32_t v[kUnroll];
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
                v[u] = ok[u] ? postop(lds[zaddr(k, c)], k, col0 + c) : 0u;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++)
                if (ok[u]) out[a[u]] = (Sout)v[u];
        }
    }
    __syncthreads();
}

// Persistent launch: workgroup b handles tiles b, b + grid, ...  The fragments
// of the current plan stay in registers across tiles of the same plan.
template <class Sin, class Sout>
__global__ void __launch_bounds__(512) mfma_pass_kernel(const PassJob* __restrict__ jobs, uint32_t njobs, PassJob inl,
                                                        uint32_t total_tiles) {
    extern __shared__ __align__(16) unsigned char lds_raw[];
    uint32_t* lds = reinterpret_cast<uint32_t*>(lds_raw);
    const uint32_t lane = threadIdx.x & 63u, mt = (threadIdx.x >> 6) & 3u, g = lane >> 4;
    v4i fa[4][4], cav[4];
    const void* cur = nullptr;
    // consecutive tiles (which share cache lines of their column segments) go to
    // workgroups on the same XCD: workgroup b runs on XCD b % 8
    for (uint32_t t = xcd_remap(blockIdx.x, gridDim.x); t < total_tiles; t += gridDim.x) {
        uint32_t j = 0;
        if (njobs) {
            uint32_t lo = 0, hi = njobs;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (jobs[mid].tile_begin <= t) lo = mid; else hi = mid;
            }
            j = lo;
        }
        const PassJob J = njobs ? jobs[j] : inl;
        if (J.z != cur) {
            // waves 4-7 of a 4096-point pass run stage B, every other wave stage A
            cur = J.z;
            const uint32_t stg = (J.log_r == 12 && (threadIdx.x >> 6) >= 4) ? 1u : 0u;
            const v4i* fr = reinterpret_cast<const v4i*>(J.z);
            const int* cr = reinterpret_cast<const int*>(reinterpret_cast<const char*>(J.z) + kMfmaCorrOff);
#pragma unroll
            for (int i = 0; i < 4; i++) {
#pragma unroll
                for (int ks = 0; ks < 4; ks++) fa[i][ks] = fr[(((stg * 4 + i) * 4 + mt) * 4 + ks) * 64 + lane];
#pragma unroll
                for (int e = 0; e < 4; e++) cav[i][e] = cr[(stg * 64 + 16 * mt + 4 * g + e) * 4 + i];
            }
        }
        mfma_tile<Sin, Sout>(J, t - J.tile_begin, lds, fa, cav);
    }
}

}  // namespace nttk
