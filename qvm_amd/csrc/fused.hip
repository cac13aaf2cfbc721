// fused.hip -- dense fused k-qubit blocks (k = 3..6) on the matrix cores.
//
// Applying a 2^k x 2^k complex U to every complement group is a tall-skinny
// complex GEMM Y = U X, X = 2^k x 2^(n-k) (one column per group).  It runs as
// split-complex MFMAs:
//     Y_re += U_re X_re + U_im (-X_im)      Y_im += U_im X_re + U_re X_im
// with v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32.  The operand layout of
// these instructions puts B[k][j] in lane j + 16k, so one lane's B values for
// the four products of a k-step are the real and imaginary part of ONE
// amplitude (row 4*ks + k of column j): every lane loads whole complex
// amplitudes straight from HBM, and the result layout hands every lane the
// real and imaginary part of the same output amplitudes, stored straight
// back.  No LDS staging of X, no barriers in the column loop.
//
// Rows of U are renumbered so that bit b of a row index is the b-th lowest
// target qubit; a k-step's four rows then differ in the two lowest targets,
// and 16 consecutive complement groups x those rows are contiguous runs.  The
// A fragments (U_re, U_im per 16x4 block and lane) are built once per
// workgroup in LDS from U in nat-tuple order.  Each wave takes 16 complement
// groups at a time, striding through the contiguous share of column groups
// that its XCD owns, with the next group's loads issued before the current
// group's MFMAs.  The four real products per complex product are
// all kept, so rounding is that of a complex matvec with FMA.
#include <algorithm>
#include <atomic>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

template <typename T> struct Mfma;
// Operand layouts (both instructions): lane = i + 16*k holds A[i][k] and
// lane = j + 16*k holds B[k][j].  Result layouts differ: register e of lane
// (j + 16*b) holds row 4*b + e for the f32 instruction (one group of four
// rows per lane block) and row b + 4*e for the f64 one (four groups of one).
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int blk, int e) { return blk + 4 * e; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int blk, int e) { return 4 * blk + e; }
};

struct FusedOff {
    uint64_t o[64];  // offset of sorted-order row a: sum of 2^sorted[b] over the bits of a
    uint32_t nat[6]; // position in the nat-tuple of the b-th lowest target
};

// A fragments in LDS: fr[(mt*KS + ks)*64 + lane] = U_s[mt*16 + i][ks*4 + k],
// lane = i + 16k, U_s = U with rows and columns in sorted-target order
// (rows past 2^K, for K = 3, are zero).  Ends with a barrier.
template <typename T, int K>
__device__ __forceinline__ void build_fragments(vec2_t<T> *fr, const double *__restrict__ U, const FusedOff &off) {
    constexpr int D = 1 << K;
    constexpr int MT = D < 16 ? 1 : D / 16;
    constexpr int KS = D / 4;
    for (int e = threadIdx.x; e < MT * KS * 64; e += 256) {
        const int ln = e & 63, blk = e >> 6;
        const int mt = blk / KS, ks = blk % KS;
        const int rs = mt * 16 + (ln & 15), cs = ks * 4 + (ln >> 4);
        vec2_t<T> v;
        v.x = T(0);
        v.y = T(0);
        if (rs < D) {
            int rn = 0, cn = 0;
#pragma unroll
            for (int b = 0; b < K; ++b) {
                rn |= ((rs >> b) & 1) << off.nat[b];
                cn |= ((cs >> b) & 1) << off.nat[b];
            }
            v.x = (T)U[2 * (rn * D + cn)];
            v.y = (T)U[2 * (rn * D + cn) + 1];
        }
        fr[e] = v;
    }
    __syncthreads();
}

template <typename T, int K, bool NTM>
__global__ __launch_bounds__(256) void k_fused(vec2_t<T> *__restrict__ psi, const double *__restrict__ U,
                                               GroupSpec g, FusedOff off) {
    using V = vec2_t<T>;
    using v4 = typename Mfma<T>::v4;
    constexpr int D = 1 << K;
    constexpr int MT = D < 16 ? 1 : D / 16;  // 16-row output blocks
    constexpr int KS = D / 4;                // 4-row k-steps
    __shared__ V fr[MT * KS * 64];
    build_fragments<T, K>(fr, U, off);

    const int lane = threadIdx.x & 63;
    const int j = lane & 15, kk = lane >> 4;
    const uint64_t ncols = g.ngroups;
    const uint64_t ncg = (ncols + 15) / 16;
    // each XCD takes a contiguous share of the column groups; the waves of
    // its workgroups stride through that share
    const XcdShare sh = xcd_share(blockIdx.x, gridDim.x, ncg);
    const uint64_t wstride = (uint64_t)sh.count * 4;
    // lane parts of the row offsets: loads take row 4*ks + kk; the result
    // rows of register e are kk + 4e (f64) or 4*kk + e (f32)
    const uint64_t lo_ld = off.o[kk];
    const uint64_t lo_st = sizeof(T) == 8 ? off.o[kk] : off.o[(4 * kk) & (D - 1)];
    const bool st_lane_ok = sizeof(T) == 8 || 4 * kk < D;

    auto load = [&](uint64_t cg, V (&xv)[KS], uint64_t &base, bool &ok) {
        const uint64_t col = cg * 16 + j;
        ok = col < ncols;
        base = group_base(g, ok ? col : 0);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            if (ok) {
                xv[ks] = ldg<NTM>(&psi[base + (off.o[4 * ks] | lo_ld)]);
            } else {
                xv[ks].x = T(0);
                xv[ks].y = T(0);
            }
        }
    };

    // k <= 5: the next column group's loads are issued before this group's
    // MFMAs (register double buffer); k = 6 has no registers to spare for it
    // and relies on the other waves of the SIMD instead.
    constexpr bool PF = K <= 5;
    uint64_t cg = sh.lo + (uint64_t)sh.rank * 4 + (threadIdx.x >> 6);
    V xn[PF ? KS : 1];
    uint64_t bn = 0;
    bool okn = false;
    if constexpr (PF) {
        if (cg < sh.hi) load(cg, xn, bn, okn);
    }
    for (; cg < sh.hi; cg += wstride) {
        V xv[KS];
        uint64_t base;
        bool ok;
        if constexpr (PF) {
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) xv[ks] = xn[ks];
            base = bn;
            ok = okn;
            if (cg + wstride < sh.hi) load(cg + wstride, xn, bn, okn);  // in flight during the MFMAs
        } else {
            load(cg, xv, base, ok);
        }

        v4 ar[MT], ai[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
            ar[mt] = v4{0, 0, 0, 0};
            ai[mt] = v4{0, 0, 0, 0};
        }
        // A fragments of k-step ks+1 are read from LDS while ks multiplies;
        // the scheduling barrier keeps the compiler from hoisting all of
        // them (k = 6 would need 256 registers for them)
        V an[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) an[mt] = fr[mt * KS * 64 + lane];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            V ac[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) ac[mt] = an[mt];
            if (ks + 1 < KS) {
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) an[mt] = fr[(mt * KS + ks + 1) * 64 + lane];
            }
            const T br = xv[ks].x, bi = xv[ks].y, nbi = -xv[ks].y;
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                ar[mt] = Mfma<T>::mma(ac[mt].x, br, ar[mt]);
                ai[mt] = Mfma<T>::mma(ac[mt].y, br, ai[mt]);
                ar[mt] = Mfma<T>::mma(ac[mt].y, nbi, ar[mt]);
                ai[mt] = Mfma<T>::mma(ac[mt].x, bi, ai[mt]);
            }
            __builtin_amdgcn_sched_barrier(0);
        }
        if (ok && st_lane_ok) {
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    // uniform part of the result row: 4e + 16mt (f64) or e + 16mt (f32)
                    const int hi = (sizeof(T) == 8 ? 4 * e : e) + 16 * mt;
                    if (hi < D) {
                        V y;
                        y.x = ar[mt][e];
                        y.y = ai[mt][e];
                        stg<NTM>(&psi[base + (off.o[hi] | lo_st)], y);
                    }
                }
            }
        }
    }
}

// 4x4 transpose of 16-lane blocks over four registers: afterwards block g of
// register k holds what block k of register g held.  Two rounds of the gfx950
// lane swaps (permlane32_swap exchanges the upper 32 lanes of its first
// operand with the lower 32 of its second, permlane16_swap the odd 16-lane
// blocks of the first with the even blocks of the second), one word at a
// time.  An involution.
__device__ __forceinline__ void xpose4_words(uint32_t &a0, uint32_t &a1, uint32_t &a2, uint32_t &a3) {
    const auto p = __builtin_amdgcn_permlane32_swap(a0, a2, false, false);
    const auto q = __builtin_amdgcn_permlane32_swap(a1, a3, false, false);
    const auto r = __builtin_amdgcn_permlane16_swap(p[0], q[0], false, false);
    const auto t = __builtin_amdgcn_permlane16_swap(p[1], q[1], false, false);
    a0 = r[0];
    a1 = r[1];
    a2 = t[0];
    a3 = t[1];
}

template <typename X>
__device__ __forceinline__ void xpose4(X &a0, X &a1, X &a2, X &a3) {
    static_assert(sizeof(X) % 4 == 0, "whole words");
    constexpr int W = sizeof(X) / 4;
    uint32_t w0[W], w1[W], w2[W], w3[W];
    __builtin_memcpy(w0, &a0, sizeof(X));
    __builtin_memcpy(w1, &a1, sizeof(X));
    __builtin_memcpy(w2, &a2, sizeof(X));
    __builtin_memcpy(w3, &a3, sizeof(X));
#pragma unroll
    for (int i = 0; i < W; ++i) xpose4_words(w0[i], w1[i], w2[i], w3[i]);
    __builtin_memcpy(&a0, w0, sizeof(X));
    __builtin_memcpy(&a1, w1, sizeof(X));
    __builtin_memcpy(&a2, w2, sizeof(X));
    __builtin_memcpy(&a3, w3, sizeof(X));
}

// The same block product with one whole row per memory instruction: a wave
// takes 64 consecutive complement groups (lane l holds column 64 cb + l) and
// each load or store moves one row of them, a 1 KiB run (FP64) when the
// columns are contiguous, where the MFMA operand layout above moves 4 rows x
// 256 B.  The four rows of a k-step are turned into the B fragments of the
// four 16-column groups by one 4x4 block transpose across lanes (xpose4),
// and the results back into rows by the same transpose, so the products, and
// their rounding, are those of k_fused.  Needs 2^(n-K) >= 64 columns.  Used
// for K = 4, and K = 5 in FP32 (the four column groups' accumulators of FP64
// K = 5 and of K = 6 leave room for one wave per SIMD).
template <typename T, int K, bool NTM>
__global__ __launch_bounds__(256) void k_fused_rows(vec2_t<T> *__restrict__ psi, const double *__restrict__ U,
                                                    GroupSpec g, FusedOff off) {
    using V = vec2_t<T>;
    using v4 = typename Mfma<T>::v4;
    constexpr int D = 1 << K;
    constexpr int MT = D < 16 ? 1 : D / 16;
    constexpr int KS = D / 4;
    __shared__ V fr[MT * KS * 64];
    build_fragments<T, K>(fr, U, off);

    const int lane = threadIdx.x & 63;
    const uint64_t ncb = g.ngroups / 64;
    const XcdShare sh = xcd_share(blockIdx.x, gridDim.x, ncb);
    const uint64_t wstride = (uint64_t)sh.count * 4;
    for (uint64_t cb = sh.lo + (uint64_t)sh.rank * 4 + (threadIdx.x >> 6); cb < sh.hi; cb += wstride) {
        const uint64_t base = group_base(g, cb * 64 + (uint64_t)lane);
        v4 ar[4][MT], ai[4][MT];
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                ar[c][mt] = v4{0, 0, 0, 0};
                ai[c][mt] = v4{0, 0, 0, 0};
            }
        // the next k-step's four rows are in flight during this one's MFMAs
        V xn[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) xn[k] = ldg<NTM>(&psi[base + off.o[k]]);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            V b[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) b[k] = xn[k];
            if (ks + 1 < KS) {
#pragma unroll
                for (int k = 0; k < 4; ++k) xn[k] = ldg<NTM>(&psi[base + off.o[4 * (ks + 1) + k]]);
            }
            // rows 4ks..4ks+3 of 64 columns -> B fragments of column groups 0..3
            xpose4(b[0], b[1], b[2], b[3]);
            V a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) a[mt] = fr[(mt * KS + ks) * 64 + lane];
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const T br = b[c].x, bi = b[c].y, nbi = -b[c].y;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    ar[c][mt] = Mfma<T>::mma(a[mt].x, br, ar[c][mt]);
                    ai[c][mt] = Mfma<T>::mma(a[mt].y, br, ai[c][mt]);
                    ar[c][mt] = Mfma<T>::mma(a[mt].y, nbi, ar[c][mt]);
                    ai[c][mt] = Mfma<T>::mma(a[mt].x, bi, ai[c][mt]);
                }
            }
        }
        // result register e of column group c, lane block b, holds row
        // 16mt + b + 4e (f64) or 16mt + 4b + e (f32); transposed across c and
        // b, register b then holds that row for all 64 columns
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                T yr[4], yi[4];
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    yr[c] = ar[c][mt][e];
                    yi[c] = ai[c][mt][e];
                }
                xpose4(yr[0], yr[1], yr[2], yr[3]);
                xpose4(yi[0], yi[1], yi[2], yi[3]);
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const int row = 16 * mt + (sizeof(T) == 8 ? b + 4 * e : 4 * b + e);
                    if (row < D) {
                        V y;
                        y.x = yr[b];
                        y.y = yi[b];
                        stg<NTM>(&psi[base + off.o[row]], y);
                    }
                }
            }
        }
    }
}

// Two registers each holding two rows of 32 columns (lanes 0-31 one row,
// lanes 32-63 the next) <-> the B fragments of two 16-column groups: one
// permlane32_swap and one permlane16_swap per word; the inverse applies them
// in the other order (each swap is an involution).
__device__ __forceinline__ void rows_to_frag2_words(uint32_t &a0, uint32_t &a1) {
    const auto p = __builtin_amdgcn_permlane32_swap(a0, a1, false, false);
    const auto r = __builtin_amdgcn_permlane16_swap(p[0], p[1], false, false);
    a0 = r[0];
    a1 = r[1];
}
__device__ __forceinline__ void frag2_to_rows_words(uint32_t &a0, uint32_t &a1) {
    const auto r = __builtin_amdgcn_permlane16_swap(a0, a1, false, false);
    const auto p = __builtin_amdgcn_permlane32_swap(r[0], r[1], false, false);
    a0 = p[0];
    a1 = p[1];
}
template <bool TO_FRAG, typename X>
__device__ __forceinline__ void xpose2(X &a0, X &a1) {
    static_assert(sizeof(X) % 4 == 0, "whole words");
    constexpr int W = sizeof(X) / 4;
    uint32_t w0[W], w1[W];
    __builtin_memcpy(w0, &a0, sizeof(X));
    __builtin_memcpy(w1, &a1, sizeof(X));
#pragma unroll
    for (int i = 0; i < W; ++i) {
        if constexpr (TO_FRAG)
            rows_to_frag2_words(w0[i], w1[i]);
        else
            frag2_to_rows_words(w0[i], w1[i]);
    }
    __builtin_memcpy(&a0, w0, sizeof(X));
    __builtin_memcpy(&a1, w1, sizeof(X));
}

// FP64 K = 5 in rows of 32 columns: a wave takes 32 consecutive complement
// groups, and each load or store moves two rows of them (512 B runs, against
// 256 B in the operand layout); half the accumulators of k_fused_rows, so the
// 32-row blocks fit.  Same products as k_fused.  Measured at n = 30 on six
// target sets above the line: 5.85-6.74 ms against 6.81-7.01 for k_fused,
// 5.5 % less in total (+20 % on 5,11,17,23,27).
template <typename T, int K, bool NTM>
__global__ __launch_bounds__(256) void k_fused_rows2(vec2_t<T> *__restrict__ psi, const double *__restrict__ U,
                                                     GroupSpec g, FusedOff off) {
    using V = vec2_t<T>;
    using v4 = typename Mfma<T>::v4;
    constexpr int D = 1 << K;
    constexpr int MT = D < 16 ? 1 : D / 16;
    constexpr int KS = D / 4;
    __shared__ V fr[MT * KS * 64];
    build_fragments<T, K>(fr, U, off);

    const int lane = threadIdx.x & 63;
    const int half = lane >> 5;
    const uint64_t ncb = g.ngroups / 32;
    const XcdShare sh = xcd_share(blockIdx.x, gridDim.x, ncb);
    const uint64_t wstride = (uint64_t)sh.count * 4;
    // lane parts of the row offsets: loads take row bit 0 = half; the rows of
    // result lane block b = 2r + half have half at bit 0 (f64) or bit 2 (f32)
    const uint64_t lh_ld = off.o[half];
    const uint64_t lh_st = off.o[sizeof(T) == 8 ? half : 4 * half];
    for (uint64_t cb = sh.lo + (uint64_t)sh.rank * 4 + (threadIdx.x >> 6); cb < sh.hi; cb += wstride) {
        const uint64_t base = group_base(g, cb * 32 + (uint64_t)(lane & 31));
        // the A fragments are read per column block: hoisted out of the loop
        // they would hold 16 registers per k-step (the barrier is on the
        // index: on the pointer it would hide the LDS address space, and the
        // reads would become flat loads that also wait for the next k-step's
        // global loads)
        uint32_t fo = (uint32_t)lane;
        asm volatile("" : "+v"(fo));
        const V *frl = fr + fo;
        v4 ar[2][MT], ai[2][MT];
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                ar[c][mt] = v4{0, 0, 0, 0};
                ai[c][mt] = v4{0, 0, 0, 0};
            }
        // register 0 holds rows 4ks, 4ks+1, register 1 rows 4ks+2, 4ks+3 (one
        // per half-wave); the next k-step's are in flight during the MFMAs.
        // A row's offset is the OR of its bits' offsets: the half-wave's bit
        // (row bit 0) is the lane part, the rest is uniform.
        V xn[2];
        xn[0] = ldg<NTM>(&psi[base + (off.o[0] | lh_ld)]);
        xn[1] = ldg<NTM>(&psi[base + (off.o[2] | lh_ld)]);
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            V b[2] = {xn[0], xn[1]};
            if (ks + 1 < KS) {
                xn[0] = ldg<NTM>(&psi[base + (off.o[4 * (ks + 1)] | lh_ld)]);
                xn[1] = ldg<NTM>(&psi[base + (off.o[4 * (ks + 1) + 2] | lh_ld)]);
            }
            xpose2<true>(b[0], b[1]);  // -> B fragments of column groups 0, 1
            V a[MT];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) a[mt] = frl[(mt * KS + ks) * 64];
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                const T br = b[c].x, bi = b[c].y, nbi = -b[c].y;
#pragma unroll
                for (int mt = 0; mt < MT; ++mt) {
                    ar[c][mt] = Mfma<T>::mma(a[mt].x, br, ar[c][mt]);
                    ai[c][mt] = Mfma<T>::mma(a[mt].y, br, ai[c][mt]);
                    ar[c][mt] = Mfma<T>::mma(a[mt].y, nbi, ar[c][mt]);
                    ai[c][mt] = Mfma<T>::mma(a[mt].x, bi, ai[c][mt]);
                }
            }
        }
        // result register e of column group c, lane block b: row 16mt + b +
        // 4e (f64) or 16mt + 4b + e (f32); back to rows, register r holds
        // lane block 2r + half's row for the wave's 32 columns
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                T yr[2] = {ar[0][mt][e], ar[1][mt][e]};
                T yi[2] = {ai[0][mt][e], ai[1][mt][e]};
                xpose2<false>(yr[0], yr[1]);
                xpose2<false>(yi[0], yi[1]);
#pragma unroll
                for (int r = 0; r < 2; ++r) {
                    // uniform part of the row of lane block b = 2r + half
                    const int hi = 16 * mt + (sizeof(T) == 8 ? 2 * r + 4 * e : 8 * r + e);
                    if (hi < D) {
                        V y;
                        y.x = yr[r];
                        y.y = yi[r];
                        stg<NTM>(&psi[base + (off.o[hi] | lh_st)], y);
                    }
                }
            }
        }
    }
}

namespace {
// Resident 256-thread workgroups per CU of a kernel, asked once per kernel
// and kept in `cache` (concurrent first calls on several handles may each ask
// and store the same value).
template <typename F>
int resident_per_cu(std::atomic<int> &cache, F kernel) {
    int v = cache.load(std::memory_order_relaxed);
    if (v == 0) {
        int nb = 0;
        v = (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kernel, 256, 0) == hipSuccess && nb > 0) ? nb : 2;
        cache.store(v, std::memory_order_relaxed);
    }
    return v;
}

template <typename T, int K>
int run_fused(qvmhip_state *s, const void *mat_dev, const uint32_t *qubits) {
    GroupSpec g;
    std::memset(&g, 0, sizeof(g));
    std::vector<uint32_t> pos(qubits, qubits + K);
    std::sort(pos.begin(), pos.end());
    g.nsorted = K;
    for (int i = 0; i < K; ++i) g.sorted[i] = pos[i];
    g.ngroups = uint64_t(1) << (s->n - K);
    FusedOff off;
    std::memset(&off, 0, sizeof(off));
    for (int b = 0; b < K; ++b)
        for (int c = 0; c < K; ++c)
            if (qubits[c] == pos[b]) off.nat[b] = (uint32_t)c;
    for (int a = 0; a < 64; ++a) {
        uint64_t o = 0;
        for (int b = 0; b < K; ++b)
            if ((a >> b) & 1) o |= uint64_t(1) << pos[b];
        off.o[a] = o;
    }
    // Every wave grid-strides over equal shares of column groups, so the grid
    // is exactly the workgroups that fit on the device at once: a second,
    // partial round would run at a fraction of the occupancy.
    static std::atomic<int> per_cu_cache{0};
    const int per_cu = resident_per_cu(per_cu_cache, k_fused<T, K, false>);
    // Non-temporal accesses when every 128 B line a k-step touches is read
    // whole in that k-step: the k-step's rows vary the two lowest targets, so
    // no other target may lie below the line size.  Otherwise a line is read
    // in two k-steps and streaming it past the cache reads it twice
    // (measured: targets 0,1,2 at k = 3, 7.7 ms cached vs 8.7 ms streamed;
    // targets 0,5,11, 6.5 vs 6.3 ms).  FP32 blocks stay cached: streamed,
    // the 30-qubit layered + QFT circuits in k = 4 / 5 blocks took 1176 /
    // 830 ms against 1119 / 786 ms cached.
    const uint32_t line_bits = sizeof(T) == 8 ? 3u : 4u;
    bool ntm = sizeof(T) == 8;
    for (int b = 2; b < K; ++b)
        if (pos[b] < line_bits) ntm = false;
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, s->device) != hipSuccess || ncu <= 0)
        ncu = 256;
    if constexpr (K == 5 && sizeof(T) == 8) {
        if (g.ngroups >= 32 && pos[0] >= line_bits) {
            static std::atomic<int> rows2_cache{0};
            const int per_cu_rows2 = resident_per_cu(rows2_cache, k_fused_rows2<T, K, true>);
            const uint64_t ncb = g.ngroups / 32;
            const uint32_t grid2 = (uint32_t)std::min<uint64_t>((ncb + 3) / 4, (uint64_t)per_cu_rows2 * (uint64_t)ncu);
            prof_begin(s, QVMHIP_K_FUSED);
            hipLaunchKernelGGL((k_fused_rows2<T, K, true>), dim3(grid2), dim3(256), 0, s->stream,
                               (vec2_t<T> *)s->amps, (const double *)mat_dev, g, off);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    if constexpr (K == 4 || (K == 5 && sizeof(T) == 4)) {
        // one row per memory instruction when every target lies above the
        // 128 B line, so that each row's lines are read whole by one
        // instruction.  Measured at n = 30 (ms per launch, this kernel vs
        // k_fused, same box): FP64 K = 4 on ten target sets 5.80-6.95 vs
        // 6.73-7.00, 5.3 % less in total, from -1 % (3,6,16,29) to +18 %
        // (12,19,20,22); FP32 K = 4 / 5 on 5,11,17,23(,27) 3.46 / 3.66 vs
        // 3.80 / 3.91.  With a target inside the line (0,5,11,17 FP64; 0,1,2,
        // 9,20 FP32) it lost 11-27 %, and at K = 3 4 %.
        if (g.ngroups >= 64 && pos[0] >= line_bits) {
            static std::atomic<int> rows_cache{0};
            const int per_cu_rows = resident_per_cu(rows_cache, k_fused_rows<T, K, false>);
            const bool nt_rows = sizeof(T) == 8;
            const uint64_t ncb = g.ngroups / 64;
            const uint32_t grid_rows =
                (uint32_t)std::min<uint64_t>((ncb + 3) / 4, (uint64_t)per_cu_rows * (uint64_t)ncu);
            prof_begin(s, QVMHIP_K_FUSED);
            if (nt_rows)
                hipLaunchKernelGGL((k_fused_rows<T, K, true>), dim3(grid_rows), dim3(256), 0, s->stream,
                                   (vec2_t<T> *)s->amps, (const double *)mat_dev, g, off);
            else
                hipLaunchKernelGGL((k_fused_rows<T, K, false>), dim3(grid_rows), dim3(256), 0, s->stream,
                                   (vec2_t<T> *)s->amps, (const double *)mat_dev, g, off);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    const uint64_t ncg = (g.ngroups + 15) / 16;
    const uint32_t grid = (uint32_t)std::min<uint64_t>((ncg + 3) / 4, (uint64_t)per_cu * (uint64_t)ncu);
    prof_begin(s, QVMHIP_K_FUSED);
    if (ntm)
        hipLaunchKernelGGL((k_fused<T, K, true>), dim3(grid), dim3(256), 0, s->stream, (vec2_t<T> *)s->amps,
                           (const double *)mat_dev, g, off);
    else
        hipLaunchKernelGGL((k_fused<T, K, false>), dim3(grid), dim3(256), 0, s->stream, (vec2_t<T> *)s->amps,
                           (const double *)mat_dev, g, off);
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T>
int run_fused_k(qvmhip_state *s, const void *mat_dev, uint32_t k, const uint32_t *qubits) {
    switch (k) {
        case 3: return run_fused<T, 3>(s, mat_dev, qubits);
        case 4: return run_fused<T, 4>(s, mat_dev, qubits);
        case 5: return run_fused<T, 5>(s, mat_dev, qubits);
        case 6: return run_fused<T, 6>(s, mat_dev, qubits);
    }
    return fail(QVMHIP_EINVAL, "fused blocks are 3..6 qubits");
}
}  // namespace

int launch_fused_dense(qvmhip_state *s, const void *mat_dev, uint32_t k, const uint32_t *qubits) {
    if (s->dtype == QVMHIP_F64) return run_fused_k<double>(s, mat_dev, k, qubits);
    return run_fused_k<float>(s, mat_dev, k, qubits);
}

}  // namespace qvmhip
