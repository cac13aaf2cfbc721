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

template <typename T, int K, bool NTM>
__global__ __launch_bounds__(256) void k_fused(vec2_t<T> *__restrict__ psi, const double *__restrict__ U,
                                               GroupSpec g, FusedOff off) {
    using V = vec2_t<T>;
    using v4 = typename Mfma<T>::v4;
    constexpr int D = 1 << K;
    constexpr int MT = D < 16 ? 1 : D / 16;  // 16-row output blocks
    constexpr int KS = D / 4;                // 4-row k-steps
    __shared__ V fr[MT * KS * 64];

    // A fragments: fr[(mt*KS + ks)*64 + lane] = U_s[mt*16 + i][ks*4 + k],
    // lane = i + 16k, U_s = U with rows and columns in sorted-target order.
    for (int e = threadIdx.x; e < MT * KS * 64; e += 256) {
        const int ln = e & 63, blk = e >> 6;
        const int mt = blk / KS, ks = blk % KS;
        const int rs = mt * 16 + (ln & 15), cs = ks * 4 + (ln >> 4);
        V v;
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

namespace {
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
    static int per_cu = 0;
    if (per_cu == 0) {
        int nb = 0;
        per_cu =
            (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_fused<T, K, false>, 256, 0) == hipSuccess && nb > 0)
                ? nb
                : 2;
    }
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
