// fused.hip -- dense fused k-qubit blocks (k = 3..6) on the matrix cores.
//
// Applying a 2^k x 2^k complex U to every complement group is a tall-skinny
// complex GEMM Y = U X, X = 2^k x 2^(n-k).  It is computed as one real GEMM
// with the interleaved real embedding
//     A[2i][2j] = Re U_ij   A[2i][2j+1] = -Im U_ij
//     A[2i+1][2j] = Im U_ij A[2i+1][2j+1] = Re U_ij
// so the state's own (re, im) interleaving is the real X (rows 2j, 2j+1 of a
// column are Re/Im of the column's amplitude j) -- no deinterleave pass.
// The four real products per complex product are all kept (no 3-multiply
// form), so rounding stays that of an ordinary complex matvec with FMA.
//
// Workgroup: 4 wavefronts, 64 complement groups (columns) per tile; wave w
// owns columns 16w..16w+15 and all 2D/16 row tiles.  The A fragments of U
// are loaded once per workgroup into registers (k <= 5) or LDS (k = 6, rows
// padded by one element so the 16 lanes of a fragment hit distinct banks).
// Each X tile is gathered with coalesced loads one tile ahead (software
// pipelined, so HBM latency overlaps the MFMAs), staged in LDS, multiplied
// with v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32, written back into the
// wave's own LDS columns and stored with the same coalesced pattern.  The
// workgroup grid-strides over tiles.
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
    uint64_t o[64];
};

constexpr int kFusedCols = 64;

template <typename T, int K>
__global__ __launch_bounds__(256) void k_fused(vec2_t<T> *__restrict__ psi, const double *__restrict__ U,
                                               GroupSpec g, FusedOff off, int group_fastest) {
    using V = vec2_t<T>;
    using v4 = typename Mfma<T>::v4;
    constexpr int D = 1 << K;
    constexpr int R2 = 2 * D;       // real rows = real cols of A
    constexpr int LU = 2 * D + 2;   // padded complex row of U (in T words)
    constexpr int LX = R2 + 2;      // padded real column of X
    constexpr int MT = R2 / 16;     // row tiles
    constexpr int KS = R2 / 4;      // k-steps
    constexpr int NCB = kFusedCols;
    constexpr int EPT = NCB * D / 256;   // amplitudes per thread per tile
    // A fragments live in registers up to k = 5 (MT*KS values per lane);
    // the 64 x 64 matrix of k = 6 is read from LDS.
    constexpr bool AREG = K <= 5;
    __shared__ T su[AREG ? 1 : D * LU];
    __shared__ T sx[NCB * LX];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lj = lane & 15, kk = lane >> 4;
    // A fragment of lane: row r = mt*16 + lj, col c = 4*ks + kk; the real
    // embedding picks Re or Im of U[r/2][c/2] and a sign from the parities.
    const int rpar = lj & 1, cpar = kk & 1;
    const int use_im = rpar != cpar ? 1 : 0;
    const T sgn = (rpar == 0 && cpar == 1) ? T(-1) : T(1);
    T afrag[AREG ? MT * KS : 1];
    if constexpr (AREG) {
#pragma unroll
        for (int mt = 0; mt < MT; ++mt)
#pragma unroll
            for (int ks = 0; ks < KS; ++ks) {
                const int ri = (mt * 16 + lj) >> 1, cj = (4 * ks + kk) >> 1;
                afrag[mt * KS + ks] = sgn * (T)U[2 * (ri * D + cj) + use_im];
            }
    } else {
        for (int e = threadIdx.x; e < D * D; e += 256) {
            const int i = e / D, j = e % D;
            su[i * LU + 2 * j] = (T)U[2 * e];
            su[i * LU + 2 * j + 1] = (T)U[2 * e + 1];
        }
    }

    // element e of a tile -> (column gi, amplitude a)
    auto col_of = [&](int e) { return group_fastest ? (e % NCB) : (e / D); };
    auto amp_of = [&](int e) { return group_fastest ? (e / NCB) : (e % D); };
    auto load_tile = [&](uint64_t tile, V *x) {
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int e = threadIdx.x + 256 * j;
            const uint64_t grp = tile * NCB + col_of(e);
            if (grp < g.ngroups) {
                x[j] = psi[group_base(g, grp) + off.o[amp_of(e)]];
            } else {
                x[j].x = T(0);
                x[j].y = T(0);
            }
        }
    };

    const uint64_t ntiles = (g.ngroups + NCB - 1) / NCB;
    V pre[EPT];
    uint64_t tile = blockIdx.x;
    if (tile < ntiles) load_tile(tile, pre);
    for (; tile < ntiles; tile += gridDim.x) {
        __syncthreads();  // previous tile fully stored
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int e = threadIdx.x + 256 * j;
            sx[col_of(e) * LX + 2 * amp_of(e)] = pre[j].x;
            sx[col_of(e) * LX + 2 * amp_of(e) + 1] = pre[j].y;
        }
        __syncthreads();
        const uint64_t next = tile + gridDim.x;
        if (next < ntiles) load_tile(next, pre);  // in flight during the MFMAs
        v4 acc[MT];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) acc[mt] = v4{0, 0, 0, 0};
        const T *xcol = &sx[(wave * 16 + lj) * LX];
#pragma unroll
        for (int ks = 0; ks < KS; ++ks) {
            const T b = xcol[4 * ks + kk];
#pragma unroll
            for (int mt = 0; mt < MT; ++mt) {
                T a;
                if constexpr (AREG) {
                    a = afrag[mt * KS + ks];
                } else {
                    const int ri = (mt * 16 + lj) >> 1, cj = (4 * ks + kk) >> 1;
                    a = sgn * su[ri * LU + 2 * cj + use_im];
                }
                acc[mt] = Mfma<T>::mma(a, b, acc[mt]);
            }
        }
        // D fragment: register e of this lane is row Mfma<T>::row(kk, e) of
        // column lj.  Each wave only overwrites its own 16 columns, which no
        // other wave reads.
        T *ycol = &sx[(wave * 16 + lj) * LX];
#pragma unroll
        for (int mt = 0; mt < MT; ++mt) {
#pragma unroll
            for (int e = 0; e < 4; ++e) ycol[mt * 16 + Mfma<T>::row(kk, e)] = acc[mt][e];
        }
        __syncthreads();
#pragma unroll
        for (int j = 0; j < EPT; ++j) {
            const int e = threadIdx.x + 256 * j;
            const uint64_t grp = tile * NCB + col_of(e);
            if (grp >= g.ngroups) continue;
            V y;
            y.x = sx[col_of(e) * LX + 2 * amp_of(e)];
            y.y = sx[col_of(e) * LX + 2 * amp_of(e) + 1];
            psi[group_base(g, grp) + off.o[amp_of(e)]] = y;
        }
    }
}

namespace {
template <typename T, int K>
int run_fused(qvmhip_state *s, const void *mat_dev, const uint32_t *qubits) {
    constexpr int D = 1 << K;
    GroupSpec g;
    std::memset(&g, 0, sizeof(g));
    std::vector<uint32_t> pos(qubits, qubits + K);
    std::sort(pos.begin(), pos.end());
    g.nsorted = K;
    for (int i = 0; i < K; ++i) g.sorted[i] = pos[i];
    g.ngroups = uint64_t(1) << (s->n - K);
    FusedOff off;
    std::memset(&off, 0, sizeof(off));
    bool has_q0 = false;
    for (int a = 0; a < D; ++a) {
        uint64_t o = 0;
        for (int b = 0; b < K; ++b)
            if ((a >> b) & 1) o |= uint64_t(1) << qubits[b];
        off.o[a] = o;
    }
    for (int b = 0; b < K; ++b) has_q0 |= qubits[b] == 0;
    const uint64_t ntiles = (g.ngroups + kFusedCols - 1) / kFusedCols;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(ntiles, 256 * 8);
    prof_begin(s, QVMHIP_K_FUSED);
    hipLaunchKernelGGL((k_fused<T, K>), dim3(grid), dim3(256), 0, s->stream, (vec2_t<T> *)s->amps,
                       (const double *)mat_dev, g, off, has_q0 ? 0 : 1);
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
