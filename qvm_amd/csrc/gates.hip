// gates.hip -- gate application kernels (plain gate-by-gate path).
//
// Every gate is one HBM sweep over the amplitudes it can change.  The host
// dispatcher splits a 2^k x 2^k matrix into control qubits (index bits whose
// 0-half is the identity) and a reduced operator on the remaining targets,
// which is one of: identity (no launch), diagonal, permutation, dense.
// Touched amplitudes per gate: dense/diagonal 1q 2^n, CNOT/SWAP 2^(n-1),
// CZ/CPHASE 2^(n-2).
//
// Arithmetic: built with -ffp-contract=off.  Each output is the left-to-right
// sum of separately rounded complex products x_j * U[r][j] with
// (a.re*b.re - a.im*b.im, a.re*b.im + a.im*b.re) -- the arithmetic of the
// reference's serial kernels (src/serial-kernels.lisp:54-67) and generic
// matrix-multiply (src/linear-algebra.lisp:97-131).  Entries the dispatcher
// skips (zero off-diagonals of controlled/diagonal/permutation gates) only
// ever add signed zeros in the reference, so FP64 results agree with it
// exactly.
#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

using cd = std::complex<double>;

// ---------------------------------------------------------------------------
// Kernel argument blocks (passed by value; all < 4 KiB).

template <typename T, int KT>
struct DenseArgs {
    GroupSpec g;
    uint64_t off[1 << KT];
    T ur[(1 << KT) * (1 << KT)];
    T ui[(1 << KT) * (1 << KT)];
};

struct OffTable {
    uint64_t o[64];
};

template <typename T> using V2Ptr = vec2_t<T> *;

template <typename T>
struct DiagFullArgs {
    uint64_t len;
    uint32_t kt;
    uint32_t tq[kMaxK];
    T dr[64];
    T di[64];
};

template <typename T>
struct SparseArgs {
    GroupSpec g;
    uint32_t nent;
    uint64_t off[64];
    uint64_t src[64];  // permutation: source offsets
    T dr[64];
    T di[64];
    // general permutation, entries in cycle order: bit 0 = first entry of a
    // cycle (its destination's old value is kept), bit 1 = last entry (it
    // takes the kept value instead of reading its source)
    uint8_t flag[64];
};

// ---------------------------------------------------------------------------
// Kernels.

// Dense k-qubit operator, k <= 3, one complement group per thread, GPT groups
// per thread kept in flight.  Consecutive threads take consecutive groups, so
// a wavefront reads contiguous runs (pair-strided indexing, coalesced for
// every target).
template <typename T, int KT, int GPT, bool NTM>
__global__ __launch_bounds__(kBlock) void k_dense(vec2_t<T> *__restrict__ psi,
                                                  const DenseArgs<T, KT> a) {
    constexpr int D = 1 << KT;
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * (kBlock * GPT);
    for (uint64_t first = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * (kBlock * GPT) + threadIdx.x;
         first < a.g.ngroups;
         first += step) {
        V x[GPT][D];
        uint64_t base[GPT];
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
            const uint64_t i = first + (uint64_t)g * kBlock;
            base[g] = group_base(a.g, i < a.g.ngroups ? i : 0);
            if (i < a.g.ngroups) {
#pragma unroll
                for (int j = 0; j < D; ++j) x[g][j] = ldg<NTM>(&psi[base[g] + a.off[j]]);
            }
        }
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
            const uint64_t i = first + (uint64_t)g * kBlock;
            if (i >= a.g.ngroups) continue;
#pragma unroll
            for (int r = 0; r < D; ++r) {
                T yr = x[g][0].x * a.ur[r * D] - x[g][0].y * a.ui[r * D];
                T yi = x[g][0].x * a.ui[r * D] + x[g][0].y * a.ur[r * D];
#pragma unroll
                for (int c = 1; c < D; ++c) {
                    const T pr = x[g][c].x * a.ur[r * D + c] - x[g][c].y * a.ui[r * D + c];
                    const T pi = x[g][c].x * a.ui[r * D + c] + x[g][c].y * a.ur[r * D + c];
                    yr = yr + pr;
                    yi = yi + pi;
                }
                V y;
                y.x = yr;
                y.y = yi;
                stg<NTM>(&psi[base[g] + a.off[r]], y);
            }
        }
    }
}

// Uncontrolled 1-qubit dense operators on qubit q, one amplitude per thread
// with non-temporal accesses, so that each wavefront reads and writes one
// contiguous run: on MI355X these in-place sweeps run at 6.6-6.8 TB/s where
// the pair-per-thread form of k_dense reaches 5.5-5.9 TB/s for q < 8.  The
// output of the amplitude with bit q = r is row r of the matrix applied to
// (lower member, upper member), with k_dense's products and sums in k_dense's
// order, so the results are bit-identical.
template <typename T>
struct Dense1Coef {
    T ur[4], ui[4];  // row-major 2x2
};

template <typename T>
__device__ __forceinline__ vec2_t<T> dense1_row(const vec2_t<T> &lo, const vec2_t<T> &hi, uint32_t r,
                                                const Dense1Coef<T> &u) {
    const T u0r = r ? u.ur[2] : u.ur[0], u0i = r ? u.ui[2] : u.ui[0];
    const T u1r = r ? u.ur[3] : u.ur[1], u1i = r ? u.ui[3] : u.ui[1];
    T yr = lo.x * u0r - lo.y * u0i;
    T yi = lo.x * u0i + lo.y * u0r;
    const T pr = hi.x * u1r - hi.y * u1i;
    const T pi = hi.x * u1i + hi.y * u1r;
    vec2_t<T> y;
    y.x = yr + pr;
    y.y = yi + pi;
    return y;
}

// q < 6: the partner (index i ^ 2^q) is in the same wavefront; a lane shuffle
// brings it.  len is a multiple of kBlock, so every lane of every wave is live.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense1_lane(vec2_t<T> *__restrict__ psi, uint64_t len, uint32_t q,
                                                        const Dense1Coef<T> u) {
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; i < len; i += step) {
        const V x = ldg<true>(&psi[i]);
        V p;
        p.x = __shfl_xor(x.x, 1 << q, 64);
        p.y = __shfl_xor(x.y, 1 << q, 64);
        const uint32_t r = (uint32_t)(i >> q) & 1u;
        stg<true>(&psi[i], dense1_row<T>(r ? p : x, r ? x : p, r, u));
    }
}

// FP32, q < 7: two consecutive amplitudes per lane (one 16-byte access, 1 KB
// per wavefront; one 8-byte amplitude per lane gives 512 B accesses, which
// sweep slower).  Element j holds amplitudes 2j and 2j + 1: for q = 0 they
// are the pair; otherwise each pairs with the same half of lane j ^ 2^(q-1).
// n4 (elements) is a multiple of kBlock, so every lane of every wave is live.
__global__ __launch_bounds__(kBlock) void k_dense1_lane2_f32(float2 *__restrict__ psi, uint64_t n4, uint32_t q,
                                                             const Dense1Coef<float> u) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    F4 *p4 = reinterpret_cast<F4 *>(psi);
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t j = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < n4; j += step) {
        const F4 v = __builtin_nontemporal_load(&p4[j]);
        const float2 a0 = make_float2(v.x, v.y), a1 = make_float2(v.z, v.w);
        float2 y0, y1;
        if (q == 0) {
            y0 = dense1_row<float>(a0, a1, 0u, u);
            y1 = dense1_row<float>(a0, a1, 1u, u);
        } else {
            const int d = 1 << (q - 1);
            float2 p0, p1;
            p0.x = __shfl_xor(a0.x, d, 64);
            p0.y = __shfl_xor(a0.y, d, 64);
            p1.x = __shfl_xor(a1.x, d, 64);
            p1.y = __shfl_xor(a1.y, d, 64);
            const uint32_t r = (uint32_t)(j >> (q - 1)) & 1u;
            y0 = dense1_row<float>(r ? p0 : a0, r ? a0 : p0, r, u);
            y1 = dense1_row<float>(r ? p1 : a1, r ? a1 : p1, r, u);
        }
        F4 o;
        o.x = y0.x;
        o.y = y0.y;
        o.z = y1.x;
        o.w = y1.y;
        __builtin_nontemporal_store(o, &p4[j]);
    }
}

// q >= 6 (any target up to n - 1): a workgroup takes 128 pairs; waves 0 and
// 2 load 64 lower members each (contiguous), waves 1 and 3 the matching upper
// members, and the partners are exchanged through LDS.  npairs is a multiple
// of 128.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dense1_xchg(vec2_t<T> *__restrict__ psi, uint64_t npairs, uint32_t q,
                                                        const Dense1Coef<T> u) {
    using V = vec2_t<T>;
    __shared__ V sm[kBlock];
    const uint32_t t = threadIdx.x, h = (t >> 6) & 1u;
    const uint64_t nchunk = npairs / 128;
    for (uint64_t c = xcd_block(blockIdx.x, gridDim.x); c < nchunk; c += gridDim.x) {
        const uint64_t pair = c * 128 + ((t >> 7) << 6) + (t & 63u);
        const uint64_t a = insert_zero(pair, q) | ((uint64_t)h << q);
        const V x = ldg<true>(&psi[a]);
        sm[t] = x;
        __syncthreads();
        const V p = sm[t ^ 64u];
        __syncthreads();  // sm is rewritten by the next chunk
        stg<true>(&psi[a], dense1_row<T>(h ? p : x, h ? x : p, h, u));
    }
}

// Dense k-qubit operator, 4 <= k <= 6: the matrix is staged in LDS
// (transposed so that the D lanes of a group read consecutive words), each
// group's D amplitudes are staged in LDS, lane r of the group produces row r.
template <typename T, int KT>
__global__ __launch_bounds__(kBlock) void k_dense_big(vec2_t<T> *__restrict__ psi,
                                                      const T *__restrict__ ut, GroupSpec g,
                                                      OffTable off) {
    constexpr int D = 1 << KT;
    constexpr int GPB = kBlock / D;
    using V = vec2_t<T>;
    __shared__ T sre[D * D];
    __shared__ T sim[D * D];
    __shared__ V sx[kBlock];
    for (int i = threadIdx.x; i < D * D; i += kBlock) {
        sre[i] = ut[i];
        sim[i] = ut[D * D + i];
    }
    const int r = threadIdx.x % D;
    const int gl = threadIdx.x / D;
    const uint64_t nbatch = (g.ngroups + GPB - 1) / GPB;
    for (uint64_t b = xcd_block(blockIdx.x, gridDim.x); b < nbatch; b += gridDim.x) {
        const uint64_t grp = b * GPB + gl;
        const bool ok = grp < g.ngroups;
        uint64_t addr = 0;
        __syncthreads();
        if (ok) {
            addr = group_base(g, grp) + off.o[r];
            sx[threadIdx.x] = psi[addr];
        }
        __syncthreads();
        if (ok) {
            const V *x = &sx[gl * D];
            T yr = x[0].x * sre[r] - x[0].y * sim[r];
            T yi = x[0].x * sim[r] + x[0].y * sre[r];
            for (int c = 1; c < D; ++c) {
                const T mr = sre[c * D + r], mi = sim[c * D + r];
                const T pr = x[c].x * mr - x[c].y * mi;
                const T pi = x[c].x * mi + x[c].y * mr;
                yr = yr + pr;
                yi = yi + pi;
            }
            V y;
            y.x = yr;
            y.y = yi;
            psi[addr] = y;
        }
    }
}

// Diagonal operator with every entry != 1 and no controls: contiguous sweep,
// amplitude i is multiplied by d[bits of i at the targets].  KT = 1..3: the
// entry is chosen from the (scalar) argument block by a compile-time chain
// of selects; KT = 0: any width, runtime loop and indexed entry.
// One amplitude per thread; the loop is written so that the store keeps its
// non-temporal flag (a conditional store inside an unrolled per-thread batch
// lost it).
template <typename T, int KT>
__device__ __forceinline__ void diag_entry(const DiagFullArgs<T> &a, uint64_t i, T &mr, T &mi) {
    if constexpr (KT > 0) {
        uint32_t j = 0;
#pragma unroll
        for (int t = 0; t < KT; ++t) j |= (uint32_t)((i >> a.tq[t]) & 1u) << t;
        mr = a.dr[0];
        mi = a.di[0];
#pragma unroll
        for (uint32_t e = 1; e < (1u << KT); ++e) {
            mr = j == e ? a.dr[e] : mr;
            mi = j == e ? a.di[e] : mi;
        }
    } else {
        uint32_t j = 0;
        for (uint32_t t = 0; t < a.kt; ++t) j |= (uint32_t)((i >> a.tq[t]) & 1u) << t;
        mr = a.dr[j];
        mi = a.di[j];
    }
}

// FP32: two amplitudes per thread as one 16-byte access (the FP64 kernel's
// one-per-lane form would make 512 B wavefront accesses).  len is even.
template <int KT>
__global__ __launch_bounds__(kBlock) void k_diag_full_f32x2(float2 *__restrict__ psi, const DiagFullArgs<float> a) {
    typedef float f4 __attribute__((ext_vector_type(4)));
    f4 *p4 = (f4 *)psi;
    const uint64_t n2 = a.len / 2;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t k = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; k < n2; k += step) {
        const f4 x = __builtin_nontemporal_load(&p4[k]);
        float m0r, m0i, m1r, m1i;
        diag_entry<float, KT>(a, 2 * k, m0r, m0i);
        diag_entry<float, KT>(a, 2 * k + 1, m1r, m1i);
        f4 y;
        y.x = x.x * m0r - x.y * m0i;
        y.y = x.x * m0i + x.y * m0r;
        y.z = x.z * m1r - x.w * m1i;
        y.w = x.z * m1i + x.w * m1r;
        __builtin_nontemporal_store(y, &p4[k]);
    }
}

template <typename T, bool NTM, int KT>
__global__ __launch_bounds__(kBlock) void k_diag_full(vec2_t<T> *__restrict__ psi,
                                                      const DiagFullArgs<T> a) {
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; i < a.len; i += step) {
        const V x = ldg<NTM>(&psi[i]);
        T mr, mi;
        diag_entry<T, KT>(a, i, mr, mi);
        V y;
        y.x = x.x * mr - x.y * mi;
        y.y = x.x * mi + x.y * mr;
        stg<NTM>(&psi[i], y);
    }
}

// Sparse diagonal: only the nent listed entries of each group change
// (controlled phases: CZ, CPHASE, PHASE, CCPHASE ...).
// ONE: a single changed entry per group (CZ, CPHASE, PHASE, CCPHASE ...),
// read from the argument block with static indices.
template <typename T, int GPT, bool NTM, bool ONE>
__global__ __launch_bounds__(kBlock) void k_diag_sparse(vec2_t<T> *__restrict__ psi,
                                                        const SparseArgs<T> a) {
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * (kBlock * GPT);
    for (uint64_t first = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * (kBlock * GPT) + threadIdx.x;
         first < a.g.ngroups;
         first += step) {
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
            const uint64_t i = first + (uint64_t)g * kBlock;
            if (i >= a.g.ngroups) continue;
            const uint64_t base = group_base(a.g, i);
            if constexpr (ONE) {
                const V x = ldg<NTM>(&psi[base + a.off[0]]);
                V y;
                y.x = x.x * a.dr[0] - x.y * a.di[0];
                y.y = x.x * a.di[0] + x.y * a.dr[0];
                stg<NTM>(&psi[base + a.off[0]], y);
                continue;
            }
            for (uint32_t e = 0; e < a.nent; ++e) {
                const V x = ldg<NTM>(&psi[base + a.off[e]]);
                V y;
                y.x = x.x * a.dr[e] - x.y * a.di[e];
                y.y = x.x * a.di[e] + x.y * a.dr[e];
                stg<NTM>(&psi[base + a.off[e]], y);
            }
        }
    }
}

// Permutation: psi'[dst_e] = psi[src_e] for the moved entries of each group
// (CNOT, SWAP, CCNOT, CSWAP ...): pure data movement, no arithmetic.
// TWO: exactly one transposition per group (CNOT, SWAP, CCNOT, CSWAP).
template <typename T, int GPT, bool NTM, bool TWO>
__global__ __launch_bounds__(kBlock) void k_perm(vec2_t<T> *__restrict__ psi,
                                                 const SparseArgs<T> a) {
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * (kBlock * GPT);
    for (uint64_t first = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * (kBlock * GPT) + threadIdx.x;
         first < a.g.ngroups;
         first += step) {
#pragma unroll
        for (int g = 0; g < GPT; ++g) {
            const uint64_t i = first + (uint64_t)g * kBlock;
            if (i >= a.g.ngroups) continue;
            const uint64_t base = group_base(a.g, i);
            if constexpr (TWO) {
                const V x0 = ldg<NTM>(&psi[base + a.src[0]]);
                const V x1 = ldg<NTM>(&psi[base + a.src[1]]);
                stg<NTM>(&psi[base + a.off[0]], x0);
                stg<NTM>(&psi[base + a.off[1]], x1);
            } else {
                // cycle by cycle with one kept value: within a cycle each
                // destination's source is the next entry's destination, read
                // before that entry overwrites it
                // (the source is read even where the kept value replaces
                // it: a choice between two loads would be merged into one
                // flat load through a scratch copy of kept)
                V kept;
                kept.x = kept.y = T(0);
                for (uint32_t e = 0; e < a.nent; ++e) {
                    const uint32_t f = a.flag[e];
                    if (f & 1u) kept = ldg<NTM>(&psi[base + a.off[e]]);
                    const V sv = ldg<NTM>(&psi[base + a.src[e]]);
                    V v;
                    v.x = (f & 2u) ? kept.x : sv.x;
                    v.y = (f & 2u) ? kept.y : sv.y;
                    stg<NTM>(&psi[base + a.off[e]], v);
                }
            }
        }
    }
}

// A single transposition (CNOT, SWAP, CCNOT, CSWAP) whose target and control
// bits are all below 6 (and the lowest below 3): every amplitude is in the
// same wavefront as its partner, so the kernel sweeps the whole state one
// amplitude per thread and takes the partner through a lane shuffle.  Amplitude i is replaced by
// amplitude i ^ flip when its target bits (mask tmask) equal o0 or o1 and its
// controls are set; the others are written back unchanged, which keeps every
// store a full contiguous run (the pair-per-group form touches half of every
// line when the bits are this low).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_lane(vec2_t<T> *__restrict__ psi, uint64_t len, uint32_t tmask,
                                                      uint32_t o0, uint32_t o1, uint32_t ctrl, uint32_t masked) {
    using V = vec2_t<T>;
    const uint32_t flip = o0 ^ o1;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; i < len; i += step) {
        const V x = ldg<true>(&psi[i]);
        V p;
        p.x = __shfl_xor(x.x, (int)flip, 64);
        p.y = __shfl_xor(x.y, (int)flip, 64);
        const uint32_t lo = (uint32_t)i & 63u;
        const bool moved = ((lo & tmask) == o0 || (lo & tmask) == o1) && (lo & ctrl) == ctrl;
        if (moved || !masked) stg<true>(&psi[i], moved ? p : x);
    }
}

// A single transposition E0 <-> E1 whose bits are all at or above 6 (CNOT,
// SWAP, CCNOT, CSWAP on high qubits), in the layout of k_dense1_xchg: a
// workgroup takes 128 consecutive groups, waves 0 and 2 load the E0 members
// of 64 groups each (a contiguous run), waves 1 and 3 the E1 members, the
// partners are exchanged through LDS and each thread stores its partner's
// amplitude at its own address.  Copies only, so results are exact.
// ngroups is a multiple of 128.  Measured at n = 30 against the pair form
// (k_perm): CNOT(q, q + 7) for q = 6, 8, 12, 15, 20, 22 2.54-2.89 ms against
// 2.59-3.01, SWAP(q, 29 - q) for q = 8, 12 2.58 / 2.62 against 2.77 / 2.80.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_xchg(vec2_t<T> *__restrict__ psi, const GroupSpec g, uint64_t e0,
                                                      uint64_t e1) {
    using V = vec2_t<T>;
    __shared__ V sm[kBlock];
    const uint32_t t = threadIdx.x, h = (t >> 6) & 1u;
    const uint64_t nchunk = g.ngroups / 128;
    for (uint64_t c = xcd_block(blockIdx.x, gridDim.x); c < nchunk; c += gridDim.x) {
        const uint64_t grp = c * 128 + ((t >> 7) << 6) + (t & 63u);
        const uint64_t a = group_base(g, grp) | (h ? e1 : e0);
        sm[t] = ldg<true>(&psi[a]);
        __syncthreads();
        const V p = sm[t ^ 64u];
        __syncthreads();  // sm is rewritten by the next chunk
        stg<true>(&psi[a], p);
    }
}

// Gates whose lowest bit lies inside a 128 B line touch every line of the
// lines they change but only some amplitudes of each, so the pair and entry
// forms above read and write partial lines.  The line forms sweep whole
// lines instead.  The gate's bits split into those inside the line (mask
// lmask) and those above it; `hg` enumerates the amplitudes whose bits above
// the line are zero (count = 2^(n - their number)), and ORing h0 (h1) selects
// the lines of entry E0 (E1).  One amplitude per thread, consecutive threads
// on consecutive amplitudes of a line; count is a multiple of kBlock, so every
// lane of a wavefront takes part in each shuffle.
//
// FP32 form of k_perm_lane with every bit below 7: two consecutive amplitudes
// per lane as one 16-byte access (element j holds amplitudes 2j, 2j + 1); the
// partner of amplitude 2j + e is in lane j ^ (flip >> 1), in the other half
// when flip has bit 0.  n4 (elements) is a multiple of kBlock.
__global__ __launch_bounds__(kBlock) void k_perm_lane2_f32(float2 *__restrict__ psi, uint64_t n4, uint32_t tmask,
                                                           uint32_t o0, uint32_t o1, uint32_t ctrl) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    F4 *p4 = reinterpret_cast<F4 *>(psi);
    const uint32_t flip = o0 ^ o1;
    const int d = (int)(flip >> 1);
    const bool odd = (flip & 1u) != 0;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t j = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < n4; j += step) {
        const F4 v = __builtin_nontemporal_load(&p4[j]);
        const float s0x = __shfl_xor(v.x, d, 64), s0y = __shfl_xor(v.y, d, 64);
        const float s1x = __shfl_xor(v.z, d, 64), s1y = __shfl_xor(v.w, d, 64);
        const uint32_t lo0 = (uint32_t)(j << 1) & 127u, lo1 = lo0 | 1u;
        const bool m0 = ((lo0 & tmask) == o0 || (lo0 & tmask) == o1) && (lo0 & ctrl) == ctrl;
        const bool m1 = ((lo1 & tmask) == o0 || (lo1 & tmask) == o1) && (lo1 & ctrl) == ctrl;
        F4 o = v;
        if (m0) {
            o.x = odd ? s1x : s0x;
            o.y = odd ? s1y : s0y;
        }
        if (m1) {
            o.z = odd ? s0x : s1x;
            o.w = odd ? s0y : s1y;
        }
        __builtin_nontemporal_store(o, &p4[j]);
    }
}

// Single-entry diagonal (CZ, CPHASE with a bit inside the line): every
// amplitude of the lines with the entry's upper bits is read, the ones whose
// in-line bits match (l0) are multiplied as in k_diag_sparse, and the lines
// are written back whole, or with `masked` only the changed amplitudes: the
// caller sets it when they form 64-byte halves of each line (in-line bit 2
// only; n = 30: CPHASE(29, 2) 2.57 -> 2.11 ms), since 16- and 32-byte pieces
// wrote slower than whole lines (CPHASE(9, 0) 2.61 -> 4.26 ms).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_line(vec2_t<T> *__restrict__ psi, const GroupSpec hg,
                                                      uint64_t count, uint64_t h0, uint32_t lmask, uint32_t l0,
                                                      T dr, T di, uint32_t masked) {
    using V = vec2_t<T>;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t j = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < count; j += step) {
        const uint64_t i = group_base(hg, j) | h0;
        const V x = ldg<true>(&psi[i]);
        V y = x;
        const bool hit = ((uint32_t)i & lmask) == l0;
        if (hit) {
            y.x = x.x * dr - x.y * di;
            y.y = x.x * di + x.y * dr;
        }
        if (hit || !masked) stg<true>(&psi[i], y);
    }
}

// FP32 single-entry sparse diagonal whose bits are all >= 1: the changed
// amplitudes come in pairs (2j, 2j + 1), so the kernel works on 16-byte pairs
// (g2: the gate's bits and controls shifted down by one over 2^(n-1) pairs,
// off2 the entry's offset in pairs), each thread one pair, non-temporal (the
// caller uses it for runs of >= 128 B); the products are those of
// k_diag_sparse.
__global__ __launch_bounds__(kBlock) void k_diag_sparse2_f32(float2 *__restrict__ psi, const GroupSpec g2,
                                                             uint64_t off2, float dr, float di) {
    typedef float F4 __attribute__((ext_vector_type(4)));
    F4 *p4 = reinterpret_cast<F4 *>(psi);
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; i < g2.ngroups;
         i += step) {
        const uint64_t e = group_base(g2, i) | off2;
        const F4 v = __builtin_nontemporal_load(&p4[e]);
        F4 y;
        y.x = v.x * dr - v.y * di;
        y.y = v.x * di + v.y * dr;
        y.z = v.z * dr - v.w * di;
        y.w = v.z * di + v.w * dr;
        __builtin_nontemporal_store(y, &p4[e]);
    }
}

// Single transposition E0 <-> E1 (CNOT, SWAP, CCNOT, CSWAP) with a bit inside
// the line and one at or above bit 6 (the all-lane case is k_perm_lane).  The
// moving amplitude at in-line position l of line A (upper bits h0) pairs with
// position l ^ (l0 ^ l1) of line B (upper bits h1): a lane shuffle.  With
// h0 == h1 both entries lie in the same lines.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_perm_line(vec2_t<T> *__restrict__ psi, const GroupSpec hg,
                                                      uint64_t count, uint64_t h0, uint64_t h1, uint32_t lmask,
                                                      uint32_t l0, uint32_t l1) {
    using V = vec2_t<T>;
    const int d = (int)(l0 ^ l1);
    const bool two = h0 != h1;
    const uint64_t step = (uint64_t)gridDim.x * kBlock;
    for (uint64_t j = (uint64_t)xcd_block(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < count; j += step) {
        const uint64_t g = group_base(hg, j);
        const uint64_t ia = g | h0, ib = g | h1;
        const V xa = ldg<true>(&psi[ia]);
        const V xb = two ? ldg<true>(&psi[ib]) : xa;
        V pa, pb;
        pa.x = __shfl_xor(xa.x, d, 64);
        pa.y = __shfl_xor(xa.y, d, 64);
        pb.x = __shfl_xor(xb.x, d, 64);
        pb.y = __shfl_xor(xb.y, d, 64);
        const uint32_t ls = (uint32_t)ia & lmask;
        if (two) {
            stg<true>(&psi[ia], ls == l0 ? pb : xa);
            stg<true>(&psi[ib], ls == l1 ? pa : xb);
        } else {
            stg<true>(&psi[ia], (ls == l0 || ls == l1) ? pa : xa);
        }
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill_zero_state(vec2_t<T> *__restrict__ psi,
                                                            uint64_t len) {
    using V = vec2_t<T>;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
        V z;
        z.x = (i == 0) ? T(1) : T(0);
        z.y = T(0);
        psi[i] = z;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_scale(vec2_t<T> *__restrict__ psi, uint64_t len,
                                                  T f) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
        vec2_t<T> v = psi[i];
        v.x = f * v.x;
        v.y = f * v.y;
        psi[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Host side: structure detection and dispatch.

namespace {

struct GateSplit {
    int kind = 0;  // 0 identity, 1 dense, 2 diagonal, 3 permutation
    uint32_t kt = 0;
    uint32_t tq[kMaxK] = {0};  // target qubits, reduced bit j -> tq[j]
    uint32_t nctrl = 0;
    uint64_t ctrl_mask = 0;
    std::vector<cd> V;             // reduced matrix, row-major D' x D'
    std::vector<uint32_t> perm;    // permutation: row j takes column perm[j]
};

bool exactly(cd z, double re, double im) { return z.real() == re && z.imag() == im; }

GateSplit split_gate(const cd *U, uint32_t k, const uint32_t *qubits, bool detect) {
    const uint32_t D = 1u << k;
    GateSplit s;
    uint32_t ctrl_bits = 0;
    if (detect) {
        for (uint32_t c = 0; c < k; ++c) {
            bool is_ctrl = true;
            for (uint32_t r = 0; r < D && is_ctrl; ++r) {
                if ((r >> c) & 1u) continue;
                for (uint32_t t = 0; t < D; ++t) {
                    const double want = (t == r) ? 1.0 : 0.0;
                    if (!exactly(U[r * D + t], want, 0.0) || !exactly(U[t * D + r], want, 0.0)) {
                        is_ctrl = false;
                        break;
                    }
                }
            }
            if (is_ctrl) ctrl_bits |= 1u << c;
        }
    }
    for (uint32_t b = 0; b < k; ++b) {
        if ((ctrl_bits >> b) & 1u) {
            s.ctrl_mask |= uint64_t(1) << qubits[b];
            ++s.nctrl;
        } else {
            s.tq[s.kt++] = qubits[b];
        }
    }
    // reduced index j -> full index
    const uint32_t Dr = 1u << s.kt;
    std::vector<uint32_t> full(Dr);
    for (uint32_t j = 0; j < Dr; ++j) {
        uint32_t f = ctrl_bits, t = 0;
        for (uint32_t b = 0; b < k; ++b) {
            if ((ctrl_bits >> b) & 1u) continue;
            if ((j >> t) & 1u) f |= 1u << b;
            ++t;
        }
        full[j] = f;
    }
    s.V.resize((size_t)Dr * Dr);
    for (uint32_t r = 0; r < Dr; ++r)
        for (uint32_t c = 0; c < Dr; ++c) s.V[(size_t)r * Dr + c] = U[(size_t)full[r] * D + full[c]];
    if (!detect) {
        s.kind = 1;
        return s;
    }
    bool diag = true, perm = true, ident = true;
    s.perm.assign(Dr, 0);
    for (uint32_t r = 0; r < Dr; ++r) {
        int ones = 0, nonzero = 0;
        for (uint32_t c = 0; c < Dr; ++c) {
            const cd z = s.V[(size_t)r * Dr + c];
            if (!exactly(z, 0.0, 0.0)) {
                ++nonzero;
                if (c != r) diag = false;
                if (exactly(z, 1.0, 0.0)) {
                    ++ones;
                    s.perm[r] = c;
                }
                if (c != r || !exactly(z, 1.0, 0.0)) ident = false;
            }
        }
        if (nonzero == 0) {
            diag = false;  // singular row: not diagonal-with-phases, use dense
            ident = false;
        }
        if (!(nonzero == 1 && ones == 1)) perm = false;
    }
    if (perm) {
        // a permutation gate moves every amplitude somewhere: rows taking the
        // same column (a singular 0/1 matrix) go to the dense kernel
        std::vector<bool> used(Dr, false);
        for (uint32_t r = 0; r < Dr && perm; ++r) {
            if (used[s.perm[r]]) perm = false;
            used[s.perm[r]] = true;
        }
    }
    if (ident)
        s.kind = 0;
    else if (diag)
        s.kind = 2;
    else if (perm)
        s.kind = 3;
    else
        s.kind = 1;
    return s;
}

void make_group_spec(GroupSpec &g, const GateSplit &s, uint32_t n) {
    std::vector<uint32_t> pos(s.tq, s.tq + s.kt);
    for (uint32_t b = 0; b < 64; ++b)
        if ((s.ctrl_mask >> b) & 1u) pos.push_back(b);
    std::sort(pos.begin(), pos.end());
    std::memset(&g, 0, sizeof(g));
    g.nsorted = (uint32_t)pos.size();
    for (size_t i = 0; i < pos.size(); ++i) g.sorted[i] = pos[i];
    g.ctrl_mask = s.ctrl_mask;
    g.ngroups = uint64_t(1) << (n - g.nsorted);
}

uint64_t deposit(uint32_t j, const uint32_t *tq, uint32_t kt) {
    uint64_t o = 0;
    for (uint32_t t = 0; t < kt; ++t)
        if ((j >> t) & 1u) o |= uint64_t(1) << tq[t];
    return o;
}

uint32_t grid_for(uint64_t work, uint64_t per_block) {
    uint64_t b = (work + per_block - 1) / per_block;
    return (uint32_t)std::min<uint64_t>(std::max<uint64_t>(b, 1), kMaxGrid);
}

// The gate's bits above the 128 B line as a GroupSpec of their own (no
// controls), and the mask of its bits inside the line, for the line forms.
struct LineSplit {
    GroupSpec hg;
    uint64_t hmask;
    uint32_t lmask;
};
LineSplit line_split(const GroupSpec &g, uint32_t n, size_t amp_bytes) {
    const uint32_t line_bits = amp_bytes == 16 ? 3u : 4u;  // 128 B lines
    LineSplit ls;
    std::memset(&ls, 0, sizeof(ls));
    for (uint32_t i = 0; i < g.nsorted; ++i) {
        if (g.sorted[i] < line_bits) {
            ls.lmask |= 1u << g.sorted[i];
        } else {
            ls.hg.sorted[ls.hg.nsorted++] = g.sorted[i];
            ls.hmask |= uint64_t(1) << g.sorted[i];
        }
    }
    ls.hg.ngroups = uint64_t(1) << (n - ls.hg.nsorted);
    return ls;
}

// Non-temporal access pays off when every contiguous run of the gate's
// access pattern spans >= 128 B: the lowest target/control qubit q gives
// runs of 2^q amplitudes.
bool long_runs(const GroupSpec &g, size_t amp_bytes) {
    const uint32_t qmin = g.nsorted ? g.sorted[0] : 63;
    return qmin >= 16 || ((uint64_t(1) << qmin) * amp_bytes >= 128);
}

template <typename T, int KT>
int run_dense_small(qvmhip_state *s, const GateSplit &gs) {
    constexpr int D = 1 << KT;
    constexpr int GPT = 1;
    DenseArgs<T, KT> a;
    std::memset(&a, 0, sizeof(a));
    make_group_spec(a.g, gs, s->n);
    for (int j = 0; j < D; ++j) a.off[j] = deposit(j, gs.tq, gs.kt);
    for (int i = 0; i < D * D; ++i) {
        a.ur[i] = (T)gs.V[i].real();
        a.ui[i] = (T)gs.V[i].imag();
    }
    if constexpr (KT == 1 && sizeof(T) == 8) {
        // FP64 only: with 8-byte FP32 amplitudes one per lane is a 512 B
        // wavefront access, slower than the pair form (5.8 vs 6.6 TB/s)
        if (a.g.ctrl_mask == 0 && s->n >= 8) {
            // one amplitude per thread: k_dense1_lane below 6, k_dense1_xchg
            // from 6 up.  With XCD-contiguous workgroups the exchange form
            // beat the pair form on every target 8..29 at n = 30 (5.08-5.84
            // against 5.18-6.11 ms, 4 % less on average)
            Dense1Coef<T> u;
            for (int i = 0; i < 4; ++i) {
                u.ur[i] = (T)gs.V[i].real();
                u.ui[i] = (T)gs.V[i].imag();
            }
            const uint32_t q = gs.tq[0];
            prof_begin(s, KT);
            if (q < 6)
                hipLaunchKernelGGL((k_dense1_lane<T>), dim3(grid_for(s->len, kBlock)), dim3(kBlock), 0, s->stream,
                                   (vec2_t<T> *)s->amps, s->len, q, u);
            else
                hipLaunchKernelGGL((k_dense1_xchg<T>), dim3(grid_for(s->len / 2, 128)), dim3(kBlock), 0, s->stream,
                                   (vec2_t<T> *)s->amps, s->len / 2, q, u);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    if constexpr (KT == 1 && sizeof(T) == 4) {
        if (a.g.ctrl_mask == 0 && s->n >= 9 && gs.tq[0] < 7) {
            Dense1Coef<float> u;
            for (int i = 0; i < 4; ++i) {
                u.ur[i] = (float)gs.V[i].real();
                u.ui[i] = (float)gs.V[i].imag();
            }
            prof_begin(s, KT);
            hipLaunchKernelGGL(k_dense1_lane2_f32, dim3(grid_for(s->len / 2, kBlock)), dim3(kBlock), 0, s->stream,
                               (float2 *)s->amps, s->len / 2, gs.tq[0], u);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    const dim3 grid(grid_for(a.g.ngroups, (uint64_t)kBlock * GPT));
    prof_begin(s, KT);
    if (long_runs(a.g, s->amp_bytes))
        hipLaunchKernelGGL((k_dense<T, KT, GPT, true>), grid, dim3(kBlock), 0, s->stream, (vec2_t<T> *)s->amps, a);
    else
        hipLaunchKernelGGL((k_dense<T, KT, GPT, false>), grid, dim3(kBlock), 0, s->stream, (vec2_t<T> *)s->amps, a);
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T, int KT>
int run_dense_big(qvmhip_state *s, const GateSplit &gs) {
    constexpr int D = 1 << KT;
    GroupSpec g;
    make_group_spec(g, gs, s->n);
    OffTable off;
    std::memset(&off, 0, sizeof(off));
    for (int j = 0; j < D; ++j) off.o[j] = deposit(j, gs.tq, gs.kt);
    // transposed matrix: [re | im], element (r, c) at c*D + r
    std::vector<T> host(2 * D * D);
    for (int r = 0; r < D; ++r)
        for (int c = 0; c < D; ++c) {
            host[c * D + r] = (T)gs.V[r * D + c].real();
            host[D * D + c * D + r] = (T)gs.V[r * D + c].imag();
        }
    QH_CHECK(hipMemcpyAsync(s->mat, host.data(), host.size() * sizeof(T), hipMemcpyHostToDevice,
                            s->stream));
    const uint64_t gpb = kBlock / D;
    const uint64_t nbatch = (g.ngroups + gpb - 1) / gpb;
    const uint32_t grid = (uint32_t)std::min<uint64_t>(nbatch, 256 * 16);
    prof_begin(s, QVMHIP_K_DENSE_BIG);
    hipLaunchKernelGGL((k_dense_big<T, KT>), dim3(grid), dim3(kBlock), 0, s->stream,
                       (vec2_t<T> *)s->amps, (const T *)s->mat, g, off);
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T>
int run_diag(qvmhip_state *s, const GateSplit &gs) {
    const uint32_t Dr = 1u << gs.kt;
    std::vector<uint32_t> ent;
    for (uint32_t j = 0; j < Dr; ++j)
        if (!exactly(gs.V[(size_t)j * Dr + j], 1.0, 0.0)) ent.push_back(j);
    if (ent.empty()) return QVMHIP_OK;
    if (gs.nctrl == 0 && ent.size() == Dr) {
        DiagFullArgs<T> a;
        std::memset(&a, 0, sizeof(a));
        a.len = s->len;
        a.kt = gs.kt;
        for (uint32_t t = 0; t < gs.kt; ++t) a.tq[t] = gs.tq[t];
        for (uint32_t j = 0; j < Dr; ++j) {
            a.dr[j] = (T)gs.V[(size_t)j * Dr + j].real();
            a.di[j] = (T)gs.V[(size_t)j * Dr + j].imag();
        }
        // one amplitude per thread and no per-lane loads of the entries:
        // measured 6.6 TB/s on MI355X for the in-place sweep, against
        // 5.9 TB/s with four amplitudes per thread and indexed entries
        V2Ptr<T> psi = (vec2_t<T> *)s->amps;
        if constexpr (sizeof(T) == 4) {
            if (s->len >= 2) {
                const dim3 g2(grid_for(a.len / 2, (uint64_t)kBlock));
                prof_begin(s, QVMHIP_K_DIAG_FULL);
                switch (gs.kt) {
                    case 1: hipLaunchKernelGGL((k_diag_full_f32x2<1>), g2, dim3(kBlock), 0, s->stream, psi, a); break;
                    case 2: hipLaunchKernelGGL((k_diag_full_f32x2<2>), g2, dim3(kBlock), 0, s->stream, psi, a); break;
                    case 3: hipLaunchKernelGGL((k_diag_full_f32x2<3>), g2, dim3(kBlock), 0, s->stream, psi, a); break;
                    default: hipLaunchKernelGGL((k_diag_full_f32x2<0>), g2, dim3(kBlock), 0, s->stream, psi, a);
                }
                prof_end(s);
                QH_LAUNCH_CHECK();
                return QVMHIP_OK;
            }
        }
        const dim3 grid(grid_for(a.len, (uint64_t)kBlock));
        prof_begin(s, QVMHIP_K_DIAG_FULL);
        switch (gs.kt) {
            case 1: hipLaunchKernelGGL((k_diag_full<T, true, 1>), grid, dim3(kBlock), 0, s->stream, psi, a); break;
            case 2: hipLaunchKernelGGL((k_diag_full<T, true, 2>), grid, dim3(kBlock), 0, s->stream, psi, a); break;
            case 3: hipLaunchKernelGGL((k_diag_full<T, true, 3>), grid, dim3(kBlock), 0, s->stream, psi, a); break;
            default: hipLaunchKernelGGL((k_diag_full<T, true, 0>), grid, dim3(kBlock), 0, s->stream, psi, a);
        }
        prof_end(s);
        QH_LAUNCH_CHECK();
        return QVMHIP_OK;
    }
    SparseArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    make_group_spec(a.g, gs, s->n);
    a.nent = (uint32_t)ent.size();
    for (size_t e = 0; e < ent.size(); ++e) {
        a.off[e] = deposit(ent[e], gs.tq, gs.kt);
        a.dr[e] = (T)gs.V[(size_t)ent[e] * Dr + ent[e]].real();
        a.di[e] = (T)gs.V[(size_t)ent[e] * Dr + ent[e]].imag();
    }
    constexpr int GPT = 1;
    const dim3 grid(grid_for(a.g.ngroups, (uint64_t)kBlock * GPT));
    const bool lr = long_runs(a.g, s->amp_bytes);
    V2Ptr<T> psi = (vec2_t<T> *)s->amps;
    // FP64 only: one 8-byte FP32 amplitude per lane makes 512 B wavefront
    // accesses, and the line form measured slower there (CPHASE(2, 9) at
    // n = 30: 1.64 vs 1.50 ms); FP64 CPHASE(1, 8) 3.14 -> 2.65 ms
    if (sizeof(T) == 8 && a.nent == 1 && !lr) {
        const LineSplit ls = line_split(a.g, s->n, s->amp_bytes);
        if (ls.hg.ngroups % kBlock == 0) {
            const uint64_t e0 = a.off[0] | a.g.ctrl_mask;
            prof_begin(s, QVMHIP_K_DIAG_SPARSE);
            hipLaunchKernelGGL((k_diag_line<T>), dim3(grid_for(ls.hg.ngroups, kBlock)), dim3(kBlock), 0, s->stream,
                               psi, ls.hg, ls.hg.ngroups, e0 & ls.hmask, ls.lmask, (uint32_t)e0 & ls.lmask, a.dr[0],
                               a.di[0], (ls.lmask & 3u) == 0 ? 1u : 0u);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    // FP32, one entry, runs of >= 128 B: 16-byte pairs per thread (n = 30:
    // CPHASE(11, 4) 0.80 -> 0.68 ms, (29, 22) 0.80 -> 0.64, (15, 8) 0.97 ->
    // 0.85; two or four groups per thread gained less, and with shorter runs
    // every variant lost)
    if (sizeof(T) == 4 && a.nent == 1 && lr) {
        GroupSpec g2 = a.g;
        for (uint32_t j = 0; j < g2.nsorted; ++j) g2.sorted[j] -= 1;
        g2.ctrl_mask >>= 1;
        g2.ngroups >>= 1;
        prof_begin(s, QVMHIP_K_DIAG_SPARSE);
        hipLaunchKernelGGL(k_diag_sparse2_f32, dim3(grid_for(g2.ngroups, (uint64_t)kBlock)), dim3(kBlock), 0,
                           s->stream, (float2 *)s->amps, g2, a.off[0] >> 1, (float)a.dr[0], (float)a.di[0]);
        prof_end(s);
        QH_LAUNCH_CHECK();
        return QVMHIP_OK;
    }
    prof_begin(s, QVMHIP_K_DIAG_SPARSE);
    if (a.nent == 1) {
        if (lr) hipLaunchKernelGGL((k_diag_sparse<T, GPT, true, true>), grid, dim3(kBlock), 0, s->stream, psi, a);
        else hipLaunchKernelGGL((k_diag_sparse<T, GPT, false, true>), grid, dim3(kBlock), 0, s->stream, psi, a);
    } else {
        if (lr) hipLaunchKernelGGL((k_diag_sparse<T, GPT, true, false>), grid, dim3(kBlock), 0, s->stream, psi, a);
        else hipLaunchKernelGGL((k_diag_sparse<T, GPT, false, false>), grid, dim3(kBlock), 0, s->stream, psi, a);
    }
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T>
int run_perm(qvmhip_state *s, const GateSplit &gs) {
    const uint32_t Dr = 1u << gs.kt;
    SparseArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    make_group_spec(a.g, gs, s->n);
    // moved entries in cycle order: d0 <- perm[d0] = d1 <- perm[d1] = d2 ...
    // <- d0 (psi'[j] = psi[perm[j]])
    std::vector<bool> seen(Dr, false);
    for (uint32_t j0 = 0; j0 < Dr; ++j0) {
        if (gs.perm[j0] == j0 || seen[j0]) continue;
        uint32_t d = j0;
        const uint32_t first = a.nent;
        do {
            seen[d] = true;
            a.off[a.nent] = deposit(d, gs.tq, gs.kt);
            a.src[a.nent] = deposit(gs.perm[d], gs.tq, gs.kt);
            a.flag[a.nent] = 0;
            ++a.nent;
            d = gs.perm[d];
        } while (d != j0);
        a.flag[first] |= 1u;
        a.flag[a.nent - 1] |= 2u;
    }
    if (a.nent == 0) return QVMHIP_OK;
    if (sizeof(T) == 4 && a.nent == 2 && s->n >= 9 && a.g.nsorted && a.g.sorted[a.g.nsorted - 1] < 7 &&
        !long_runs(a.g, s->amp_bytes)) {
        uint32_t tmask = 0;
        for (uint32_t t = 0; t < gs.kt; ++t) tmask |= 1u << gs.tq[t];
        prof_begin(s, QVMHIP_K_PERM);
        hipLaunchKernelGGL(k_perm_lane2_f32, dim3(grid_for(s->len / 2, kBlock)), dim3(kBlock), 0, s->stream,
                           (float2 *)s->amps, s->len / 2, tmask, (uint32_t)a.off[0], (uint32_t)a.off[1],
                           (uint32_t)a.g.ctrl_mask);
        prof_end(s);
        QH_LAUNCH_CHECK();
        return QVMHIP_OK;
    }
    if (a.nent == 2 && s->n >= 8 && a.g.nsorted && a.g.sorted[a.g.nsorted - 1] < 6 &&
        !long_runs(a.g, s->amp_bytes)) {
        // every bit involved is a lane bit and the moved amplitudes come in
        // runs shorter than a 128 B line: a whole-state sweep with shuffles
        // beats the half-state pair form (n = 30: 5.3 vs 6.4 ms FP64, 2.8 vs
        // 3.2 ms FP32); from 128 B runs on the pair form moves half the
        // bytes (3.0 ms FP64)
        uint32_t tmask = 0;
        for (uint32_t t = 0; t < gs.kt; ++t) tmask |= 1u << gs.tq[t];
        prof_begin(s, QVMHIP_K_PERM);
        // FP64 with every bit >= 2: the moved amplitudes are 64-byte halves
        // of lines, and only they are stored (as in k_diag_line)
        const uint32_t masked = (sizeof(T) == 8 && ((tmask | (uint32_t)a.g.ctrl_mask) & 3u) == 0) ? 1u : 0u;
        hipLaunchKernelGGL((k_perm_lane<T>), dim3(grid_for(s->len, kBlock)), dim3(kBlock), 0, s->stream,
                           (vec2_t<T> *)s->amps, s->len, tmask, (uint32_t)a.off[0], (uint32_t)a.off[1],
                           (uint32_t)a.g.ctrl_mask, masked);
        prof_end(s);
        QH_LAUNCH_CHECK();
        return QVMHIP_OK;
    }
    constexpr int GPT = 1;
    const dim3 grid(grid_for(a.g.ngroups, (uint64_t)kBlock * GPT));
    const bool lr = long_runs(a.g, s->amp_bytes);
    V2Ptr<T> psi = (vec2_t<T> *)s->amps;
    if (sizeof(T) == 8 && a.nent == 2 && a.g.nsorted && a.g.sorted[0] >= 6 && a.g.ngroups % 128 == 0) {
        prof_begin(s, QVMHIP_K_PERM);
        hipLaunchKernelGGL((k_perm_xchg<T>), dim3(grid_for(a.g.ngroups, 128)), dim3(kBlock), 0, s->stream,
                           (vec2_t<T> *)s->amps, a.g, a.off[0], a.off[1]);
        prof_end(s);
        QH_LAUNCH_CHECK();
        return QVMHIP_OK;
    }
    // FP64 (as for diagonals): CNOT(2, 9) at n = 30 3.14 -> 2.60 ms, SWAP(0,
    // 29) 6.71 -> 5.49 ms
    if (sizeof(T) == 8 && a.nent == 2 && !lr) {
        const LineSplit ls = line_split(a.g, s->n, s->amp_bytes);
        if (ls.hg.ngroups % kBlock == 0) {
            const uint64_t e0 = a.off[0] | a.g.ctrl_mask, e1 = a.off[1] | a.g.ctrl_mask;
            prof_begin(s, QVMHIP_K_PERM);
            hipLaunchKernelGGL((k_perm_line<T>), dim3(grid_for(ls.hg.ngroups, kBlock)), dim3(kBlock), 0, s->stream,
                               psi, ls.hg, ls.hg.ngroups, e0 & ls.hmask, e1 & ls.hmask, ls.lmask,
                               (uint32_t)e0 & ls.lmask, (uint32_t)e1 & ls.lmask);
            prof_end(s);
            QH_LAUNCH_CHECK();
            return QVMHIP_OK;
        }
    }
    prof_begin(s, QVMHIP_K_PERM);
    if (a.nent == 2) {
        if (lr) hipLaunchKernelGGL((k_perm<T, GPT, true, true>), grid, dim3(kBlock), 0, s->stream, psi, a);
        else hipLaunchKernelGGL((k_perm<T, GPT, false, true>), grid, dim3(kBlock), 0, s->stream, psi, a);
    } else {
        if (lr) hipLaunchKernelGGL((k_perm<T, GPT, true, false>), grid, dim3(kBlock), 0, s->stream, psi, a);
        else hipLaunchKernelGGL((k_perm<T, GPT, false, false>), grid, dim3(kBlock), 0, s->stream, psi, a);
    }
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T>
int run_split(qvmhip_state *s, const GateSplit &gs) {
    switch (gs.kind) {
        case 0:
            return QVMHIP_OK;
        case 2:
            return run_diag<T>(s, gs);
        case 3:
            return run_perm<T>(s, gs);
        default:
            break;
    }
    switch (gs.kt) {
        case 0: {  // dense 1x1 is a phase on the control subspace
            GateSplit d = gs;
            d.kind = 2;
            return run_diag<T>(s, d);
        }
        case 1:
            return run_dense_small<T, 1>(s, gs);
        case 2:
            return run_dense_small<T, 2>(s, gs);
        case 3:
            return run_dense_small<T, 3>(s, gs);
        case 4:
            return run_dense_big<T, 4>(s, gs);
        case 5:
            return run_dense_big<T, 5>(s, gs);
        case 6:
            return run_dense_big<T, 6>(s, gs);
    }
    return fail(QVMHIP_EINVAL, "gate wider than 6 qubits");
}

}  // namespace

int launch_gate(qvmhip_state *s, const double *U, uint32_t k, const uint32_t *qubits,
                int force_dense) {
    const GateSplit gs = split_gate(reinterpret_cast<const cd *>(U), k, qubits, !force_dense);
    if (s->dtype == QVMHIP_F64) return run_split<double>(s, gs);
    return run_split<float>(s, gs);
}

int launch_diagonal(qvmhip_state *s, const double *d, uint32_t k, const uint32_t *qubits) {
    const uint32_t D = 1u << k;
    std::vector<cd> U((size_t)D * D, cd(0, 0));
    for (uint32_t j = 0; j < D; ++j) U[(size_t)j * D + j] = cd(d[2 * j], d[2 * j + 1]);
    return launch_gate(s, reinterpret_cast<const double *>(U.data()), k, qubits, 0);
}

int launch_permutation(qvmhip_state *s, const uint32_t *perm, uint32_t k, const uint32_t *qubits) {
    const uint32_t D = 1u << k;
    std::vector<char> seen(D, 0);
    for (uint32_t j = 0; j < D; ++j) {
        if (perm[j] >= D || seen[perm[j]]) return fail(QVMHIP_EINVAL, "not a permutation");
        seen[perm[j]] = 1;
    }
    std::vector<cd> U((size_t)D * D, cd(0, 0));
    for (uint32_t j = 0; j < D; ++j) U[(size_t)j * D + perm[j]] = cd(1, 0);
    return launch_gate(s, reinterpret_cast<const double *>(U.data()), k, qubits, 0);
}

int launch_fill_zero_state(qvmhip_state *s) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((s->len + kBlock - 1) / kBlock, 256 * 64);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_fill_zero_state<double>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (double2 *)s->amps, s->len);
    else
        hipLaunchKernelGGL(k_fill_zero_state<float>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (float2 *)s->amps, s->len);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int launch_scale(qvmhip_state *s, double factor) {
    const uint32_t grid = (uint32_t)std::min<uint64_t>((s->len + kBlock - 1) / kBlock, 256 * 64);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_scale<double>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (double2 *)s->amps, s->len, factor);
    else
        hipLaunchKernelGGL(k_scale<float>, dim3(grid), dim3(kBlock), 0, s->stream, (float2 *)s->amps,
                           s->len, (float)factor);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

}  // namespace qvmhip
