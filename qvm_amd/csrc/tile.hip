// tile.hip -- gate sequences applied to LDS-resident blocks of the state.
//
// A tile pass names m "tile qubits".  For every assignment of the other n-m
// qubits, one workgroup loads the 2^m amplitudes spanned by the tile qubits
// into LDS, applies the pass's gates one after another (each a 1- or 2-qubit
// op on local positions, with __syncthreads between ops) and writes the block
// back: the whole gate sequence costs one HBM read and one write of the state.
// Tile qubit sets that include qubits 0..L-1 give contiguous runs of 2^L
// amplitudes per row, so loads and stores stay coalesced; high-order gate
// targets are brought in simply by naming them as tile qubits (the load is a
// strided gather of those runs), which is the qubit permutation into an
// LDS-resident block.
//
// Arithmetic per gate is the same as the plain kernels (gates.hip): separate
// complex products summed left to right, built with -ffp-contract=off.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

struct TileArgs {
    GroupSpec g;        // tile base: inject zeros at the m tile positions
    uint64_t toff_bits; // unused padding for alignment
    uint32_t tq[16];    // global positions of the tile qubits (ascending)
    uint64_t roff[32];  // offsets of the per-thread rounds r (bits >= log2(NT))
};

template <int NT>
__device__ inline uint64_t thread_offset(const TileArgs &a, uint32_t t) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; (1 << b) < NT; ++b)
        if ((t >> b) & 1u) o |= uint64_t(1) << a.tq[b];
    return o;
}

__device__ inline uint32_t ins0(uint32_t x, uint32_t q) {
    const uint32_t lo = x & ((1u << q) - 1u);
    return ((x >> q) << (q + 1)) | lo;
}

template <typename T>
__device__ inline void cmul(T ar, T ai, T br, T bi, T &r, T &i) {
    r = ar * br - ai * bi;
    i = ar * bi + ai * br;
}

// One tile: load the 2^M amplitudes of tile `tile` into LDS, apply the op
// list, store.  Shared by the single-state and the batched kernels.
template <typename T, int M, int NT>
__device__ inline void tile_body(vec2_t<T> *__restrict__ psi, const TileArgs &a,
                                 const qvmhip_tile_op *__restrict__ ops, uint32_t nops, uint64_t tile,
                                 vec2_t<T> *sa, uint64_t toff) {
    using V = vec2_t<T>;
    constexpr int S = 1 << M;
    constexpr int R = S / NT;  // amplitudes per thread
    const uint32_t t = threadIdx.x;
    {
        const uint64_t base = group_base(a.g, tile) + toff;

        // load: element l = t + r*NT (only this thread reads these slots back in
        // the previous tile's store phase, so no barrier is needed before it)
        {
            V x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = psi[base + a.roff[r]];
#pragma unroll
            for (int r = 0; r < R; ++r) sa[t + r * NT] = x[r];
        }
        __syncthreads();

        for (uint32_t oi = 0; oi < nops; ++oi) {
            const qvmhip_tile_op &op = ops[oi];
            const uint32_t kind = op.kind;
            const uint32_t q0 = op.q0, q1 = op.q1, cm = op.ctrl_mask;
            if (kind == QVMHIP_TILE_DENSE1) {
                const T u00r = (T)op.m[0], u00i = (T)op.m[1], u01r = (T)op.m[2], u01i = (T)op.m[3];
                const T u10r = (T)op.m[4], u10i = (T)op.m[5], u11r = (T)op.m[6], u11i = (T)op.m[7];
#pragma unroll
                for (uint32_t p = t; p < S / 2; p += NT) {
                    const uint32_t ia = ins0(p, q0), ib = ia | (1u << q0);
                    if ((ia & cm) != cm) continue;
                    const V xa = sa[ia], xb = sa[ib];
                    T p0r, p0i, p1r, p1i;
                    V ya, yb;
                    cmul(xa.x, xa.y, u00r, u00i, p0r, p0i);
                    cmul(xb.x, xb.y, u01r, u01i, p1r, p1i);
                    ya.x = p0r + p1r;
                    ya.y = p0i + p1i;
                    cmul(xa.x, xa.y, u10r, u10i, p0r, p0i);
                    cmul(xb.x, xb.y, u11r, u11i, p1r, p1i);
                    yb.x = p0r + p1r;
                    yb.y = p0i + p1i;
                    sa[ia] = ya;
                    sa[ib] = yb;
                }
            } else if (kind == QVMHIP_TILE_DENSE2) {
                const uint32_t lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
#pragma unroll
                for (uint32_t p = t; p < S / 4; p += NT) {
                    const uint32_t i0 = ins0(ins0(p, lo), hi);
                    if ((i0 & cm) != cm) continue;
                    const uint32_t idx[4] = {i0, i0 | (1u << q0), i0 | (1u << q1), i0 | (1u << q0) | (1u << q1)};
                    V x[4];
#pragma unroll
                    for (int j = 0; j < 4; ++j) x[j] = sa[idx[j]];
#pragma unroll
                    for (int row = 0; row < 4; ++row) {
                        T yr, yi;
                        cmul(x[0].x, x[0].y, (T)op.m[8 * row], (T)op.m[8 * row + 1], yr, yi);
#pragma unroll
                        for (int c = 1; c < 4; ++c) {
                            T pr, pi;
                            cmul(x[c].x, x[c].y, (T)op.m[8 * row + 2 * c], (T)op.m[8 * row + 2 * c + 1], pr, pi);
                            yr = yr + pr;
                            yi = yi + pi;
                        }
                        V y;
                        y.x = yr;
                        y.y = yi;
                        sa[idx[row]] = y;
                    }
                }
            } else if (kind == QVMHIP_TILE_DIAG1 || kind == QVMHIP_TILE_DIAG2) {
                const bool two = kind == QVMHIP_TILE_DIAG2;
#pragma unroll
                for (uint32_t l = t; l < S; l += NT) {
                    if ((l & cm) != cm) continue;
                    uint32_t j = (l >> q0) & 1u;
                    if (two) j |= ((l >> q1) & 1u) << 1;
                    const V x = sa[l];
                    V y;
                    cmul(x.x, x.y, (T)op.m[2 * j], (T)op.m[2 * j + 1], y.x, y.y);
                    sa[l] = y;
                }
            } else if (kind == QVMHIP_TILE_CPERM) {
                // target q0, control q1: swap (c=1,t=0) <-> (c=1,t=1)
                const uint32_t lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
#pragma unroll
                for (uint32_t p = t; p < S / 4; p += NT) {
                    const uint32_t i0 = ins0(ins0(p, lo), hi) | (1u << q1);
                    if ((i0 & cm) != cm) continue;
                    const uint32_t i1 = i0 | (1u << q0);
                    const V x0 = sa[i0], x1 = sa[i1];
                    sa[i0] = x1;
                    sa[i1] = x0;
                }
            } else if (kind == QVMHIP_TILE_SWAP) {
                const uint32_t lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
#pragma unroll
                for (uint32_t p = t; p < S / 4; p += NT) {
                    const uint32_t i0 = ins0(ins0(p, lo), hi);
                    if ((i0 & cm) != cm) continue;
                    const uint32_t ia = i0 | (1u << q0), ib = i0 | (1u << q1);
                    const V xa = sa[ia], xb = sa[ib];
                    sa[ia] = xb;
                    sa[ib] = xa;
                }
            } else if (kind == QVMHIP_TILE_CPHASE) {
                const uint32_t lo = q0 < q1 ? q0 : q1, hi = q0 < q1 ? q1 : q0;
                const T mr = (T)op.m[0], mi = (T)op.m[1];
#pragma unroll
                for (uint32_t p = t; p < S / 4; p += NT) {
                    const uint32_t i0 = ins0(ins0(p, lo), hi) | (1u << q0) | (1u << q1);
                    if ((i0 & cm) != cm) continue;
                    const V x = sa[i0];
                    V y;
                    cmul(x.x, x.y, mr, mi, y.x, y.y);
                    sa[i0] = y;
                }
            }
            __syncthreads();
        }

        // store
        {
            V x[R];
#pragma unroll
            for (int r = 0; r < R; ++r) x[r] = sa[t + r * NT];
#pragma unroll
            for (int r = 0; r < R; ++r) psi[base + a.roff[r]] = x[r];
        }
    }
}

template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_tile(vec2_t<T> *__restrict__ psi, const TileArgs a,
                                             const qvmhip_tile_op *__restrict__ ops, uint32_t nops) {
    __shared__ vec2_t<T> sa[1 << M];
    const uint64_t toff = thread_offset<NT>(a, threadIdx.x);
    for (uint64_t tile = blockIdx.x; tile < a.g.ngroups; tile += gridDim.x) {
        tile_body<T, M, NT>(psi, a, ops, nops, tile, sa, toff);
    }
}

// Batched form: many states of one buffer, one entry per state.  Entry e
// owns blocks [block_start[e], block_start[e] + ntiles[e]) of the launch.
struct TileEntry {
    TileArgs args;
    uint64_t amp_offset;   // first amplitude of the state in the batch buffer
    uint64_t block_start;  // first block of the launch that works on this entry
    const qvmhip_tile_op *ops;
    uint32_t nops;
    uint32_t pad;
};

template <typename T, int M, int NT>
__global__ __launch_bounds__(NT) void k_tile_batch(vec2_t<T> *__restrict__ base, const TileEntry *__restrict__ ent,
                                                   uint32_t nent, uint64_t total_blocks) {
    __shared__ vec2_t<T> sa[1 << M];
    for (uint64_t b = blockIdx.x; b < total_blocks; b += gridDim.x) {
        uint32_t lo = 0, hi = nent;  // last entry with block_start <= b
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (ent[mid].block_start <= b)
                lo = mid;
            else
                hi = mid;
        }
        const TileEntry &e = ent[lo];
        const uint64_t toff = thread_offset<NT>(e.args, threadIdx.x);
        tile_body<T, M, NT>(base + e.amp_offset, e.args, e.ops, e.nops, b - e.block_start, sa, toff);
        __syncthreads();  // the next block id may reuse sa with other threads' slots
    }
}

namespace {

constexpr int threads_for(int m) { return m >= 12 ? 512 : 256; }

// Host-side tile arguments of an n-qubit state and tile qubits tq[0..m).
void make_tile_args(uint32_t n, uint32_t m, const uint32_t *tq, TileArgs &a) {
    const int NT = threads_for((int)m);
    const int LOGNT = NT == 512 ? 9 : 8;
    std::memset(&a, 0, sizeof(a));
    a.g.nsorted = m;
    for (uint32_t i = 0; i < m; ++i) {
        a.g.sorted[i] = tq[i];
        a.tq[i] = tq[i];
    }
    a.g.ctrl_mask = 0;
    a.g.ngroups = uint64_t(1) << (n - m);
    const int R = (1 << m) / NT;
    for (int r = 0; r < R; ++r) {
        uint64_t o = 0;
        const uint32_t l = (uint32_t)r << LOGNT;
        for (uint32_t b = 0; b < m; ++b)
            if ((l >> b) & 1u) o |= uint64_t(1) << tq[b];
        a.roff[r] = o;
    }
}

template <typename T, int M>
int run_tile(qvmhip_state *s, const uint32_t *tq, uint32_t m, const qvmhip_tile_op *ops_dev,
             uint32_t nops) {
    constexpr int NT = threads_for(M);
    TileArgs a;
    make_tile_args(s->n, m, tq, a);
    prof_begin(s, QVMHIP_K_TILE);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(a.g.ngroups, kMaxGrid);
    hipLaunchKernelGGL((k_tile<T, M, NT>), dim3(grid), dim3(NT), 0, s->stream,
                       (vec2_t<T> *)s->amps, a, ops_dev, nops);
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T>
int run_tile_m(qvmhip_state *s, uint32_t m, const uint32_t *tq, const qvmhip_tile_op *ops_dev,
               uint32_t nops) {
    switch (m) {
        case 8: return run_tile<T, 8>(s, tq, m, ops_dev, nops);
        case 9: return run_tile<T, 9>(s, tq, m, ops_dev, nops);
        case 10: return run_tile<T, 10>(s, tq, m, ops_dev, nops);
        case 11: return run_tile<T, 11>(s, tq, m, ops_dev, nops);
        case 12: return run_tile<T, 12>(s, tq, m, ops_dev, nops);
        case 13:
            if (sizeof(T) == 4) return run_tile<T, 13>(s, tq, m, ops_dev, nops);
            break;
    }
    return fail(QVMHIP_EINVAL, "tile size m must be 8..12 (F64) or 8..13 (F32)");
}

}  // namespace

int launch_tile_pass(qvmhip_state *s, uint32_t m, const uint32_t *tile_qubits, uint32_t nops,
                     const qvmhip_tile_op *ops, int ops_on_device) {
    const qvmhip_tile_op *dev = ops;
    qvmhip_tile_op *tmp = nullptr;
    if (!ops_on_device) {
        // Operands go through the matrix slot (64 KiB = 240 ops); longer lists
        // get a temporary buffer released after the stream drains.
        const size_t bytes = sizeof(qvmhip_tile_op) * (size_t)nops;
        if (bytes <= 64 * 1024) {
            QH_CHECK(hipMemcpyAsync(s->mat, ops, bytes, hipMemcpyHostToDevice, s->stream));
            dev = (const qvmhip_tile_op *)s->mat;
        } else {
            QH_CHECK(hipMalloc((void **)&tmp, bytes));
            QH_CHECK(hipMemcpyAsync(tmp, ops, bytes, hipMemcpyHostToDevice, s->stream));
            dev = tmp;
        }
    }
    int rc = s->dtype == QVMHIP_F64 ? run_tile_m<double>(s, m, tile_qubits, dev, nops)
                                    : run_tile_m<float>(s, m, tile_qubits, dev, nops);
    if (tmp) {
        hipStreamSynchronize(s->stream);
        hipFree(tmp);
    }
    return rc;
}

}  // namespace qvmhip

namespace qvmhip {

size_t tile_entry_bytes() { return sizeof(TileEntry); }

void build_tile_entry(void *entry, uint32_t n, uint32_t m, const uint32_t *tq, uint64_t amp_offset,
                      uint64_t block_start, const qvmhip_tile_op *ops_dev, uint32_t nops, uint64_t *ntiles) {
    TileEntry e;
    std::memset(&e, 0, sizeof(e));
    make_tile_args(n, m, tq, e.args);
    e.amp_offset = amp_offset;
    e.block_start = block_start;
    e.ops = ops_dev;
    e.nops = nops;
    *ntiles = e.args.g.ngroups;
    std::memcpy(entry, &e, sizeof(e));
}

namespace {
template <typename T, int M>
void tile_batch_go(hipStream_t st, void *base, const void *ent, uint32_t nent, uint64_t total) {
    constexpr int NT = threads_for(M);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(total, kMaxGrid);
    hipLaunchKernelGGL((k_tile_batch<T, M, NT>), dim3(grid), dim3(NT), 0, st, (vec2_t<T> *)base,
                       (const TileEntry *)ent, nent, total);
}
template <typename T>
int tile_batch_m(hipStream_t st, uint32_t m, void *base, const void *ent, uint32_t nent, uint64_t total) {
    switch (m) {
        case 8: tile_batch_go<T, 8>(st, base, ent, nent, total); break;
        case 9: tile_batch_go<T, 9>(st, base, ent, nent, total); break;
        case 10: tile_batch_go<T, 10>(st, base, ent, nent, total); break;
        case 11: tile_batch_go<T, 11>(st, base, ent, nent, total); break;
        case 12: tile_batch_go<T, 12>(st, base, ent, nent, total); break;
        case 13:
            if (sizeof(T) == 4) {
                tile_batch_go<T, 13>(st, base, ent, nent, total);
                break;
            }
            return fail(QVMHIP_EINVAL, "tile size 13 is FP32 only");
        default:
            return fail(QVMHIP_EINVAL, "tile size m must be 8..13");
    }
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}
}  // namespace

int launch_tile_batch(hipStream_t st, int dtype, uint32_t m, void *base, const void *entries_dev, uint32_t nent,
                      uint64_t total_blocks) {
    return dtype == QVMHIP_F64 ? tile_batch_m<double>(st, m, base, entries_dev, nent, total_blocks)
                               : tile_batch_m<float>(st, m, base, entries_dev, nent, total_blocks);
}

}  // namespace qvmhip
