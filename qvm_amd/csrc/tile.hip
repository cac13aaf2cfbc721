// tile.hip -- gate sequences applied to register-resident blocks of the state.
//
// A tile pass names m "tile qubits".  For every assignment of the other n-m
// qubits one workgroup loads the 2^m amplitudes spanned by the tile qubits,
// applies the pass's gates one after another and writes the block back: the
// whole gate sequence costs one HBM read and one write of the state.  Tile
// qubit sets that include qubits 0..L-1 give contiguous runs of 2^L
// amplitudes, so loads and stores stay coalesced; high-order gate targets are
// brought in by naming them as tile qubits (the load is a strided gather of
// those runs), which is the qubit permutation into an on-chip block.
//
// Inside the workgroup the block lives in registers.  Slot s = t + NT*r of
// the block is register r of thread t, so the low TB = m - RB slot bits are
// thread bits and the high RB bits are register bits (R = 2^RB amplitudes per
// thread).  A gate whose non-diagonal targets sit on register bits runs in
// registers with no data exchange; diagonal gates and control tests run on any
// bit.  When a gate needs a target that sits on a thread bit, the block is
// re-laid out through LDS (REMAP: every thread stores its registers, and loads
// back the slots of a new bit permutation).  Uncontrolled SWAPs are pure
// relabelings of that permutation and move no data.  The host scheduler
// (schedule_tile_ops) turns a pass's op list, written against tile-local bit
// positions, into this slot-space program: it picks which local bits become
// register bits at each REMAP by next use (farthest next use is evicted), and
// ends the pass with the identity layout so the store is coalesced.
//
// Arithmetic per gate is the same as the plain kernels (gates.hip): separate
// complex products summed left to right, built with -ffp-contract=off, so an
// in-order pass is bit-identical to gate-by-gate application.
#include <algorithm>
#include <array>
#include <complex>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

// Slot-space op kinds that only the scheduler emits (after the public ones).
constexpr uint32_t kTileXReg = 16;   // X on register bit (q0 - TB) where ctrl_mask holds
constexpr uint32_t kTileRSwap = 17;  // exchange register bits q0 and q1 (both < RB)
constexpr uint32_t kTileRemap = 18;  // relayout through LDS; new slot bit j <- old bit src[j]
constexpr uint32_t kTileRDiag = 19;  // multiply register r by entry r (m holds R complex entries)
// Uncontrolled 1-qubit ops on register bit (q0 - TB) whose entries have zero
// parts, with those parts left out: every entry real (m[0..3] = u00, u01,
// u10, u11: H, RY), or real diagonal and imaginary off-diagonal (m[0..3] =
// Re u00, Im u01, Im u10, Re u11: RX).
constexpr uint32_t kTileD1Real = 20;
constexpr uint32_t kTileD1RX = 21;
// FMA passes: a run of ZZ phases with one angle (diag(a, b, b, a) on qubit
// pairs, the MaxCut cost layer of QAOA) multiplies amplitude x by
// T[cut(x)], cut(x) = the number of the run's pairs whose bits differ in x,
// T[c] = a^(E - c) b^c.  CUT_PHASE holds the pairs by where their ends sit
// (byte layout in cut_phase_op); the ceil((E + 1) / 16) ops after it are
// DATA ops holding T, which the kernel reads and skips.
constexpr uint32_t kTileCutPhase = 22;
constexpr uint32_t kTileData = 23;
constexpr int kCutMaxEdges = 32;

// Register bits per thread (measured on MI355X): 3 (8 amplitudes) for FP64
// and for exact FP32, 4 (16) for FP32 with FMA contraction; fewer for small
// tiles so a workgroup keeps at least 64 threads.
__host__ __device__ constexpr int tile_reg_bits(int m, bool f64, bool fma) {
    return (m - 6) < ((!f64 && fma) ? 4 : 3) ? (m - 6) : ((!f64 && fma) ? 4 : 3);
}

struct TileArgs {
    GroupSpec g;        // tile base: inject zeros at the m tile positions
    uint32_t tq[16];    // global positions of the tile qubits (ascending)
    uint64_t roff[32];  // global offset of register r (slot bits >= TB)
};

template <int TB>
__device__ inline uint64_t thread_offset(const TileArgs &a, uint32_t t) {
    uint64_t o = 0;
#pragma unroll
    for (int b = 0; b < TB; ++b)
        if ((t >> b) & 1u) o |= uint64_t(1) << a.tq[b];
    return o;
}

__device__ __forceinline__ double qfma(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float qfma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Complex product.  FMA = false: four separately rounded products, as the
// plain kernels; FMA = true: one product and one fused multiply-add per part.
template <typename T, bool FMA = false>
__device__ __forceinline__ void cmul(T ar, T ai, T br, T bi, T &r, T &i) {
    if constexpr (FMA) {
        r = qfma(ar, br, -(ai * bi));
        i = qfma(ar, bi, ai * br);
    } else {
        r = ar * br - ai * bi;
        i = ar * bi + ai * br;
    }
}

// y = a*u + b*v (complex): the sum of two rounded products, or (FMA) one
// product and three fused multiply-adds per part.
template <typename T, bool FMA>
__device__ __forceinline__ void cmac2(T ar, T ai, T ur, T ui, T br, T bi, T vr, T vi, T &yr, T &yi) {
    if constexpr (FMA) {
        yr = qfma(ar, ur, qfma(-ai, ui, qfma(br, vr, -(bi * vi))));
        yi = qfma(ar, ui, qfma(ai, ur, qfma(br, vi, bi * vr)));
    } else {
        T p0r, p0i, p1r, p1i;
        cmul(ar, ai, ur, ui, p0r, p0i);
        cmul(br, bi, vr, vi, p1r, p1i);
        yr = p0r + p1r;
        yi = p0i + p1i;
    }
}

// Operand i of a scheduled op: FP64 passes keep the doubles of the public op;
// for FP32 passes the scheduler stores the values as floats in the same
// bytes, so the kernel reads them without a conversion.
template <typename T>
__device__ __forceinline__ T mval(const qvmhip_tile_op &op, int i) {
    if constexpr (sizeof(T) == 8)
        return op.m[i];
    else
        return ((const float *)op.m)[i];
}

// register index with a zero inserted at bit I
template <int I>
__device__ constexpr int rins0(int p) {
    return ((p >> I) << (I + 1)) | (p & ((1 << I) - 1));
}

template <typename T>
__device__ __forceinline__ vec2_t<T> vsel(bool c, vec2_t<T> a, vec2_t<T> b) {
    vec2_t<T> r;
    r.x = c ? a.x : b.x;
    r.y = c ? a.y : b.y;
    return r;
}

// The register ops compute every pair unconditionally and, for controlled
// ops (CTRL), keep the old values by selection where the control bits are not
// all 1 (lane part lane_ok, register part rcm): straight-line code, so the
// pairs of all registers interleave.

template <typename T, int R, int I, bool CTRL, bool FMA>
__device__ __forceinline__ void reg_dense1(vec2_t<T> (&x)[R], const qvmhip_tile_op &op, bool lane_ok,
                                           uint32_t rcm) {
    using V = vec2_t<T>;
    const T u00r = mval<T>(op, 0), u00i = mval<T>(op, 1), u01r = mval<T>(op, 2), u01i = mval<T>(op, 3);
    const T u10r = mval<T>(op, 4), u10i = mval<T>(op, 5), u11r = mval<T>(op, 6), u11i = mval<T>(op, 7);
#pragma unroll
    for (int p = 0; p < R / 2; ++p) {
        const int ra = rins0<I>(p), rb = ra | (1 << I);
        const V xa = x[ra], xb = x[rb];
        V ya, yb;
        cmac2<T, FMA>(xa.x, xa.y, u00r, u00i, xb.x, xb.y, u01r, u01i, ya.x, ya.y);
        cmac2<T, FMA>(xa.x, xa.y, u10r, u10i, xb.x, xb.y, u11r, u11i, yb.x, yb.y);
        if constexpr (CTRL) {
            const bool ok = lane_ok && ((uint32_t)ra & rcm) == rcm;
            x[ra] = vsel<T>(ok, ya, xa);
            x[rb] = vsel<T>(ok, yb, xb);
        } else {
            x[ra] = ya;
            x[rb] = yb;
        }
    }
}

// 4x4 on register bits 0 (matrix bit 0) and 1 (matrix bit 1)
template <typename T, int R, bool CTRL, bool FMA>
__device__ __forceinline__ void reg_dense2(vec2_t<T> (&x)[R], const qvmhip_tile_op &op, bool lane_ok,
                                           uint32_t rcm) {
    using V = vec2_t<T>;
#pragma unroll
    for (int p = 0; p < R / 4; ++p) {
        const int r0 = p << 2;
        V in[4], out[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) in[j] = x[r0 | j];
#pragma unroll
        for (int row = 0; row < 4; ++row) {
            T yr, yi;
            cmul<T, FMA>(in[0].x, in[0].y, mval<T>(op, 8 * row), mval<T>(op, 8 * row + 1), yr, yi);
#pragma unroll
            for (int c = 1; c < 4; ++c) {
                const T ur = mval<T>(op, 8 * row + 2 * c), ui = mval<T>(op, 8 * row + 2 * c + 1);
                if constexpr (FMA) {
                    yr = qfma(in[c].x, ur, qfma(-in[c].y, ui, yr));
                    yi = qfma(in[c].x, ui, qfma(in[c].y, ur, yi));
                } else {
                    T pr, pi;
                    cmul(in[c].x, in[c].y, ur, ui, pr, pi);
                    yr = yr + pr;
                    yi = yi + pi;
                }
            }
            out[row].x = yr;
            out[row].y = yi;
        }
        const bool ok = !CTRL || (lane_ok && ((uint32_t)r0 & rcm) == rcm);
#pragma unroll
        for (int j = 0; j < 4; ++j) x[r0 | j] = CTRL ? vsel<T>(ok, out[j], in[j]) : out[j];
    }
}

// The 1-qubit product of reg_dense1 with the terms whose matrix factor is 0
// left out.  x*0 is an exact zero and adding an exact zero leaves a nonzero
// value as it is, so every output is the same rounded value as the general
// form's, in both the separately rounded and the FMA form (an exact-zero
// result may differ in its sign only).  Half the multiplies of reg_dense1.
template <typename T, int R, int I, bool FMA>
__device__ __forceinline__ void reg_dense1_real(vec2_t<T> (&x)[R], T u00, T u01, T u10, T u11) {
    using V = vec2_t<T>;
#pragma unroll
    for (int p = 0; p < R / 2; ++p) {
        const int ra = rins0<I>(p), rb = ra | (1 << I);
        const V xa = x[ra], xb = x[rb];
        V ya, yb;
        if constexpr (FMA) {
            ya.x = qfma(xa.x, u00, xb.x * u01);
            ya.y = qfma(xa.y, u00, xb.y * u01);
            yb.x = qfma(xa.x, u10, xb.x * u11);
            yb.y = qfma(xa.y, u10, xb.y * u11);
        } else {
            ya.x = xa.x * u00 + xb.x * u01;
            ya.y = xa.y * u00 + xb.y * u01;
            yb.x = xa.x * u10 + xb.x * u11;
            yb.y = xa.y * u10 + xb.y * u11;
        }
        x[ra] = ya;
        x[rb] = yb;
    }
}

// u00 = a, u01 = i b, u10 = i c, u11 = d (a, b, c, d real):
// ya = (a xa.re - b xb.im, a xa.im + b xb.re), yb = (d xb.re - c xa.im,
// d xb.im + c xa.re), each sum in the general form's order.
template <typename T, int R, int I, bool FMA>
__device__ __forceinline__ void reg_dense1_rx(vec2_t<T> (&x)[R], T a, T b, T c, T d) {
    using V = vec2_t<T>;
#pragma unroll
    for (int p = 0; p < R / 2; ++p) {
        const int ra = rins0<I>(p), rb = ra | (1 << I);
        const V xa = x[ra], xb = x[rb];
        V ya, yb;
        if constexpr (FMA) {
            ya.x = qfma(xa.x, a, -(xb.y * b));
            ya.y = qfma(xa.y, a, xb.x * b);
            yb.x = qfma(-xa.y, c, xb.x * d);
            yb.y = qfma(xa.x, c, xb.y * d);
        } else {
            ya.x = xa.x * a - xb.y * b;
            ya.y = xa.y * a + xb.x * b;
            yb.x = xb.x * d - xa.y * c;
            yb.y = xa.x * c + xb.y * d;
        }
        x[ra] = ya;
        x[rb] = yb;
    }
}

template <typename T, int R, int I>
__device__ __forceinline__ void reg_x(vec2_t<T> (&x)[R], bool lane_ok, uint32_t rcm) {
    using V = vec2_t<T>;
#pragma unroll
    for (int p = 0; p < R / 2; ++p) {
        const int ra = rins0<I>(p), rb = ra | (1 << I);
        const bool ok = lane_ok && ((uint32_t)ra & rcm) == rcm;
        const V xa = x[ra], xb = x[rb];
        x[ra] = vsel<T>(ok, xb, xa);
        x[rb] = vsel<T>(ok, xa, xb);
    }
}

// controlled SWAP of register bits 0 and 1
template <typename T, int R>
__device__ __forceinline__ void reg_swap01(vec2_t<T> (&x)[R], bool lane_ok, uint32_t rcm) {
    using V = vec2_t<T>;
#pragma unroll
    for (int p = 0; p < R / 4; ++p) {
        const int r0 = p << 2;
        const bool ok = lane_ok && ((uint32_t)r0 & rcm) == rcm;
        const V a = x[r0 | 1], b = x[r0 | 2];
        x[r0 | 1] = vsel<T>(ok, b, a);
        x[r0 | 2] = vsel<T>(ok, a, b);
    }
}

// relabel register bits A < B (no arithmetic)
template <int A, int B>
__device__ constexpr int bit_swapped(int r) {
    return (((r >> A) & 1) == ((r >> B) & 1)) ? r : (r ^ (1 << A) ^ (1 << B));
}

template <typename T, int R, int A, int B>
__device__ __forceinline__ void reg_rswap(vec2_t<T> (&x)[R]) {
    using V = vec2_t<T>;
    V y[R];
#pragma unroll
    for (int r = 0; r < R; ++r) y[r] = x[bit_swapped<A, B>(r)];
#pragma unroll
    for (int r = 0; r < R; ++r) x[r] = y[r];
}

template <typename T, int R>
__device__ __forceinline__ void reg_rswap_any(vec2_t<T> (&x)[R], uint32_t a, uint32_t b) {
    const uint32_t code = a * 8 + b;
    switch (code) {
#define QH_RS(A, B)                                                 \
    case A * 8 + B:                                                 \
        if constexpr ((1 << B) < R) reg_rswap<T, R, A, B>(x);       \
        break;
        QH_RS(0, 1) QH_RS(0, 2) QH_RS(0, 3) QH_RS(0, 4) QH_RS(1, 2) QH_RS(1, 3) QH_RS(1, 4)
        QH_RS(2, 3) QH_RS(2, 4) QH_RS(3, 4)
#undef QH_RS
        default: break;
    }
}

// 4-way select of a complex entry of m by a (per-lane) index j
template <typename T>
__device__ __forceinline__ void pick4(const qvmhip_tile_op &op, uint32_t j, T &dr, T &di) {
    const T r0 = mval<T>(op, 0), i0 = mval<T>(op, 1), r1 = mval<T>(op, 2), i1 = mval<T>(op, 3);
    const T r2 = mval<T>(op, 4), i2 = mval<T>(op, 5), r3 = mval<T>(op, 6), i3 = mval<T>(op, 7);
    const T ra = (j & 1u) ? r1 : r0, ia = (j & 1u) ? i1 : i0;
    const T rb = (j & 1u) ? r3 : r2, ib = (j & 1u) ? i3 : i2;
    dr = (j & 2u) ? rb : ra;
    di = (j & 2u) ? ib : ia;
}

// Register r takes entry (c1) when its bit RS is set, (c0) otherwise;
// registers whose register-control bits (rcm) are not all set are skipped (a
// uniform test).  Lane controls are folded into the entries by the caller.
template <typename T, int R, int RS, bool FMA>
__device__ __forceinline__ void diag_rbit(vec2_t<T> (&x)[R], T c0r, T c0i, T c1r, T c1i, uint32_t rcm) {
    if constexpr ((1 << RS) < R) {
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (((uint32_t)r & rcm) == rcm) {
                const bool one = ((r >> RS) & 1) != 0;
                vec2_t<T> y;
                cmul<T, FMA>(x[r].x, x[r].y, one ? c1r : c0r, one ? c1i : c0i, y.x, y.y);
                x[r] = y;
            }
        }
    }
}

// Register r times entry r: a diagonal whose index bits (and controls) are
// all register bits, with its per-register entries worked out by the
// scheduler.
template <typename T, int R, bool FMA>
__device__ __forceinline__ void reg_rdiag(vec2_t<T> (&x)[R], const qvmhip_tile_op &op) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        vec2_t<T> y;
        cmul<T, FMA>(x[r].x, x[r].y, mval<T>(op, 2 * r), mval<T>(op, 2 * r + 1), y.x, y.y);
        x[r] = y;
    }
}

// diagonal 1q / 2q and CPHASE on any slot bits.  The diagonal index has a
// lane part (target bits that are thread bits, the same for all registers)
// and a register part (uniform per register).  No case needs a per-register
// vector select: the lane part picks one entry (or two) per lane once per op,
// a lane control (or CPHASE condition) turns a lane's entry into 1, register
// controls skip whole registers, and the register part is a scalar choice or,
// for one lane and one register bit, a compile-time one (switch on the
// register bit).  A lane whose condition fails multiplies by exactly 1, which
// leaves its amplitude's value unchanged.
template <typename T, int R, int TB, bool CTRL, bool FMA>
__device__ __forceinline__ void reg_diag(vec2_t<T> (&x)[R], const qvmhip_tile_op &op, uint32_t t, bool lane_ok,
                                         uint32_t rcm_in) {
    using V = vec2_t<T>;
    const uint32_t kind = op.kind, q0 = op.q0, q1 = op.q1;
    constexpr uint32_t lanes = (1u << TB) - 1u;
    const uint32_t rcm = CTRL ? rcm_in : 0u;
    if (kind == QVMHIP_TILE_CPHASE) {
        const uint32_t need = op.ctrl_mask | (1u << q0) | (1u << q1);
        const uint32_t rn = need >> TB, ln = need & lanes;
        T fr = mval<T>(op, 0), fi = mval<T>(op, 1);
        if (ln) {
            const bool lok = (t & ln) == ln;
            fr = lok ? fr : T(1);
            fi = lok ? fi : T(0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (((uint32_t)r & rn) == rn) {
                V y;
                cmul<T, FMA>(x[r].x, x[r].y, fr, fi, y.x, y.y);
                x[r] = y;
            }
        }
        return;
    }
    const bool two = kind == QVMHIP_TILE_DIAG2;
    const bool reg0 = q0 >= TB, reg1 = two && q1 >= TB;
    if (reg0 && (!two || reg1)) {
        // index bits all on registers (the scheduler turns such ops into
        // register diagonals unless a control bit is a lane bit)
        const uint32_t rs0 = q0 - TB, rs1 = two ? q1 - TB : 0u;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (((uint32_t)r & rcm) == rcm) {
                const uint32_t jr = (((uint32_t)r >> rs0) & 1u) | (two ? ((((uint32_t)r >> rs1) & 1u) << 1) : 0u);
                T dr, di;
                pick4<T>(op, jr, dr, di);
                if constexpr (CTRL) {
                    dr = lane_ok ? dr : T(1);
                    di = lane_ok ? di : T(0);
                }
                V y;
                cmul<T, FMA>(x[r].x, x[r].y, dr, di, y.x, y.y);
                x[r] = y;
            }
        }
        return;
    }
    // lane part of the index
    const uint32_t jl = (reg0 ? 0u : ((t >> q0) & 1u)) | ((two && !reg1) ? (((t >> q1) & 1u) << 1) : 0u);
    if (!reg0 && !reg1) {
        // index bits all on lanes: one entry per lane for every register
        T dr, di;
        if (two) {
            pick4<T>(op, jl, dr, di);
        } else {
            dr = jl ? mval<T>(op, 2) : mval<T>(op, 0);
            di = jl ? mval<T>(op, 3) : mval<T>(op, 1);
        }
        if constexpr (CTRL) {
            dr = lane_ok ? dr : T(1);
            di = lane_ok ? di : T(0);
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            if (((uint32_t)r & rcm) == rcm) {
                V y;
                cmul<T, FMA>(x[r].x, x[r].y, dr, di, y.x, y.y);
                x[r] = y;
            }
        }
        return;
    }
    // one lane bit and one register bit: two entries per lane, and the
    // register bit (a compile-time position below) chooses between them
    const uint32_t rs = reg0 ? q0 - TB : q1 - TB;
    const uint32_t w = reg0 ? 1u : 2u;
    T c0r, c0i, c1r, c1i;
    pick4<T>(op, jl, c0r, c0i);
    pick4<T>(op, jl | w, c1r, c1i);
    if constexpr (CTRL) {
        c0r = lane_ok ? c0r : T(1);
        c0i = lane_ok ? c0i : T(0);
        c1r = lane_ok ? c1r : T(1);
        c1i = lane_ok ? c1i : T(0);
    }
    switch (rs) {
        case 0: diag_rbit<T, R, 0, FMA>(x, c0r, c0i, c1r, c1i, rcm); break;
        case 1: diag_rbit<T, R, 1, FMA>(x, c0r, c0i, c1r, c1i, rcm); break;
        case 2: diag_rbit<T, R, 2, FMA>(x, c0r, c0i, c1r, c1i, rcm); break;
        case 3: diag_rbit<T, R, 3, FMA>(x, c0r, c0i, c1r, c1i, rcm); break;
        case 4: diag_rbit<T, R, 4, FMA>(x, c0r, c0i, c1r, c1i, rcm); break;
        default: break;
    }
}

// Diagonal on local slot bit q0 whose entries also depend on state qubits
// outside the tile: `sel` (their bits in the block's base index, uniform)
// picks the entry pair (2 sel, 2 sel + 1), and q0 picks within it.
template <typename T, int R, int TB, bool FMA>
__device__ __forceinline__ void reg_gdiag(vec2_t<T> (&x)[R], const qvmhip_tile_op &op, uint32_t sel, uint32_t t) {
    using V = vec2_t<T>;
    const T r0 = mval<T>(op, 4 * sel), i0 = mval<T>(op, 4 * sel + 1);
    const T r1 = mval<T>(op, 4 * sel + 2), i1 = mval<T>(op, 4 * sel + 3);
    const uint32_t q0 = op.q0;
    if (q0 >= TB) {
        const uint32_t rs = q0 - TB;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool one = (((uint32_t)r >> rs) & 1u) != 0;  // uniform per register
            V y;
            cmul<T, FMA>(x[r].x, x[r].y, one ? r1 : r0, one ? i1 : i0, y.x, y.y);
            x[r] = y;
        }
        return;
    }
    const bool one = ((t >> q0) & 1u) != 0;
    const T dr = one ? r1 : r0, di = one ? i1 : i0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
        V y;
        cmul<T, FMA>(x[r].x, x[r].y, dr, di, y.x, y.y);
        x[r] = y;
    }
}

// CUT_PHASE (the layout of its bytes is described at cut_phase_op): the
// phase table T of the DATA ops after it goes to LDS, each thread counts the
// cut pairs of its amplitudes and multiplies register r by T[cut].
template <typename T, int R, int TB, bool FMA>
__device__ __forceinline__ void reg_cut_phase(vec2_t<T> (&x)[R], const qvmhip_tile_op *__restrict__ op,
                                              vec2_t<T> *sa, uint32_t t, uint64_t gb) {
    const uint8_t *b = (const uint8_t *)op->m;
    const uint32_t E = b[0];
    // values derived from t stay inside this op (hoisted out of the op loop
    // they would hold registers through every other op)
    asm volatile("" : "+v"(t));
    __syncthreads();  // earlier readers of sa are done
    if (t <= E) {
        const qvmhip_tile_op &d = op[1 + t / 16];
        vec2_t<T> v;
        v.x = mval<T>(d, 2 * (t & 15u));
        v.y = mval<T>(d, 2 * (t & 15u) + 1);
        sa[t] = v;
    }
    uint32_t c = 0, A = 0;
#pragma unroll 4  // by four: one scalar load per four pairs (a full unroll computed the pair terms ahead and spilled)
    for (uint32_t i = 0; i < E; ++i) {
        {
            const uint32_t cls = b[72 + i], e0 = b[8 + 2 * i], e1 = b[9 + 2 * i];
            if (cls == 0u) c += ((t >> e0) ^ (t >> e1)) & 1u;                          // LL
            else if (cls == 1u) c += ((t >> e0) ^ (uint32_t)(gb >> e1)) & 1u;          // LO
            else if (cls == 2u) A |= ((t >> e0) & 1u) << i;                            // LR
            else if (cls == 3u) A |= ((uint32_t)(gb >> e1) & 1u) << i;                 // RO
            else if (cls == 5u) c += (uint32_t)((gb >> e0) ^ (gb >> e1)) & 1u;         // OO
        }
    }
    const uint32_t *B = (const uint32_t *)(b + 128);
    __syncthreads();  // the table is in LDS
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const vec2_t<T> f = sa[c + (uint32_t)__popc(A ^ B[r])];
        vec2_t<T> y;
        cmul<T, FMA>(x[r].x, x[r].y, f.x, f.y, y.x, y.y);
        x[r] = y;
        __builtin_amdgcn_sched_barrier(0);  // one table read at a time: R reads ahead hold 2R registers
    }
}

// LDS address of slot s: register row r = s >> TB of NT slots is padded by
// one element, so a thread's store of register r is t + r*(NT+1) (a fixed
// offset from t), and a lane bit that a REMAP takes from register bit b of the
// old layout steps by (NT+1)*2^b, i.e. 2^b modulo the bank count: with the
// scheduler pairing vacated lane bit p with evicted register bit p, the lanes
// of one LDS access hit distinct banks.
template <int TB>
__device__ __forceinline__ uint32_t lds_addr(uint32_t s) {
    return s + (s >> TB);
}

// src4: the op's first 16 operand bytes as four words; byte j is the old
// slot bit of new slot bit j.  FP64 blocks go through LDS one component at a
// time (real parts, then imaginary parts): the exchange buffer is then 2^M
// doubles, half a block, and five workgroups fit on a CU instead of four
// (measured at n = 30: 8.23 -> 7.84 ms per exact pass, 7.50 -> 7.13 FMA).
// FP32 blocks move whole amplitudes (split, they lost 2.5-6 %).
template <typename T, int R, int TB>
__device__ __forceinline__ void reg_remap(vec2_t<T> (&x)[R], T *sa, const uint32_t (&src4)[4], uint32_t t) {
    constexpr uint32_t NT = 1u << TB;
    auto src = [&](int j) -> uint32_t { return (src4[j >> 2] >> (8 * (j & 3))) & 0xFFu; };
    uint32_t base = 0;
#pragma unroll
    for (int j = 0; j < TB; ++j)
        if ((t >> j) & 1u) base |= 1u << src(j);
    uint32_t pw[8];
#pragma unroll
    for (int b = 0; b < 8; ++b) pw[b] = (1 << b) < R ? (1u << src(TB + b)) : 0u;
    auto addr = [&](int r) -> uint32_t {
        uint32_t off = 0;
#pragma unroll
        for (int b = 0; (1 << b) < R; ++b)
            if ((r >> b) & 1) off |= pw[b];
        return lds_addr<TB>(base | off);
    };
    if constexpr (sizeof(T) != 8) {
        vec2_t<T> *sv = reinterpret_cast<vec2_t<T> *>(sa);
        __syncthreads();  // readers of the previous layout are done with sa
#pragma unroll
        for (int r = 0; r < R; ++r) sv[t + (uint32_t)r * (NT + 1)] = x[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) x[r] = sv[addr(r)];
        return;
    }
    __syncthreads();  // readers of the previous layout are done with sa
#pragma unroll
    for (int r = 0; r < R; ++r) sa[t + (uint32_t)r * (NT + 1)] = x[r].x;
    __syncthreads();
    T re[R];
#pragma unroll
    for (int r = 0; r < R; ++r) re[r] = sa[addr(r)];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) sa[t + (uint32_t)r * (NT + 1)] = x[r].y;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) {
        x[r].x = re[r];
        x[r].y = sa[addr(r)];
    }
}

// One tile: load the 2^M amplitudes of tile `tile` into registers, run the
// scheduled op list, store.  Shared by the single-state and batched kernels.
template <typename T, int M, int RB>
__device__ __forceinline__ void tile_load(const vec2_t<T> *__restrict__ psi, const TileArgs &a, uint64_t gb,
                                          uint64_t toff, vec2_t<T> (&x)[1 << RB]) {
    const vec2_t<T> *p = psi + gb + toff;
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) x[r] = ldg<true>(&p[a.roff[r]]);
}

template <typename T, int M, int RB>
__device__ __forceinline__ void tile_store(vec2_t<T> *__restrict__ psi, const TileArgs &a, uint64_t gb,
                                           uint64_t toff, const vec2_t<T> (&x)[1 << RB]) {
    // recompute the store addresses instead of keeping them live across
    // the op loop (the barrier is on the offset: on the pointer itself it
    // would hide the global address space, and flat stores count against
    // the LDS and scalar-load waits of the next tile)
    uint64_t o = gb + toff;
    asm volatile("" : "+v"(o));
    vec2_t<T> *p = psi + o;
#pragma unroll
    for (int r = 0; r < (1 << RB); ++r) stg<true>(&p[a.roff[r]], x[r]);
}

// Run a scheduled op list on the block held in x.
// CUT: the op list holds CUT_PHASE ops (FMA passes only).  Kernels without
// them are built without that code: its registers and code size slow every
// other op down.
template <typename T, int M, int RB, bool FMA, bool CUT>
__device__ inline void tile_ops(vec2_t<T> (&x)[1 << RB], const qvmhip_tile_op *__restrict__ ops, uint32_t nops,
                                T *sa, uint64_t gb) {
    constexpr int R = 1 << RB, TB = M - RB;
    const uint32_t t = threadIdx.x;
    for (uint32_t oi = 0; oi < nops; ++oi) {
        const qvmhip_tile_op &op = ops[oi];
        // the header and the first 32 operand bytes in one scalar load: the
        // reduced 1q ops (most ops of a pass) then start after a single wait
        struct Head48 {
            uint32_t w[12];
        };
        const Head48 h48 = *reinterpret_cast<const Head48 *>(&op);
        asm volatile("" ::"s"(h48.w[1]), "s"(h48.w[2]), "s"(h48.w[4]), "s"(h48.w[5]), "s"(h48.w[6]),
                     "s"(h48.w[7]), "s"(h48.w[8]), "s"(h48.w[9]), "s"(h48.w[10]), "s"(h48.w[11]));
        const uint4 hdr = make_uint4(h48.w[0], h48.w[1], h48.w[2], h48.w[3]);
        auto coef = [&](int k) -> T {  // operand k of a reduced 1q op
            if constexpr (sizeof(T) == 8)
                return __hiloint2double((int)h48.w[5 + 2 * k], (int)h48.w[4 + 2 * k]);
            else
                return __uint_as_float(h48.w[4 + k]);
        };
        const uint32_t kind = hdr.x, cm = hdr.w;
        const bool lane_ok = (t & cm) == (cm & ((1u << TB) - 1u));
        const uint32_t rcm = cm >> TB;
        const bool ctrl = cm != 0;
        switch (kind) {
            case QVMHIP_TILE_DENSE1:
#define QH_D1(I)                                                             \
    case I:                                                                  \
        if constexpr (I < RB) {                                              \
            if (ctrl)                                                        \
                reg_dense1<T, R, I, true, FMA>(x, op, lane_ok, rcm);            \
            else                                                             \
                reg_dense1<T, R, I, false, FMA>(x, op, lane_ok, rcm);            \
        }                                                                    \
        break;
                switch (hdr.y - TB) {
                    QH_D1(0) QH_D1(1) QH_D1(2) QH_D1(3) QH_D1(4)
                    default: break;
                }
#undef QH_D1
                break;
            case kTileD1Real:
#define QH_D1R(I)                                                        \
    case I:                                                              \
        if constexpr (I < RB) reg_dense1_real<T, R, I, FMA>(x, coef(0), coef(1), coef(2), coef(3)); \
        break;
                switch (hdr.y - TB) {
                    QH_D1R(0) QH_D1R(1) QH_D1R(2) QH_D1R(3) QH_D1R(4)
                    default: break;
                }
#undef QH_D1R
                break;
            case kTileD1RX:
#define QH_D1X(I)                                                        \
    case I:                                                              \
        if constexpr (I < RB) reg_dense1_rx<T, R, I, FMA>(x, coef(0), coef(1), coef(2), coef(3)); \
        break;
                switch (hdr.y - TB) {
                    QH_D1X(0) QH_D1X(1) QH_D1X(2) QH_D1X(3) QH_D1X(4)
                    default: break;
                }
#undef QH_D1X
                break;
            case QVMHIP_TILE_DENSE2:
                if (ctrl)
                    reg_dense2<T, R, true, FMA>(x, op, lane_ok, rcm);
                else
                    reg_dense2<T, R, false, FMA>(x, op, lane_ok, rcm);
                break;
            case kTileXReg:
                switch (hdr.y - TB) {
                    case 0: reg_x<T, R, 0>(x, lane_ok, rcm); break;
                    case 1: reg_x<T, R, 1>(x, lane_ok, rcm); break;
                    case 2: if constexpr (RB > 2) reg_x<T, R, 2>(x, lane_ok, rcm); break;
                    case 3: if constexpr (RB > 3) reg_x<T, R, 3>(x, lane_ok, rcm); break;
                    case 4: if constexpr (RB > 4) reg_x<T, R, 4>(x, lane_ok, rcm); break;
                    default: break;
                }
                break;
            case QVMHIP_TILE_SWAP:
                reg_swap01<T, R>(x, lane_ok, rcm);
                break;
            case kTileRSwap:
                reg_rswap_any<T, R>(x, hdr.y, hdr.z);
                break;
            case kTileRemap:
                {
                    const uint32_t src4[4] = {h48.w[4], h48.w[5], h48.w[6], h48.w[7]};
                    reg_remap<T, R, TB>(x, sa, src4, t);
                }
                break;
            case kTileRDiag:
                reg_rdiag<T, R, FMA>(x, op);
                break;
            case QVMHIP_TILE_GDIAG: {
                const uint32_t d = hdr.z;
                uint32_t sel = (uint32_t)(gb >> (d & 63u)) & 1u;
                if ((d >> 16) == 2u) sel |= ((uint32_t)(gb >> ((d >> 8) & 63u)) & 1u) << 1;
                reg_gdiag<T, R, TB, FMA>(x, op, sel, t);
                break;
            }
            default:  // DIAG1, DIAG2, CPHASE; CUT_PHASE (no case label of its own: each label
                      // deepens the dispatch of every op)
                if constexpr (FMA && CUT) {
                    if (kind == kTileCutPhase) {
                        reg_cut_phase<T, R, TB, FMA>(x, ops + oi, reinterpret_cast<vec2_t<T> *>(sa), t, gb);
                        oi += ((const uint8_t *)op.m)[1];  // its DATA ops
                        break;
                    }
                }
                if (ctrl)
                    reg_diag<T, R, TB, true, FMA>(x, op, t, lane_ok, rcm);
                else
                    reg_diag<T, R, TB, false, FMA>(x, op, t, lane_ok, rcm);
                break;
        }
    }

}

template <typename T, int M, int RB, bool FMA, bool CUT>
__global__ __launch_bounds__(1 << (M - RB)) void k_tile(vec2_t<T> *__restrict__ psi, const TileArgs a,
                                                       const qvmhip_tile_op *__restrict__ ops, uint32_t nops) {
    using V = vec2_t<T>;
    // the remap buffer: 2^M + 2^RB reals for FP64 (one component at a time),
    // as many amplitudes for FP32
    __shared__ __align__(16) T sa[(sizeof(T) == 8 ? 1 : 2) * ((1 << M) + (1 << RB))];
    const uint64_t toff = thread_offset<M - RB>(a, threadIdx.x);
    // One tile per workgroup (grid-striding only past the dispatch limit):
    // many resident workgroups overlap each other's loads, ops and stores.
    for (uint64_t tile = xcd_block(blockIdx.x, gridDim.x); tile < a.g.ngroups; tile += gridDim.x) {
        V x[1 << RB];
        const uint64_t gb = group_base(a.g, tile);
        tile_load<T, M, RB>(psi, a, gb, toff, x);
        tile_ops<T, M, RB, FMA, CUT>(x, ops, nops, sa, gb);
        tile_store<T, M, RB>(psi, a, gb, toff, x);
    }
}

// Batched form: many states of one buffer, one entry per state.  Work block
// b of a launch belongs to entry blk_ent[b] (a host-built map, one load
// instead of a search) and covers tiles [(b - block_start) * CH, ... + CH) of
// that entry's state.
// The op lists of all entries live in one device buffer passed to the kernel
// as a restrict argument; an entry holds its byte offset there.  (A pointer
// read from memory would not be known to be unaliased with the state, and
// the op loads would become per-lane vector loads with every op field in
// vector registers instead of scalar loads into SGPRs.)
struct TileEntry {
    TileArgs args;
    uint64_t amp_offset;   // first amplitude of the state in the batch buffer
    uint64_t block_start;  // first work block of the launch that works on this entry
    uint64_t op_off;       // byte offset of the entry's scheduled ops in the op buffer
    // state qubits that no earlier pass or gate of the run has touched: the
    // run starts from |0...0>, so a block whose base index has one of them set
    // is all zeros before this pass, and every tile op keeps it zero
    uint64_t zero_mask;
    uint32_t nops;
    uint32_t from_zero;    // 1: the state is |0...0>, built in registers instead of loaded
};

// ZERO: some entry of the launch starts from |0...0> or has all-zero
// blocks; launches without (every level after the first few) are built
// without that code, which costs the exact FP64 forms 60 VGPRs.
// FP64 FMA forms held to five waves per SIMD (96 VGPRs, 11 spilled; left
// alone they take 108 and four waves, 5 % slower on the 512-instance QAOA
// batch); the exact forms would spill ~100 values at five.
template <typename T, int M, int RB, bool FMA, bool CUT, bool ZERO>
__global__ __launch_bounds__(1 << (M - RB))
__attribute__((amdgpu_waves_per_eu(FMA && sizeof(T) == 8 ? 5 : 1))) void k_tile_batch(
    vec2_t<T> *__restrict__ base, const TileEntry *__restrict__ ent, const uint32_t *__restrict__ blk_ent,
    uint64_t total_blocks, uint32_t tiles_per_block, const char *__restrict__ op_buf) {
    using V = vec2_t<T>;
    // the remap buffer: 2^M + 2^RB reals for FP64 (one component at a time),
    // as many amplitudes for FP32
    __shared__ __align__(16) T sa[(sizeof(T) == 8 ? 1 : 2) * ((1 << M) + (1 << RB))];
    // round robin over the XCDs (no xcd_block): work blocks of different
    // entries differ in cost, and contiguous shares per XCD measured 5-6 %
    // slower on the 512-instance QAOA batch
    for (uint64_t b = blockIdx.x; b < total_blocks; b += gridDim.x) {
        const TileEntry &e = ent[blk_ent[b]];
        const uint64_t toff = thread_offset<M - RB>(e.args, threadIdx.x);
        const uint64_t first = (b - e.block_start) * tiles_per_block;
        const uint64_t last = first + tiles_per_block < e.args.g.ngroups ? first + tiles_per_block : e.args.g.ngroups;
        V *psi = base + e.amp_offset;
        for (uint64_t tile = first; tile < last; ++tile) {
            V x[1 << RB];
            const uint64_t gb = group_base(e.args.g, tile);
            if (!ZERO) {
                tile_load<T, M, RB>(psi, e.args, gb, toff, x);
                tile_ops<T, M, RB, FMA, CUT>(x, (const qvmhip_tile_op *)(op_buf + e.op_off), e.nops, sa, gb);
                tile_store<T, M, RB>(psi, e.args, gb, toff, x);
                continue;
            }
            if (gb & e.zero_mask) {
                // an all-zero block stays zero: the first pass of a run stores
                // it (the buffer holds the previous run), later passes leave
                // the zeros in place
                if (e.from_zero) {
#pragma unroll
                    for (int r = 0; r < (1 << RB); ++r) x[r].x = x[r].y = T(0);
                    tile_store<T, M, RB>(psi, e.args, gb, toff, x);
                }
                continue;
            }
            if (e.from_zero) {
                // the first pass of a batch run: |0...0> needs no read
#pragma unroll
                for (int r = 0; r < (1 << RB); ++r) {
                    x[r].x = (gb + toff + e.args.roff[r] == 0) ? T(1) : T(0);
                    x[r].y = T(0);
                }
            } else {
                tile_load<T, M, RB>(psi, e.args, gb, toff, x);
            }
            tile_ops<T, M, RB, FMA, CUT>(x, (const qvmhip_tile_op *)(op_buf + e.op_off), e.nops, sa, gb);
            tile_store<T, M, RB>(psi, e.args, gb, toff, x);
        }
    }
}

// ---------------------------------------------------------------------------
// Host side.

int check_tile_ops(int dtype, uint32_t m, uint32_t nops, const qvmhip_tile_op *ops) {
    if (m < 8 || m > 13 || (m == 13 && dtype == QVMHIP_F64))
        return fail(QVMHIP_EINVAL, "tile size m must be 8..12 (F64) or 8..13 (F32)");
    for (uint32_t o = 0; o < nops; ++o) {
        const qvmhip_tile_op &op = ops[o];
        if (op.kind < QVMHIP_TILE_DENSE1 || op.kind > QVMHIP_TILE_GDIAG)
            return fail(QVMHIP_EINVAL, "unknown tile op kind");
        if (op.kind == QVMHIP_TILE_GDIAG) {
            const uint32_t ng = op.q1 >> 16;
            if (op.q0 >= m) return fail(QVMHIP_EINVAL, "tile op qubit out of range");
            if (ng < 1 || ng > 2) return fail(QVMHIP_EINVAL, "GDIAG reads one or two outside qubits");
            if (op.ctrl_mask) return fail(QVMHIP_EINVAL, "GDIAG takes no controls");
            continue;
        }
        const bool two = op.kind != QVMHIP_TILE_DENSE1 && op.kind != QVMHIP_TILE_DIAG1;
        if (op.q0 >= m || (two && (op.q1 >= m || op.q1 == op.q0)))
            return fail(QVMHIP_EINVAL, "tile op qubit out of range");
        if (m < 32 && (op.ctrl_mask >> m)) return fail(QVMHIP_EINVAL, "tile op control outside the tile");
        if ((op.ctrl_mask >> op.q0) & 1u || (two && ((op.ctrl_mask >> op.q1) & 1u)))
            return fail(QVMHIP_EINVAL, "tile op control overlaps its targets");
    }
    return QVMHIP_OK;
}

int check_tile_globals(uint32_t n, uint32_t m, const uint32_t *tq, uint32_t nops, const qvmhip_tile_op *ops) {
    for (uint32_t o = 0; o < nops; ++o) {
        const qvmhip_tile_op &op = ops[o];
        if (op.kind != QVMHIP_TILE_GDIAG) continue;
        const uint32_t ng = op.q1 >> 16;
        const uint32_t g[2] = {op.q1 & 0xffu, (op.q1 >> 8) & 0xffu};
        for (uint32_t k = 0; k < ng; ++k) {
            if (g[k] >= n) return fail(QVMHIP_EINVAL, "GDIAG qubit out of range");
            for (uint32_t i = 0; i < m; ++i)
                if (tq[i] == g[k]) return fail(QVMHIP_EINVAL, "GDIAG outside qubit is a tile qubit");
        }
        if (ng == 2 && g[0] == g[1]) return fail(QVMHIP_EINVAL, "GDIAG outside qubits repeat");
    }
    return QVMHIP_OK;
}

namespace {

// A scheduled DIAG1 / DIAG2 / CPHASE whose index bits and controls are all
// register bits becomes a register diagonal: entry r is the factor register
// r takes (1 where the controls are not met).
bool to_register_diagonal(qvmhip_tile_op &o, int tb, int rb) {
    if (o.kind != QVMHIP_TILE_DIAG1 && o.kind != QVMHIP_TILE_DIAG2 && o.kind != QVMHIP_TILE_CPHASE) return false;
    const bool two = o.kind != QVMHIP_TILE_DIAG1;
    const uint32_t lanes = (1u << tb) - 1u;
    if ((int)o.q0 < tb || (two && (int)o.q1 < tb) || (o.ctrl_mask & lanes)) return false;
    const uint32_t rcm = o.ctrl_mask >> tb, b0 = o.q0 - tb, b1 = two ? o.q1 - tb : 0u;
    double f[32];
    for (int r = 0; r < (1 << rb); ++r) {
        double re = 1.0, im = 0.0;
        const uint32_t ur = (uint32_t)r;
        if ((ur & rcm) == rcm) {
            const uint32_t j0 = (ur >> b0) & 1u, j1 = two ? (ur >> b1) & 1u : 0u;
            if (o.kind == QVMHIP_TILE_CPHASE) {
                if (j0 && j1) {
                    re = o.m[0];
                    im = o.m[1];
                }
            } else {
                const uint32_t j = j0 | (j1 << 1);
                re = o.m[2 * j];
                im = o.m[2 * j + 1];
            }
        }
        f[2 * r] = re;
        f[2 * r + 1] = im;
    }
    o.kind = kTileRDiag;
    o.q0 = o.q1 = o.ctrl_mask = 0;
    std::memset(o.m, 0, sizeof(o.m));
    std::memcpy(o.m, f, sizeof(double) * 2 * (1 << rb));
    return true;
}

// An uncontrolled 1-qubit op whose entries are all real, or whose diagonal is
// real and off-diagonal imaginary, becomes the matching reduced kind.
void to_split_dense1(qvmhip_tile_op &o) {
    const double *u = o.m;  // u00 = (u[0], u[1]), u01 = (u[2], u[3]), u10 = (u[4], u[5]), u11 = (u[6], u[7])
    double v[4];
    if (u[1] == 0.0 && u[3] == 0.0 && u[5] == 0.0 && u[7] == 0.0) {
        o.kind = kTileD1Real;
        v[0] = u[0], v[1] = u[2], v[2] = u[4], v[3] = u[6];
    } else if (u[1] == 0.0 && u[2] == 0.0 && u[4] == 0.0 && u[7] == 0.0) {
        o.kind = kTileD1RX;
        v[0] = u[0], v[1] = u[3], v[2] = u[5], v[3] = u[6];
    } else {
        return;
    }
    std::memset(o.m, 0, sizeof(o.m));
    std::memcpy(o.m, v, sizeof(v));
}

// A public op that is a ZZ phase diag(a, b, b, a) between two qubits, each
// a tile bit (local position) or a state qubit outside the tile.
struct ZZEdge {
    std::complex<double> a, b;
    int end[2];  // local position, or 64 + outside state qubit
};

bool zz_edge(const qvmhip_tile_op &op, ZZEdge &e) {
    auto c = [&](int i) { return std::complex<double>(op.m[2 * i], op.m[2 * i + 1]); };
    if (op.kind == QVMHIP_TILE_DIAG2 && op.ctrl_mask == 0) {
        if (!(c(0) == c(3) && c(1) == c(2))) return false;
        e.a = c(0), e.b = c(1);
        e.end[0] = (int)op.q0, e.end[1] = (int)op.q1;
        return true;
    }
    if (op.kind != QVMHIP_TILE_GDIAG) return false;
    const uint32_t ng = op.q1 >> 16, g0 = op.q1 & 0xFFu, g1 = (op.q1 >> 8) & 0xFFu;
    if (ng == 1) {  // entry index: local bit | outside bit << 1
        if (!(c(0) == c(3) && c(1) == c(2))) return false;
        e.a = c(0), e.b = c(1);
        e.end[0] = (int)op.q0, e.end[1] = 64 + (int)g0;
        return true;
    }
    // both ends outside: entries must not depend on the local bit
    for (int i = 0; i < 8; i += 2)
        if (!(c(i) == c(i + 1))) return false;
    if (!(c(0) == c(6) && c(2) == c(4))) return false;
    e.a = c(0), e.b = c(2);
    e.end[0] = 64 + (int)g0, e.end[1] = 64 + (int)g1;
    return true;
}

// Local positions an op reads or changes (a ZZ phase with both ends outside
// the tile touches none).
uint32_t touched_bits(const qvmhip_tile_op &op) {
    ZZEdge e;
    if (zz_edge(op, e)) {
        uint32_t msk = 0;
        for (int k = 0; k < 2; ++k)
            if (e.end[k] < 64) msk |= 1u << e.end[k];
        return msk;
    }
    uint32_t msk = (1u << op.q0) | op.ctrl_mask;
    if (op.kind == QVMHIP_TILE_DENSE2 || op.kind == QVMHIP_TILE_DIAG2 || op.kind == QVMHIP_TILE_CPERM ||
        op.kind == QVMHIP_TILE_SWAP || op.kind == QVMHIP_TILE_CPHASE)
        msk |= 1u << op.q1;
    return msk;
}

bool diagonal_public(uint32_t kind) {
    return kind == QVMHIP_TILE_DIAG1 || kind == QVMHIP_TILE_DIAG2 || kind == QVMHIP_TILE_CPHASE ||
           kind == QVMHIP_TILE_GDIAG;
}

// Ops in layers: op i goes one layer after the latest earlier op it does
// not commute with (a shared bit, unless both are diagonal); within a layer
// the ZZ phases come first, grouped by angle, in their original order.
std::vector<qvmhip_tile_op> layer_order(const qvmhip_tile_op *in, uint32_t nin) {
    std::vector<uint32_t> level(nin, 0), bits(nin);
    std::vector<bool> diag(nin), zz(nin);
    std::vector<ZZEdge> edge(nin);
    for (uint32_t i = 0; i < nin; ++i) {
        bits[i] = touched_bits(in[i]);
        diag[i] = diagonal_public(in[i].kind);
        zz[i] = zz_edge(in[i], edge[i]);
        for (uint32_t j = 0; j < i; ++j)
            if ((bits[i] & bits[j]) && !(diag[i] && diag[j])) level[i] = std::max(level[i], level[j] + 1);
    }
    // angle group of each ZZ phase: index of the first ZZ phase of its level with the same entries
    std::vector<uint32_t> group(nin, UINT32_MAX);
    for (uint32_t i = 0; i < nin; ++i) {
        if (!zz[i]) continue;
        group[i] = i;
        for (uint32_t j = 0; j < i; ++j) {
            if (zz[j] && level[j] == level[i] && edge[j].a == edge[i].a && edge[j].b == edge[i].b) {
                group[i] = group[j];
                break;
            }
        }
    }
    std::vector<uint32_t> order(nin);
    for (uint32_t i = 0; i < nin; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
        if (level[x] != level[y]) return level[x] < level[y];
        if (zz[x] != zz[y]) return (bool)zz[x];
        if (zz[x]) return group[x] < group[y];
        return false;
    });
    std::vector<qvmhip_tile_op> out(nin);
    for (uint32_t i = 0; i < nin; ++i) out[i] = in[order[i]];
    return out;
}

// Local bits an op needs on register bits (its non-diagonal targets).
uint32_t reg_need(const qvmhip_tile_op &op) {
    switch (op.kind) {
        case QVMHIP_TILE_DENSE1: return 1u << op.q0;
        case QVMHIP_TILE_DENSE2: return (1u << op.q0) | (1u << op.q1);
        case QVMHIP_TILE_CPERM: return 1u << op.q0;
        case QVMHIP_TILE_SWAP: return op.ctrl_mask ? ((1u << op.q0) | (1u << op.q1)) : 0u;
        default: return 0u;
    }
}

struct Layout {
    int m, tb, rb;
    int l2s[16], s2l[16];  // local bit <-> slot bit
    void swap_slots(int a, int b) {
        const int la = s2l[a], lb = s2l[b];
        s2l[a] = lb;
        s2l[b] = la;
        l2s[la] = b;
        l2s[lb] = a;
    }
    bool identity() const {
        for (int i = 0; i < m; ++i)
            if (l2s[i] != i) return false;
        return true;
    }
    uint32_t slot_mask(uint32_t local_mask) const {
        uint32_t o = 0;
        for (int i = 0; i < m; ++i)
            if ((local_mask >> i) & 1u) o |= 1u << l2s[i];
        return o;
    }
};

qvmhip_tile_op remap_op(const Layout &from, const int *new_s2l) {
    qvmhip_tile_op o;
    std::memset(&o, 0, sizeof(o));
    o.kind = kTileRemap;
    uint8_t src[16] = {0};
    for (int j = 0; j < from.m; ++j) src[j] = (uint8_t)from.l2s[new_s2l[j]];
    std::memcpy(o.m, src, sizeof(src));
    return o;
}

// Whether two consecutive ops are ZZ phases of one angle.
bool has_cut_run(const std::vector<qvmhip_tile_op> &ops) {
    ZZEdge prev, e;
    bool have = false;
    for (const qvmhip_tile_op &o : ops) {
        const bool z = zz_edge(o, e);
        if (z && have && e.a == prev.a && e.b == prev.b) return true;
        have = z;
        prev = e;
    }
    return false;
}

// CUT_PHASE for a run of ZZ phases of one angle under layout L, followed by
// its DATA ops.  Bytes of m: [0] E (pairs, <= 32), [1] number of DATA ops;
// pair i: ends at bytes 8 + 2i and 9 + 2i, class at byte 72 + i; bytes
// 128..191: B[r] (uint32, r < 16).  Classes by where the ends sit: LL (two
// thread bits), LO (thread bit, outside qubit), LR (thread bit, register
// bit), RO (register bit, outside qubit), RR (two register bits), OO (two
// outside qubits).  For thread t, register r and block base gb:
//   cut = sum over LL, LO, OO pairs of (bit ^ bit) + popc(A ^ B[r])
// where bit i of A is t's thread bit (LR) or gb's outside bit (RO) of pair
// i, and bit i of B[r] is r's register bit (LR, RO) or the xor of r's two
// register bits (RR).
enum : uint8_t { kCutLL = 0, kCutLO = 1, kCutLR = 2, kCutRO = 3, kCutRR = 4, kCutOO = 5 };

void cut_phase_op(const Layout &L, const std::vector<ZZEdge> &run, std::vector<qvmhip_tile_op> &out) {
    struct End {
        int kind;  // 0 thread bit, 1 register bit, 2 outside qubit
        int pos;
    };
    auto where = [&](int end) -> End {
        if (end >= 64) return {2, end - 64};
        const int sbit = L.l2s[end];
        return sbit < L.tb ? End{0, sbit} : End{1, sbit - L.tb};
    };
    qvmhip_tile_op o;
    std::memset(&o, 0, sizeof(o));
    o.kind = kTileCutPhase;
    uint8_t *b = (uint8_t *)o.m;
    const int E = (int)run.size(), ndata = (E + 1 + 15) / 16;
    b[0] = (uint8_t)E;
    b[1] = (uint8_t)ndata;
    uint32_t B[16] = {0};
    for (int i = 0; i < E; ++i) {
        End x = where(run[i].end[0]), y = where(run[i].end[1]);
        if (x.kind > y.kind) std::swap(x, y);
        uint8_t cls;
        if (x.kind == 0) cls = y.kind == 0 ? kCutLL : (y.kind == 1 ? kCutLR : kCutLO);
        else if (x.kind == 1) cls = y.kind == 1 ? kCutRR : kCutRO;
        else cls = kCutOO;
        b[72 + i] = cls;
        b[8 + 2 * i] = (uint8_t)x.pos;
        b[9 + 2 * i] = (uint8_t)y.pos;
        for (int r = 0; r < (1 << L.rb); ++r) {
            uint32_t v = 0;
            if (cls == kCutLR) v = (uint32_t)(r >> y.pos) & 1u;
            else if (cls == kCutRO) v = (uint32_t)(r >> x.pos) & 1u;
            else if (cls == kCutRR) v = (uint32_t)((r >> x.pos) ^ (r >> y.pos)) & 1u;
            B[r] |= v << i;
        }
    }
    std::memcpy(b + 128, B, sizeof(B));
    out.push_back(o);
    // T[c] = a^(E - c) b^c
    for (int d = 0; d < ndata; ++d) {
        qvmhip_tile_op t;
        std::memset(&t, 0, sizeof(t));
        t.kind = kTileData;
        for (int k = 0; k < 16 && 16 * d + k <= E; ++k) {
            const int c = 16 * d + k;
            std::complex<double> v(1.0, 0.0);
            for (int j = 0; j < E - c; ++j) v *= run[0].a;
            for (int j = 0; j < c; ++j) v *= run[0].b;
            t.m[2 * k] = v.real();
            t.m[2 * k + 1] = v.imag();
        }
        out.push_back(t);
    }
}

}  // namespace

// Turn a pass's op list (tile-local bit positions) into the slot-space
// program the kernel runs.  `pin` places the targets of 2-qubit dense ops and
// controlled SWAPs on register bits 0 and 1.
void schedule_tile_ops(int dtype, bool fma, uint32_t m, const qvmhip_tile_op *in_ops, uint32_t nin,
                       std::vector<qvmhip_tile_op> &out) {
    Layout L;
    L.m = (int)m;
    L.rb = tile_reg_bits((int)m, dtype == QVMHIP_F64, fma);
    L.tb = L.m - L.rb;
    for (int i = 0; i < L.m; ++i) L.l2s[i] = L.s2l[i] = i;
    // FMA passes do not keep gate-by-gate rounding: commuting ops are
    // regrouped into layers so that ZZ phases of one angle become one op
    std::vector<qvmhip_tile_op> layered;
    if (fma) {
        // the layering compares every pair of ops: passes longer than 4096
        // ops keep their order (and still merge the runs they hold)
        if (nin <= 4096) layered = layer_order(in_ops, nin);
        if (nin > 4096 || !has_cut_run(layered)) layered.assign(in_ops, in_ops + nin);  // nothing to merge
    }
    const qvmhip_tile_op *in = fma ? layered.data() : in_ops;
    out.clear();
    out.reserve(nin + nin / 2 + 2);
    std::vector<uint32_t> need(nin);
    for (uint32_t i = 0; i < nin; ++i) need[i] = reg_need(in[i]);
    // next_use[i][l]: first op index >= i that needs local bit l on a register
    std::vector<std::array<uint32_t, 16>> next_use(nin + 1);
    next_use[nin].fill(UINT32_MAX);
    for (int i = (int)nin - 1; i >= 0; --i) {
        next_use[i] = next_use[i + 1];
        for (int l = 0; l < L.m; ++l)
            if ((need[i] >> l) & 1u) next_use[i][l] = (uint32_t)i;
    }
    auto reg_local_mask = [&]() {
        uint32_t msk = 0;
        for (int b = 0; b < L.rb; ++b) msk |= 1u << L.s2l[L.tb + b];
        return msk;
    };
    for (uint32_t i = 0; i < nin; ++i) {
        const qvmhip_tile_op &op = in[i];
        if (fma) {
            ZZEdge e0;
            if (zz_edge(op, e0)) {
                // the run of ZZ phases of this angle that starts here
                std::vector<ZZEdge> run{e0};
                uint32_t k = i + 1;
                ZZEdge e;
                while (k < nin && (int)run.size() < kCutMaxEdges && zz_edge(in[k], e) && e.a == e0.a &&
                       e.b == e0.b) {
                    run.push_back(e);
                    ++k;
                }
                if (run.size() >= 2) {
                    cut_phase_op(L, run, out);
                    i = k - 1;
                    continue;
                }
            }
        }
        if (op.kind == QVMHIP_TILE_SWAP && op.ctrl_mask == 0) {
            L.swap_slots(L.l2s[op.q0], L.l2s[op.q1]);  // relabel only
            continue;
        }
        const bool pin2 = op.kind == QVMHIP_TILE_DENSE2 || op.kind == QVMHIP_TILE_SWAP;
        if (need[i] & ~reg_local_mask()) {
            // New register set: this op's targets, then the local bits with the
            // nearest next use (ties keep bits that are already registers).
            std::vector<int> order(L.m);
            for (int l = 0; l < L.m; ++l) order[l] = l;
            const uint32_t cur = reg_local_mask();
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
                const uint64_t ka = ((uint64_t)next_use[i][a] << 1) | (((cur >> a) & 1u) ? 0u : 1u);
                const uint64_t kb = ((uint64_t)next_use[i][b] << 1) | (((cur >> b) & 1u) ? 0u : 1u);
                return ka < kb;
            });
            uint32_t nset = 0;
            for (int k = 0; k < L.rb; ++k) nset |= 1u << order[k];
            int ns2l[16];
            for (int j = 0; j < L.m; ++j) ns2l[j] = L.s2l[j];
            // register bits: keep residents in place, then fill freed positions
            bool taken[16] = {false};
            int placed_l[16];
            for (int l = 0; l < 16; ++l) placed_l[l] = -1;
            if (pin2) {
                ns2l[L.tb + 0] = (int)op.q0;
                ns2l[L.tb + 1] = (int)op.q1;
                taken[0] = taken[1] = true;
                placed_l[op.q0] = L.tb + 0;
                placed_l[op.q1] = L.tb + 1;
            }
            for (int b = 0; b < L.rb; ++b) {
                const int l = L.s2l[L.tb + b];
                if (((nset >> l) & 1u) && placed_l[l] < 0 && !taken[b]) {
                    ns2l[L.tb + b] = l;
                    taken[b] = true;
                    placed_l[l] = L.tb + b;
                }
            }
            std::vector<int> entering;
            for (int l = 0; l < L.m; ++l)
                if (((nset >> l) & 1u) && placed_l[l] < 0) entering.push_back(l);
            for (int b = 0, e = 0; b < L.rb; ++b)
                if (!taken[b]) {
                    ns2l[L.tb + b] = entering[e];
                    placed_l[entering[e]] = L.tb + b;
                    taken[b] = true;
                    ++e;
                }
            // thread bits: the slots vacated by entering bits take the evicted
            // bits; pair vacated thread bit p with an evicted bit whose old
            // register position is p when there is one (distinct LDS banks)
            std::vector<int> vacated, evicted;
            for (int p = 0; p < L.tb; ++p)
                if ((nset >> L.s2l[p]) & 1u) vacated.push_back(p);
            for (int b = 0; b < L.rb; ++b) {
                const int l = L.s2l[L.tb + b];
                if (!((nset >> l) & 1u)) evicted.push_back(l);
            }
            std::vector<bool> used(evicted.size(), false);
            for (int p : vacated) {
                int pick = -1;
                for (size_t e = 0; e < evicted.size(); ++e)
                    if (!used[e] && L.l2s[evicted[e]] - L.tb == p) pick = (int)e;
                if (pick < 0)
                    for (size_t e = 0; e < evicted.size() && pick < 0; ++e)
                        if (!used[e]) pick = (int)e;
                used[pick] = true;
                ns2l[p] = evicted[pick];
            }
            out.push_back(remap_op(L, ns2l));
            for (int j = 0; j < L.m; ++j) {
                L.s2l[j] = ns2l[j];
                L.l2s[ns2l[j]] = j;
            }
        } else if (pin2) {
            // both targets are registers; move them to register bits 0 and 1
            for (int want = 0; want < 2; ++want) {
                const int l = want == 0 ? (int)op.q0 : (int)op.q1;
                const int b = L.l2s[l] - L.tb;
                if (b != want) {
                    qvmhip_tile_op r;
                    std::memset(&r, 0, sizeof(r));
                    r.kind = kTileRSwap;
                    r.q0 = (uint32_t)std::min(b, want);
                    r.q1 = (uint32_t)std::max(b, want);
                    out.push_back(r);
                    L.swap_slots(L.tb + b, L.tb + want);
                }
            }
        }
        qvmhip_tile_op o = op;
        o.ctrl_mask = L.slot_mask(op.ctrl_mask);
        o.q0 = (uint32_t)L.l2s[op.q0];
        const bool two = op.kind != QVMHIP_TILE_DENSE1 && op.kind != QVMHIP_TILE_DIAG1 &&
                         op.kind != QVMHIP_TILE_GDIAG;
        if (op.kind != QVMHIP_TILE_GDIAG) o.q1 = two ? (uint32_t)L.l2s[op.q1] : 0u;  // GDIAG: outside qubits
        if (op.kind == QVMHIP_TILE_CPERM) {  // X on the target where the control is 1
            o.kind = kTileXReg;
            o.ctrl_mask |= 1u << L.l2s[op.q1];
            o.q1 = 0;
        }
        if (o.kind == QVMHIP_TILE_DENSE1 && o.ctrl_mask == 0) to_split_dense1(o);
        if (to_register_diagonal(o, L.tb, L.rb)) {
            // consecutive register diagonals multiply into one when the
            // pass contracts with FMAs (no promise of gate-by-gate rounding)
            if (fma && !out.empty() && out.back().kind == kTileRDiag) {
                qvmhip_tile_op &prev = out.back();
                for (int r = 0; r < (1 << L.rb); ++r) {
                    const double ar = prev.m[2 * r], ai = prev.m[2 * r + 1];
                    const double br = o.m[2 * r], bi = o.m[2 * r + 1];
                    prev.m[2 * r] = ar * br - ai * bi;
                    prev.m[2 * r + 1] = ar * bi + ai * br;
                }
                continue;
            }
        }
        out.push_back(o);
    }
    if (!L.identity()) {
        int ns2l[16];
        for (int j = 0; j < L.m; ++j) ns2l[j] = j;
        out.push_back(remap_op(L, ns2l));
    }
    if (dtype == QVMHIP_F32) {
        for (qvmhip_tile_op &o : out) {
            if (o.kind == kTileRemap || o.kind == kTileRSwap || o.kind == kTileXReg || o.kind == kTileCutPhase)
                continue;
            float f[32];
            for (int i = 0; i < 32; ++i) f[i] = (float)o.m[i];
            std::memset(o.m, 0, sizeof(o.m));
            std::memcpy(o.m, f, sizeof(f));
        }
    }
}

namespace {

// Host-side tile arguments of an n-qubit state and tile qubits tq[0..m).
void make_tile_args(uint32_t n, uint32_t m, const uint32_t *tq, int tb, TileArgs &a) {
    std::memset(&a, 0, sizeof(a));
    a.g.nsorted = m;
    for (uint32_t i = 0; i < m; ++i) {
        a.g.sorted[i] = tq[i];
        a.tq[i] = tq[i];
    }
    a.g.ctrl_mask = 0;
    a.g.ngroups = uint64_t(1) << (n - m);
    const int R = 1 << (m - tb);
    for (int r = 0; r < R; ++r) {
        uint64_t o = 0;
        const uint32_t l = (uint32_t)r << tb;
        for (uint32_t b = 0; b < m; ++b)
            if ((l >> b) & 1u) o |= uint64_t(1) << tq[b];
        a.roff[r] = o;
    }
}

template <typename T, int M, bool FMA, bool CUT>
int run_tile(qvmhip_state *s, const uint32_t *tq, uint32_t m, const qvmhip_tile_op *ops_dev,
             uint32_t nops) {
    constexpr int RB = tile_reg_bits(M, sizeof(T) == 8, FMA);
    constexpr int NT = 1 << (M - RB);
    TileArgs a;
    make_tile_args(s->n, m, tq, M - RB, a);
    prof_begin(s, QVMHIP_K_TILE);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(a.g.ngroups, kMaxGrid);
    hipLaunchKernelGGL((k_tile<T, M, RB, FMA, CUT>), dim3(grid), dim3(NT), 0, s->stream, (vec2_t<T> *)s->amps, a,
                       ops_dev, nops);
    prof_end(s);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

template <typename T, int M>
int run_tile_f(qvmhip_state *s, const uint32_t *tq, uint32_t m, const qvmhip_tile_op *ops_dev, uint32_t nops,
               bool cut) {
    if (!(s->flags & QVMHIP_FLAG_FMA_TILES)) return run_tile<T, M, false, false>(s, tq, m, ops_dev, nops);
    return cut ? run_tile<T, M, true, true>(s, tq, m, ops_dev, nops)
               : run_tile<T, M, true, false>(s, tq, m, ops_dev, nops);
}

template <typename T>
int run_tile_m(qvmhip_state *s, uint32_t m, const uint32_t *tq, const qvmhip_tile_op *ops_dev,
               uint32_t nops, bool cut) {
    switch (m) {
        case 8: return run_tile_f<T, 8>(s, tq, m, ops_dev, nops, cut);
        case 9: return run_tile_f<T, 9>(s, tq, m, ops_dev, nops, cut);
        case 10: return run_tile_f<T, 10>(s, tq, m, ops_dev, nops, cut);
        case 11: return run_tile_f<T, 11>(s, tq, m, ops_dev, nops, cut);
        case 12: return run_tile_f<T, 12>(s, tq, m, ops_dev, nops, cut);
        case 13:
            if constexpr (sizeof(T) == 4) return run_tile_f<T, 13>(s, tq, m, ops_dev, nops, cut);
            break;
    }
    return fail(QVMHIP_EINVAL, "tile size m must be 8..12 (F64) or 8..13 (F32)");
}

}  // namespace

bool tile_ops_have_cut(const qvmhip_tile_op *ops, size_t n) {
    for (size_t i = 0; i < n; ++i)
        if (ops[i].kind == kTileCutPhase) return true;
    return false;
}

int launch_tile_pass(qvmhip_state *s, uint32_t m, const uint32_t *tile_qubits, uint32_t nops,
                     const qvmhip_tile_op *ops, int ops_on_device, bool cut) {
    const qvmhip_tile_op *dev = ops;
    qvmhip_tile_op *tmp = nullptr;
    if (!ops_on_device) {
        std::vector<qvmhip_tile_op> sched;
        schedule_tile_ops(s->dtype, (s->flags & QVMHIP_FLAG_FMA_TILES) != 0, m, ops, nops, sched);
        nops = (uint32_t)sched.size();
        cut = tile_ops_have_cut(sched.data(), sched.size());
        // Operands go through the matrix slot (64 KiB = 240 ops); longer lists
        // get a temporary buffer released after the stream drains.  The slot
        // is reused by the next call, so wait for any earlier reader first.
        const size_t bytes = sizeof(qvmhip_tile_op) * (size_t)nops;
        if (bytes <= 64 * 1024) {
            QH_CHECK(hipMemcpyAsync(s->mat, sched.data(), bytes, hipMemcpyHostToDevice, s->stream));
            QH_CHECK(hipStreamSynchronize(s->stream));  // sched is a host temporary
            dev = (const qvmhip_tile_op *)s->mat;
        } else {
            QH_CHECK(hipMalloc((void **)&tmp, bytes));
            QH_CHECK(hipMemcpyAsync(tmp, sched.data(), bytes, hipMemcpyHostToDevice, s->stream));
            QH_CHECK(hipStreamSynchronize(s->stream));
            dev = tmp;
        }
    }
    int rc = s->dtype == QVMHIP_F64 ? run_tile_m<double>(s, m, tile_qubits, dev, nops, cut)
                                    : run_tile_m<float>(s, m, tile_qubits, dev, nops, cut);
    if (tmp) {
        (void)hipStreamSynchronize(s->stream);
        (void)hipFree(tmp);
    }
    return rc;
}

size_t tile_entry_bytes() { return sizeof(TileEntry); }

uint64_t build_tile_entry(void *entry, int dtype, bool fma, uint32_t n, uint32_t m, const uint32_t *tq,
                          uint64_t amp_offset, uint64_t block_start, uint64_t op_off, uint32_t nops,
                          uint32_t tiles_per_block, bool from_zero, uint64_t zero_mask) {
    TileEntry e;
    std::memset(&e, 0, sizeof(e));
    make_tile_args(n, m, tq, (int)m - tile_reg_bits((int)m, dtype == QVMHIP_F64, fma), e.args);
    e.amp_offset = amp_offset;
    e.block_start = block_start;
    e.op_off = op_off;
    e.nops = nops;
    e.from_zero = from_zero ? 1u : 0u;
    e.zero_mask = zero_mask;
    std::memcpy(entry, &e, sizeof(e));
    return (e.args.g.ngroups + tiles_per_block - 1) / tiles_per_block;
}

namespace {
template <typename T, int M, bool FMA, bool CUT>
void tile_batch_go(hipStream_t st, void *base, const void *ent, const uint32_t *blk_ent, uint64_t total,
                   uint32_t tpb, const void *op_buf, bool zero) {
    constexpr int RB = tile_reg_bits(M, sizeof(T) == 8, FMA);
    constexpr int NT = 1 << (M - RB);
    const uint32_t grid = (uint32_t)std::min<uint64_t>(total, kMaxGrid);
    if (zero)
        hipLaunchKernelGGL((k_tile_batch<T, M, RB, FMA, CUT, true>), dim3(grid), dim3(NT), 0, st, (vec2_t<T> *)base,
                           (const TileEntry *)ent, blk_ent, total, tpb, (const char *)op_buf);
    else
        hipLaunchKernelGGL((k_tile_batch<T, M, RB, FMA, CUT, false>), dim3(grid), dim3(NT), 0, st,
                           (vec2_t<T> *)base, (const TileEntry *)ent, blk_ent, total, tpb, (const char *)op_buf);
}
template <typename T, bool FMA, bool CUT>
int tile_batch_m(hipStream_t st, uint32_t m, void *base, const void *ent, const uint32_t *blk_ent, uint64_t total,
                 uint32_t tpb, const void *op_buf, bool zero) {
    switch (m) {
        case 8: tile_batch_go<T, 8, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero); break;
        case 9: tile_batch_go<T, 9, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero); break;
        case 10: tile_batch_go<T, 10, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero); break;
        case 11: tile_batch_go<T, 11, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero); break;
        case 12: tile_batch_go<T, 12, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero); break;
        case 13:
            if constexpr (sizeof(T) == 4) {
                tile_batch_go<T, 13, FMA, CUT>(st, base, ent, blk_ent, total, tpb, op_buf, zero);
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

int launch_tile_batch(hipStream_t st, int dtype, bool fma, bool cut, bool zero, uint32_t m, void *base,
                      const void *entries_dev, const uint32_t *blk_ent_dev, uint64_t total_blocks,
                      uint32_t tiles_per_block, const void *op_buf) {
    const bool f64 = dtype == QVMHIP_F64;
#define QH_TB(TT, F, C) \
    tile_batch_m<TT, F, C>(st, m, base, entries_dev, blk_ent_dev, total_blocks, tiles_per_block, op_buf, zero)
    if (!fma) return f64 ? QH_TB(double, false, false) : QH_TB(float, false, false);
    if (!cut) return f64 ? QH_TB(double, true, false) : QH_TB(float, true, false);
    return f64 ? QH_TB(double, true, true) : QH_TB(float, true, true);
#undef QH_TB
}

}  // namespace qvmhip
