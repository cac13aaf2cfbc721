// reduce.hip -- probability reductions, measurement collapse and sampling.
//
// Reductions accumulate |psi_i|^2 in FP64 for both state dtypes: each thread
// sums a grid-strided slice, a wavefront reduces with cross-lane shuffles, the
// waves of a block combine in LDS, and one block adds the block partials in a
// fixed order (deterministic: same input -> same bits).
//
// Sampling (measure-all and multi-shot) is a two-level search over the
// cumulative distribution C(b) = sum_{i<=b} |psi_i|^2:
//   1. chunk sums S[c] over chunks of 256 amplitudes (one wavefront per chunk),
//   2. exclusive prefix P[c] of S (device-wide scan),
//   3. per draw p, one wavefront: binary search over P for the chunk, then a
//      wavefront prefix scan of the chunk's 256 probabilities, C(b) =
//      P[c] + in-chunk prefix, and a ballot picks the first b that satisfies
//      the rule: C(b) > p for measure-all (src/measurement.lisp:179-225) or
//      C(b) >= p for multi-shot sampling (src/measurement.lisp:246-288).
//   The result is clamped to len-1 when no b qualifies, as in the reference.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

constexpr int kWave = 64;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Inclusive prefix over the 64 lanes (Hillis-Steele with shuffles).
__device__ inline double wave_inclusive_scan(double v, int lane) {
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const double t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

__device__ inline double block_sum(double v, double *sm) {
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    v = wave_sum(v);
    if (lane == 0) sm[wid] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0) {
        for (int w = 0; w < (int)(blockDim.x / kWave); ++w) t += sm[w];
    }
    return t;  // valid in thread 0
}

template <typename T>
__device__ inline double prob(const vec2_t<T> &a) {
    const double re = (double)a.x, im = (double)a.y;
    return re * re + im * im;
}

// out[0, count) = v (clears counters with a kernel, so the clear can sit in a
// captured graph).
__global__ void k_fill_u64(uint64_t *__restrict__ out, uint64_t count, uint64_t v) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (uint64_t)gridDim.x * blockDim.x)
        out[i] = v;
}

// mode 0: all amplitudes; 1: bit q = 1 half; 2: bit q = 0 half.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_prob_partial(const vec2_t<T> *__restrict__ psi,
                                                         uint64_t count, int mode, uint32_t q,
                                                         double *__restrict__ partial) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const uint64_t qbit = uint64_t(1) << q;
    double acc = 0.0;
    uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    for (; i + 3 * stride < count; i += 4 * stride) {
        vec2_t<T> v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            uint64_t a = i + u * stride;
            if (mode == 1) a = insert_zero(a, q) | qbit;
            if (mode == 2) a = insert_zero(a, q);
            v[u] = psi[a];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) acc += prob<T>(v[u]);
    }
    for (; i < count; i += stride) {
        uint64_t a = i;
        if (mode == 1) a = insert_zero(a, q) | qbit;
        if (mode == 2) a = insert_zero(a, q);
        acc += prob<T>(psi[a]);
    }
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

// sum over edges of <Z_a Z_b>: weight(i) = nedges - 2 * #cut edges.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_zz_partial(const vec2_t<T> *__restrict__ psi,
                                                       uint64_t count, const uint32_t *edges,
                                                       uint32_t nedges, double *__restrict__ partial) {
    __shared__ double sm[kBlock / kWave];
    __shared__ uint32_t se[2 * 256];
    for (uint32_t e = threadIdx.x; e < 2 * nedges && e < 512; e += kBlock) se[e] = edges[e];
    __syncthreads();
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        int cut = 0;
        for (uint32_t e = 0; e < nedges; ++e) cut += (int)(((i >> se[2 * e]) ^ (i >> se[2 * e + 1])) & 1u);
        acc += prob<T>(psi[i]) * (double)((int)nedges - 2 * cut);
    }
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_dot_partial(const vec2_t<TA> *__restrict__ a,
                                                        const vec2_t<TB> *__restrict__ b,
                                                        uint64_t count, double *__restrict__ partial) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    double re = 0.0, im = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const double ar = a[i].x, ai = a[i].y, br = b[i].x, bi = b[i].y;
        re += ar * br + ai * bi;  // conj(a) * b
        im += ar * bi - ai * br;
    }
    const double tr = block_sum(re, sm);
    __syncthreads();
    const double ti = block_sum(im, sm);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = tr;
        partial[2 * blockIdx.x + 1] = ti;
    }
}

// max |a_i - b_i| over re and im parts, in FP64.
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_maxdiff_partial(const vec2_t<TA> *__restrict__ a,
                                                            const vec2_t<TB> *__restrict__ b,
                                                            uint64_t count, double *__restrict__ partial) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    double m = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const double dr = fabs((double)a[i].x - (double)b[i].x);
        const double di = fabs((double)a[i].y - (double)b[i].y);
        m = fmax(m, fmax(dr, di));
        if (dr != dr || di != di) m = INFINITY;  // NaN counts as a mismatch
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, kWave));
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    if (lane == 0) sm[wid] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t = fmax(t, sm[w]);
        partial[blockIdx.x] = t;
    }
}

// sum |(|a_i|^2 - |b_i|^2)|: twice the total-variation distance of the two
// outcome distributions.
template <typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_tv_partial(const vec2_t<TA> *__restrict__ a,
                                                       const vec2_t<TB> *__restrict__ b,
                                                       uint64_t count, double *__restrict__ partial) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    double acc = 0.0;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride)
        acc += fabs(prob<TA>(a[i]) - prob<TB>(b[i]));
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) partial[blockIdx.x] = t;
}

__global__ __launch_bounds__(kBlock) void k_final_max(const double *__restrict__ partial, int n,
                                                      double *__restrict__ out) {
    if (threadIdx.x == 0) {
        double m = 0.0;
        for (int i = 0; i < n; ++i) m = fmax(m, partial[i]);
        out[0] = m;
    }
}

// One block adds `n` partials (stride `stride`, `width` interleaved values) in
// a fixed order.
__global__ __launch_bounds__(kBlock) void k_final_sum(const double *__restrict__ partial, int n,
                                                      int width, double *__restrict__ out) {
    __shared__ double sm[kBlock / kWave];
    for (int w = 0; w < width; ++w) {
        double acc = 0.0;
        for (int i = threadIdx.x; i < n; i += kBlock) acc += partial[(size_t)i * width + w];
        const double t = block_sum(acc, sm);
        if (threadIdx.x == 0) out[w] = t;
        __syncthreads();
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_force_measure(vec2_t<T> *__restrict__ psi, uint64_t len,
                                                          uint32_t q, int bit, double inv_norm,
                                                          int compiled_form) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    const T f = (T)inv_norm;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
        vec2_t<T> v = psi[i];
        const int ab = (int)((i >> q) & 1u);
        if (compiled_form) {
            // src/compile-gate.lisp:248-254: amp *= inv_norm * [bit matches]
            const T g = f * (T)(ab == bit ? 1 : 0);
            v.x = g * v.x;
            v.y = g * v.y;
        } else if (ab != bit) {
            // src/measurement.lisp:29-39: annihilated half -> 0, kept half *= inv_norm
            v.x = T(0);
            v.y = T(0);
        } else {
            v.x = f * v.x;
            v.y = f * v.y;
        }
        psi[i] = v;
    }
}

// ---------------------------------------------------------------------------
// Sampling.

// Lane l of the chunk's wavefront holds amplitudes 4l .. 4l+3.  The chunk
// total S[c] is lane 63's inclusive scan value; k_sample_search rebuilds the
// same lane sums, so its in-chunk prefix agrees with S[c] to rounding.
template <typename T>
__device__ inline void chunk_lane_probs(const vec2_t<T> *__restrict__ psi, uint64_t len,
                                        uint64_t chunk, int lane, double p[4]) {
    const uint64_t first = (chunk << kChunkLog2) + (uint64_t)lane * 4;
#pragma unroll
    for (int u = 0; u < 4; ++u) p[u] = (first + u < len) ? prob<T>(psi[first + u]) : 0.0;
}

// The four amplitudes of chunk_lane_probs, unconverted (zeros past len).
template <typename T>
__device__ inline void chunk_lane_load(const vec2_t<T> *__restrict__ psi, uint64_t len, uint64_t chunk, int lane,
                                       vec2_t<T> v[4]) {
    const uint64_t first = (chunk << kChunkLog2) + (uint64_t)lane * 4;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        v[u].x = v[u].y = T(0);
        if (first + u < len) v[u] = psi[first + u];
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_chunk_sums(const vec2_t<T> *__restrict__ psi,
                                                       uint64_t len, uint64_t nchunks,
                                                       double *__restrict__ sums) {
    const int lane = threadIdx.x % kWave;
    const uint64_t step = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t chunk = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; chunk < nchunks;
         chunk += step) {
        double p[4];
        chunk_lane_probs<T>(psi, len, chunk, lane, p);
        const double local = ((p[0] + p[1]) + p[2]) + p[3];
        const double incl = wave_inclusive_scan(local, lane);
        if (lane == kWave - 1) sums[chunk] = incl;
    }
}

// Device-wide exclusive scan of doubles, three phases.  Phase 1: totals of
// segments of kScanSeg values.
constexpr int kScanSeg = 1024;

__global__ __launch_bounds__(kBlock) void k_scan_seg_totals(const double *__restrict__ in, uint64_t n,
                                                            double *__restrict__ tot) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kScanSeg;
    double acc = 0.0;
    // each thread: 4 consecutive values, sequential
    const uint64_t j = base + (uint64_t)threadIdx.x * 4;
#pragma unroll
    for (int u = 0; u < 4; ++u)
        if (j + u < n) acc += in[j + u];
    // totals must be the same sums phase 3 builds: lane scan order
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const double incl = wave_inclusive_scan(acc, lane);
    if (lane == kWave - 1) sm[wid] = incl;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) t += sm[w];
        tot[blockIdx.x] = t;
    }
}

// Phase 2: one block, exclusive scan of the segment totals in place.
__global__ __launch_bounds__(kBlock) void k_scan_tops(double *__restrict__ tot, uint64_t nseg) {
    __shared__ double sm[kBlock];
    // each thread owns a contiguous run
    const uint64_t per = (nseg + kBlock - 1) / kBlock;
    const uint64_t lo = (uint64_t)threadIdx.x * per;
    const uint64_t hi = std::min<uint64_t>(lo + per, nseg);
    double acc = 0.0;
    for (uint64_t i = lo; i < hi; ++i) acc += tot[i];
    sm[threadIdx.x] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double run = 0.0;
        for (int t = 0; t < kBlock; ++t) {
            const double v = sm[t];
            sm[t] = run;
            run += v;
        }
    }
    __syncthreads();
    double run = sm[threadIdx.x];
    for (uint64_t i = lo; i < hi; ++i) {
        const double v = tot[i];
        tot[i] = run;
        run += v;
    }
}

// Phase 3: exclusive prefix of every value = segment offset + in-segment scan.
__global__ __launch_bounds__(kBlock) void k_scan_apply(const double *__restrict__ in, uint64_t n,
                                                       const double *__restrict__ seg_off,
                                                       double *__restrict__ out) {
    __shared__ double sm[kBlock / kWave];
    const uint64_t base = (uint64_t)blockIdx.x * kScanSeg;
    const uint64_t j = base + (uint64_t)threadIdx.x * 4;
    double v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) v[u] = (j + u < n) ? in[j + u] : 0.0;
    const double acc = ((v[0] + v[1]) + v[2]) + v[3];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    const double incl = wave_inclusive_scan(acc, lane);
    const double up = __shfl_up(incl, 1, kWave);
    if (lane == kWave - 1) sm[wid] = incl;
    __syncthreads();
    double woff = 0.0;
    for (int w = 0; w < wid; ++w) woff += sm[w];
    double run = seg_off[blockIdx.x] + (woff + (lane ? up : 0.0));
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        if (j + u < n) out[j + u] = run;
        run += v[u];
    }
}

// The index a draw p selects (rule_ge: first index with cdf >= p, else
// first with cdf > p; len - 1 when none): the last chunk whose exclusive
// prefix is below p by binary search, then the lanes' in-chunk prefix.
// Called by a whole wavefront with the same p; the result is valid in
// every lane.  *near (every lane) reports whether p lies within eps of the
// two cumulative values that decided the draw: C(b) and C(b - 1) as this
// search formed them (C(b - 1) of a chunk's first index is the chunk's
// exclusive prefix; for the clamped result, C(len - 1)).  eps <= 0 flags
// nothing.
template <typename T>
__device__ inline uint64_t sample_one(const vec2_t<T> *__restrict__ psi, uint64_t len, uint64_t nchunks,
                                      const double *__restrict__ pre, double p, int rule_ge, int lane,
                                      double eps, bool *near) {
    uint64_t lo = 0, hi = nchunks;  // invariant: answer in [lo, hi)
    while (hi - lo > 1) {
        const uint64_t mid = lo + (hi - lo) / 2;
        const double v = pre[mid];
        const bool ok = rule_ge ? (v < p) : (v <= p);
        if (ok)
            lo = mid;
        else
            hi = mid;
    }
    double last = 0.0;
    for (uint64_t c = lo; c < nchunks; ++c) {
        double pr[4];
        chunk_lane_probs<T>(psi, len, c, lane, pr);
        const double local = ((pr[0] + pr[1]) + pr[2]) + pr[3];
        const double incl = wave_inclusive_scan(local, lane);
        const double up = __shfl_up(incl, 1, kWave);
        const double pc = pre[c];
        double run = lane ? up : 0.0;  // exclusive lane offset
        double C[4];
        int hit = -1;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            run += pr[k];
            C[k] = pc + run;
            const bool ok = rule_ge ? (C[k] >= p) : (C[k] > p);
            const uint64_t idx = (c << kChunkLog2) + (uint64_t)lane * 4 + k;
            if (ok && hit < 0 && idx < len) hit = k;
        }
        const double prev_lane = __shfl_up(C[3], 1, kWave);
        const unsigned long long mask = __ballot(hit >= 0);
        if (mask) {
            const int first = __ffsll((long long)mask) - 1;
            double cb = 0.0, cp = 0.0;
#pragma unroll
            for (int k = 0; k < 4; ++k)
                if (k == hit) {
                    cb = C[k];
                    cp = k > 0 ? C[k > 0 ? k - 1 : 0] : (lane ? prev_lane : pc);
                }
            const bool nr = eps > 0.0 && (fabs(p - cb) < eps || fabs(p - cp) < eps);
            *near = __shfl((int)nr, first, kWave) != 0;
            const int k = __shfl(hit, first, kWave);
            return (c << kChunkLog2) + (uint64_t)first * 4 + (uint64_t)k;
        }
        last = __shfl(C[3], kWave - 1, kWave);
    }
    *near = eps > 0.0 && fabs(p - last) < eps;
    return len - 1;
}

// One wavefront per draw.  near_flags (optional) receives one byte per draw,
// near_count (optional) the number of near-boundary draws (added to).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_sample_search(const vec2_t<T> *__restrict__ psi,
                                                          uint64_t len, uint64_t nchunks,
                                                          const double *__restrict__ pre,
                                                          const double *__restrict__ u,
                                                          uint64_t nshots, int rule_ge,
                                                          uint64_t *__restrict__ out, double eps,
                                                          uint8_t *__restrict__ near_flags,
                                                          unsigned long long *__restrict__ near_count) {
    const int lane = threadIdx.x % kWave;
    const uint64_t sstep = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t shot = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; shot < nshots;
         shot += sstep) {
        bool nr = false;
        const uint64_t result = sample_one<T>(psi, len, nchunks, pre, u[shot], rule_ge, lane, eps, &nr);
        if (lane == 0) {
            out[shot] = result;
            if (near_flags) near_flags[shot] = nr ? 1 : 0;
            if (nr && near_count) atomicAdd(near_count, 1ull);
        }
    }
}

// ---------------------------------------------------------------------------
// Batched finish: every instance of a batch buffer at once.  A work block
// covers kFinishAmps amplitudes (64 sampling chunks) of one instance: each
// wave builds the chunk sums of 16 chunks and, reading the same
// probabilities, accumulates sum_x |psi_x|^2 (|E| - 2 cut(x)), the MaxCut
// <H_C> partial of the block.  A block first sets up per-lane edge tables,
// so it takes many chunks.

constexpr int kFinishChunksPerWave = 16;
constexpr uint64_t kFinishAmps = uint64_t(kBlock) * 4 * kFinishChunksPerWave;

template <typename T>
__global__ __launch_bounds__(kBlock) void k_bfinish_partial(const vec2_t<T> *__restrict__ buf,
                                                            const BatchFinishArgs a) {
    __shared__ double sm[kBlock / kWave];
    __shared__ uint32_t se[512];
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    for (uint64_t b = blockIdx.x; b < a.blocks; b += gridDim.x) {
        const uint32_t i = a.blk_inst[b];
        const uint64_t lb = b - a.blk_base[i];
        const uint32_t ne = a.nedges[i];
        __syncthreads();  // the previous block's edge and partial reads are done
        for (uint32_t k = threadIdx.x; k < 2 * ne; k += kBlock) se[k] = a.edges[a.edge_off[i] + k];
        __syncthreads();
        const vec2_t<T> *psi = buf + a.inst_off[i];
        const uint64_t len = a.inst_len[i];
        const uint64_t nch = a.chunk_base[i + 1] - a.chunk_base[i];
        // cut(x) for x = chunk * 256 + lo: edges with both endpoints in the
        // low 8 bits depend on lo only (win), edges with both above on the
        // chunk only (hh), and the cross edges are popcount(A(lo) ^ B(chunk))
        // with one bit per edge -- per amplitude a XOR and a popcount
        // instead of a loop over the edges
        const uint32_t nll = a.ecount[3 * i], ncr = a.ecount[3 * i + 1];
        const bool fast = ncr <= 64;
        int win[4] = {0, 0, 0, 0};
        uint64_t A[4] = {0, 0, 0, 0};
        if (ne && fast) {
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                const uint32_t lo = (uint32_t)lane * 4 + k;
                for (uint32_t e = 0; e < nll; ++e) win[k] += (int)(((lo >> se[2 * e]) ^ (lo >> se[2 * e + 1])) & 1u);
                for (uint32_t j = 0; j < ncr; ++j) A[k] |= (uint64_t)((lo >> se[2 * (nll + j)]) & 1u) << j;
            }
        }
        double acc = 0.0;
        const uint64_t c0 = lb * (kBlock / kWave * kFinishChunksPerWave) + (uint64_t)wid * kFinishChunksPerWave;
        // the next chunk's amplitudes are loaded while this one is summed
        vec2_t<T> cur[4], nxt[4];
        if (c0 < nch) chunk_lane_load<T>(psi, len, c0, lane, cur);
#pragma unroll 1
        for (int k4 = 0; k4 < kFinishChunksPerWave; ++k4) {
            const uint64_t c = c0 + k4;
            if (c >= nch) break;  // uniform per wave
            if (k4 + 1 < kFinishChunksPerWave && c + 1 < nch) chunk_lane_load<T>(psi, len, c + 1, lane, nxt);
            double p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] = prob<T>(cur[u]);
#pragma unroll
            for (int u = 0; u < 4; ++u) cur[u] = nxt[u];
            if (a.chunk_sum) {
                const double incl = wave_inclusive_scan(((p[0] + p[1]) + p[2]) + p[3], lane);
                if (lane == kWave - 1) a.chunk_sum[a.chunk_base[i] + c] = incl;
            }
            if (ne && fast) {
                // the chunk part: one mask of the cross edges' upper bits and
                // the count of cut edges above the chunk bits (uniform)
                uint64_t B = 0;
                int hh = 0;
                for (uint32_t j = 0; j < ncr; ++j)
                    B |= (uint64_t)((c >> (se[2 * (nll + j) + 1] - kChunkLog2)) & 1u) << j;
                for (uint32_t e = nll + ncr; e < ne; ++e)
                    hh += (int)(((c >> (se[2 * e] - kChunkLog2)) ^ (c >> (se[2 * e + 1] - kChunkLog2))) & 1u);
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int cut = win[k] + hh + __popcll(A[k] ^ B);
                    acc += p[k] * (double)((int)ne - 2 * cut);
                }
            } else if (ne) {
                const uint64_t x0 = (c << kChunkLog2) + (uint64_t)lane * 4;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const uint64_t x = x0 + k;
                    int cut = 0;
                    for (uint32_t e = 0; e < ne; ++e) cut += (int)(((x >> se[2 * e]) ^ (x >> se[2 * e + 1])) & 1u);
                    acc += p[k] * (double)((int)ne - 2 * cut);
                }
            }
        }
        if (a.zz_part) {
            const double t = block_sum(acc, sm);
            if (threadIdx.x == 0) a.zz_part[b] = t;
        }
    }
}

// One block per instance: its <H_C> from its block partials, in order.
__global__ __launch_bounds__(kBlock) void k_bfinish_zz(const BatchFinishArgs a) {
    __shared__ double sm[kBlock / kWave];
    const uint32_t i = blockIdx.x;
    if (a.nedges[i] == 0) return;  // uniform: the whole block leaves
    double acc = 0.0;
    for (uint64_t j = a.blk_base[i] + threadIdx.x; j < a.blk_base[i + 1]; j += kBlock) acc += a.zz_part[j];
    const double t = block_sum(acc, sm);
    if (threadIdx.x == 0) a.exp_out[i] = t;
}

// One block per instance: exclusive prefix of its chunk sums, 1024 at a
// time with a running carry.
__global__ __launch_bounds__(kBlock) void k_bfinish_scan(const BatchFinishArgs a) {
    __shared__ double sm[kBlock / kWave];
    const uint32_t i = blockIdx.x;
    const uint64_t base = a.chunk_base[i], n = a.chunk_base[i + 1] - base;
    const int lane = threadIdx.x % kWave, wid = threadIdx.x / kWave;
    double carry = 0.0;
    for (uint64_t off = 0; off < n; off += 4 * kBlock) {
        const uint64_t j = off + (uint64_t)threadIdx.x * 4;
        double v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = (j + u < n) ? a.chunk_sum[base + j + u] : 0.0;
        const double incl = wave_inclusive_scan(((v[0] + v[1]) + v[2]) + v[3], lane);
        const double up = __shfl_up(incl, 1, kWave);
        if (lane == kWave - 1) sm[wid] = incl;
        __syncthreads();
        double woff = 0.0, total = 0.0;
        for (int w = 0; w < kBlock / kWave; ++w) {
            if (w < wid) woff += sm[w];
            total += sm[w];
        }
        double run = carry + (woff + (lane ? up : 0.0));
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (j + u < n) a.chunk_pre[base + j + u] = run;
            run += v[u];
        }
        carry += total;
        __syncthreads();  // sm is rewritten by the next round
    }
}

// One wavefront per (instance, shot), rule cdf >= p.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bfinish_sample(const vec2_t<T> *__restrict__ buf,
                                                           const BatchFinishArgs a) {
    const int lane = threadIdx.x % kWave;
    const uint64_t total = (uint64_t)a.count * a.shots;
    const uint64_t sstep = (uint64_t)gridDim.x * (kBlock / kWave);
    for (uint64_t s = (uint64_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; s < total; s += sstep) {
        const uint32_t i = (uint32_t)(s / a.shots);
        const uint64_t base = a.chunk_base[i];
        bool nr = false;
        const uint64_t r = sample_one<T>(buf + a.inst_off[i], a.inst_len[i], a.chunk_base[i + 1] - base,
                                         a.chunk_pre + base, a.u[s], 1, lane, 0.0, &nr);
        if (lane == 0) a.shot_out[s] = r;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_collapse_basis(vec2_t<T> *__restrict__ psi, uint64_t len,
                                                           const uint64_t *__restrict__ index) {
    const uint64_t b = *index;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
        vec2_t<T> z;
        z.x = (i == b) ? T(1) : T(0);
        z.y = T(0);
        psi[i] = z;
    }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_probs(const vec2_t<T> *__restrict__ psi, uint64_t start,
                                                  uint64_t count, double *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride)
        out[i] = prob<T>(psi[start + i]);
}

__global__ __launch_bounds__(kBlock) void k_widen(const float2 *__restrict__ in, uint64_t count,
                                                  double2 *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const float2 v = in[i];
        out[i] = make_double2((double)v.x, (double)v.y);
    }
}

__global__ __launch_bounds__(kBlock) void k_narrow(const double2 *__restrict__ in, uint64_t count,
                                                   float2 *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const double2 v = in[i];
        out[i] = make_float2((float)v.x, (float)v.y);
    }
}

// dst += src, elementwise (the running sum of a Kraus-list density evolution).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_accumulate(vec2_t<T> *__restrict__ dst, const vec2_t<T> *__restrict__ src,
                                                       uint64_t len) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < len; i += stride) {
        vec2_t<T> a = dst[i];
        const vec2_t<T> b = src[i];
        a.x = a.x + b.x;
        a.y = a.y + b.y;
        dst[i] = a;
    }
}

// out[i] = Re psi[i * (dim + 1)], i < dim: the diagonal of a density matrix
// stored row-major as a vector of dim^2 amplitudes.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_diag_gather(const vec2_t<T> *__restrict__ psi, uint64_t dim,
                                                        double *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < dim; i += stride)
        out[i] = (double)psi[i * (dim + 1)].x;
}

// Host export of a range as 64-bit words: amplitudes as (re, im) doubles or
// probabilities re*re + im*im, in native or big-endian byte order (the
// server's octet-stream replies, app/src/utilities.lisp:40-87).
template <typename T, bool PROBS, bool BE>
__global__ __launch_bounds__(kBlock) void k_export(const vec2_t<T> *__restrict__ psi, uint64_t start,
                                                   uint64_t count, uint64_t *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < count; i += stride) {
        const vec2_t<T> a = psi[start + i];
        if constexpr (PROBS) {
            const uint64_t w = __double_as_longlong(prob<T>(a));
            out[i] = BE ? __builtin_bswap64(w) : w;
        } else {
            const uint64_t wr = __double_as_longlong((double)a.x), wi = __double_as_longlong((double)a.y);
            out[2 * i] = BE ? __builtin_bswap64(wr) : wr;
            out[2 * i + 1] = BE ? __builtin_bswap64(wi) : wi;
        }
    }
}

// ---------------------------------------------------------------------------
// Host launchers.

namespace {
uint32_t red_grid(uint64_t count) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((count + kBlock - 1) / kBlock, kRedBlocks));
}
uint32_t sweep_grid(uint64_t count) {
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>((count + kBlock - 1) / kBlock, 256 * 64));
}
}  // namespace

int reduce_probability(qvmhip_state *s, int mode, uint32_t q, double *out_dev) {
    const uint64_t count = mode == 0 ? s->len : s->len / 2;
    const uint32_t grid = red_grid(count);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_prob_partial<double>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (const double2 *)s->amps, count, mode, q, s->red);
    else
        hipLaunchKernelGGL(k_prob_partial<float>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (const float2 *)s->amps, count, mode, q, s->red);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, s->stream, s->red, (int)grid, 1, out_dev);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int reduce_norm2(qvmhip_state *s, double *out_dev) { return reduce_probability(s, 0, 0, out_dev); }

int launch_force_measure(qvmhip_state *s, uint32_t q, int bit, double inv_norm, int compiled_form) {
    const uint32_t grid = sweep_grid(s->len);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_force_measure<double>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (double2 *)s->amps, s->len, q, bit, inv_norm, compiled_form);
    else
        hipLaunchKernelGGL(k_force_measure<float>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (float2 *)s->amps, s->len, q, bit, inv_norm, compiled_form);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

// Chunk sums + exclusive chunk prefix of a len-amplitude state at `amps`
// into chunk_sum / chunk_pre (nchunks = ceil(len / 256) entries each);
// scan_tmp holds ceil(nchunks / kScanSeg) segment totals.
int launch_chunk_prefix(hipStream_t st, int dtype, const void *amps, uint64_t len, double *chunk_sum,
                        double *scan_tmp, double *chunk_pre) {
    const uint64_t nch = (len + (uint64_t(1) << kChunkLog2) - 1) >> kChunkLog2;
    const uint32_t g1 = (uint32_t)std::min<uint64_t>((nch + (kBlock / kWave) - 1) / (kBlock / kWave), kMaxGrid);
    if (dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_chunk_sums<double>, dim3(g1), dim3(kBlock), 0, st, (const double2 *)amps, len, nch,
                           chunk_sum);
    else
        hipLaunchKernelGGL(k_chunk_sums<float>, dim3(g1), dim3(kBlock), 0, st, (const float2 *)amps, len, nch,
                           chunk_sum);
    QH_LAUNCH_CHECK();
    const uint64_t nseg = (nch + kScanSeg - 1) / kScanSeg;
    hipLaunchKernelGGL(k_scan_seg_totals, dim3((uint32_t)nseg), dim3(kBlock), 0, st, chunk_sum, nch, scan_tmp);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_tops, dim3(1), dim3(kBlock), 0, st, scan_tmp, nseg);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_scan_apply, dim3((uint32_t)nseg), dim3(kBlock), 0, st, chunk_sum, nch, scan_tmp,
                       chunk_pre);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

// One search per draw over a prefix built by launch_chunk_prefix; u, out,
// near_flags and near_count are device pointers (the last two optional).
int launch_sample_search(hipStream_t st, int dtype, const void *amps, uint64_t len, const double *chunk_pre,
                         const double *u, uint64_t nshots, int rule_ge, uint64_t *out, double eps,
                         uint8_t *near_flags, unsigned long long *near_count) {
    if (nshots == 0) return QVMHIP_OK;
    const uint64_t nch = (len + (uint64_t(1) << kChunkLog2) - 1) >> kChunkLog2;
    const uint64_t wpb = kBlock / kWave;
    const uint32_t nb = (uint32_t)std::min<uint64_t>((nshots + wpb - 1) / wpb, kMaxGrid);
    if (dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_sample_search<double>, dim3(nb), dim3(kBlock), 0, st, (const double2 *)amps, len, nch,
                           chunk_pre, u, nshots, rule_ge, out, eps, near_flags, near_count);
    else
        hipLaunchKernelGGL(k_sample_search<float>, dim3(nb), dim3(kBlock), 0, st, (const float2 *)amps, len, nch,
                           chunk_pre, u, nshots, rule_ge, out, eps, near_flags, near_count);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

static int build_chunk_prefix(qvmhip_state *s) {
    return launch_chunk_prefix(s->stream, s->dtype, s->amps, s->len, s->chunk_sum, s->scan_tmp, s->chunk_pre);
}

static int ensure_shots(qvmhip_state *s, uint64_t nshots) {
    if (nshots <= s->shot_cap) return QVMHIP_OK;
    if (s->shot_out) (void)hipFree(s->shot_out);
    if (s->shot_u) (void)hipFree(s->shot_u);
    if (s->shot_near) (void)hipFree(s->shot_near);
    s->shot_out = nullptr;
    s->shot_u = nullptr;
    s->shot_near = nullptr;
    s->shot_cap = 0;
    QH_CHECK(hipMalloc((void **)&s->shot_out, nshots * sizeof(uint64_t)));
    QH_CHECK(hipMalloc((void **)&s->shot_u, nshots * sizeof(double)));
    QH_CHECK(hipMalloc((void **)&s->shot_near, nshots));
    s->shot_cap = nshots;
    return QVMHIP_OK;
}

// Draws nshots samples; u, out, near_flags and n_near are host pointers
// (out may be null when the caller only needs the device copy in
// s->shot_out; near_flags and n_near are optional).  Draws within eps of the
// cumulative values that decided them are flagged and counted (eps <= 0:
// none).
int sample_shots(qvmhip_state *s, const double *u, uint64_t nshots, uint64_t *out, int rule_ge, double eps,
                 uint8_t *near_flags, uint64_t *n_near) {
    if (n_near) *n_near = 0;
    if (nshots == 0) return QVMHIP_OK;
    // Draws, results and flags travel through the pinned staging buffer in
    // batches of 17 bytes per draw.
    const uint64_t batch = std::min<uint64_t>(nshots, s->stage_bytes / 17);
    int rc = ensure_shots(s, batch);
    if (rc) return rc;
    rc = build_chunk_prefix(s);
    if (rc) return rc;
    unsigned long long *count_dev = (unsigned long long *)(s->red + kRedBlocks * 2 + 8);
    if (n_near) {
        hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, s->stream, (uint64_t *)count_dev, uint64_t(1),
                           uint64_t(0));
        QH_LAUNCH_CHECK();
    }
    double *stage_u = (double *)s->stage;
    uint64_t *stage_out = (uint64_t *)((char *)s->stage + batch * 8);
    uint8_t *stage_near = (uint8_t *)s->stage + batch * 16;
    for (uint64_t first = 0; first < nshots; first += batch) {
        const uint64_t c = std::min<uint64_t>(batch, nshots - first);
        QH_CHECK(hipStreamSynchronize(s->stream));  // staging buffer is free
        std::memcpy(stage_u, u + first, c * sizeof(double));
        QH_CHECK(hipMemcpyAsync(s->shot_u, stage_u, c * sizeof(double), hipMemcpyHostToDevice, s->stream));
        rc = launch_sample_search(s->stream, s->dtype, s->amps, s->len, s->chunk_pre, s->shot_u, c, rule_ge,
                                  s->shot_out, eps, near_flags ? s->shot_near : nullptr,
                                  n_near ? count_dev : nullptr);
        if (rc) return rc;
        if (out || near_flags) {
            if (out)
                QH_CHECK(hipMemcpyAsync(stage_out, s->shot_out, c * sizeof(uint64_t), hipMemcpyDeviceToHost,
                                        s->stream));
            if (near_flags)
                QH_CHECK(hipMemcpyAsync(stage_near, s->shot_near, c, hipMemcpyDeviceToHost, s->stream));
            QH_CHECK(hipStreamSynchronize(s->stream));
            if (out) std::memcpy(out + first, stage_out, c * sizeof(uint64_t));
            if (near_flags) std::memcpy(near_flags + first, stage_near, c);
        }
    }
    if (n_near) {
        QH_CHECK(hipMemcpyAsync(s->host, count_dev, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
        QH_CHECK(hipStreamSynchronize(s->stream));
        std::memcpy(n_near, s->host, sizeof(uint64_t));
    }
    return QVMHIP_OK;
}

int sample_to_device(qvmhip_state *s, const double *u_dev, uint64_t nshots, uint64_t *out_dev, int rule_ge) {
    if (nshots == 0) return QVMHIP_OK;
    int rc = build_chunk_prefix(s);
    if (rc) return rc;
    return launch_sample_search(s->stream, s->dtype, s->amps, s->len, s->chunk_pre, u_dev, nshots, rule_ge, out_dev,
                                0.0, nullptr, nullptr);
}

int expectation_zz_to_device(qvmhip_state *s, const uint32_t *edges_dev, uint32_t nedges, double *out_dev) {
    if (nedges > 256) return fail(QVMHIP_EINVAL, "at most 256 edges");
    const uint32_t grid = red_grid(s->len);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_zz_partial<double>, dim3(grid), dim3(kBlock), 0, s->stream, (const double2 *)s->amps,
                           s->len, edges_dev, nedges, s->red);
    else
        hipLaunchKernelGGL(k_zz_partial<float>, dim3(grid), dim3(kBlock), 0, s->stream, (const float2 *)s->amps,
                           s->len, edges_dev, nedges, s->red);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, s->stream, s->red, (int)grid, 1, out_dev);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int collapse_to_basis(qvmhip_state *s, const uint64_t *index_dev) {
    const uint32_t grid = sweep_grid(s->len);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_collapse_basis<double>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (double2 *)s->amps, s->len, index_dev);
    else
        hipLaunchKernelGGL(k_collapse_basis<float>, dim3(grid), dim3(kBlock), 0, s->stream,
                           (float2 *)s->amps, s->len, index_dev);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int expectation_zz(qvmhip_state *s, const uint32_t *edges, uint32_t nedges, double *out) {
    if (nedges > 256) return fail(QVMHIP_EINVAL, "at most 256 edges");
    // the edges go to the state's matrix slot (2 KiB of its 128 KiB), in
    // stream order after every earlier use of the slot: no allocation
    uint32_t *dedges = (uint32_t *)s->mat;
    if (nedges)
        QH_CHECK(hipMemcpyAsync(dedges, edges, sizeof(uint32_t) * 2 * nedges, hipMemcpyHostToDevice, s->stream));
    int rc = expectation_zz_to_device(s, dedges, nedges, s->red + kRedBlocks * 2);
    if (rc) return rc;
    QH_CHECK(hipMemcpyAsync(s->host, s->red + kRedBlocks * 2, sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
    QH_CHECK(hipStreamSynchronize(s->stream));
    *out = s->host[0];
    return QVMHIP_OK;
}

template <template <typename, typename> class Launch>
static void dispatch2(qvmhip_state *a, qvmhip_state *b, uint32_t grid) {
    if (a->dtype == QVMHIP_F64 && b->dtype == QVMHIP_F64) Launch<double, double>::go(a, b, grid);
    else if (a->dtype == QVMHIP_F64) Launch<double, float>::go(a, b, grid);
    else if (b->dtype == QVMHIP_F64) Launch<float, double>::go(a, b, grid);
    else Launch<float, float>::go(a, b, grid);
}

template <typename TA, typename TB> struct DotLaunch {
    static void go(qvmhip_state *a, qvmhip_state *b, uint32_t grid) {
        hipLaunchKernelGGL((k_dot_partial<TA, TB>), dim3(grid), dim3(kBlock), 0, a->stream,
                           (const vec2_t<TA> *)a->amps, (const vec2_t<TB> *)b->amps, a->len, a->red);
    }
};

template <typename TA, typename TB> struct DiffLaunch {
    static void go(qvmhip_state *a, qvmhip_state *b, uint32_t grid) {
        hipLaunchKernelGGL((k_maxdiff_partial<TA, TB>), dim3(grid), dim3(kBlock), 0, a->stream,
                           (const vec2_t<TA> *)a->amps, (const vec2_t<TB> *)b->amps, a->len, a->red);
    }
};

template <typename TA, typename TB> struct TvLaunch {
    static void go(qvmhip_state *a, qvmhip_state *b, uint32_t grid) {
        hipLaunchKernelGGL((k_tv_partial<TA, TB>), dim3(grid), dim3(kBlock), 0, a->stream,
                           (const vec2_t<TA> *)a->amps, (const vec2_t<TB> *)b->amps, a->len, a->red);
    }
};

int tv_distance(qvmhip_state *a, qvmhip_state *b, double *out) {
    const uint32_t grid = red_grid(a->len);
    dispatch2<TvLaunch>(a, b, grid);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, a->stream, a->red, (int)grid, 1,
                       a->red + kRedBlocks * 2);
    QH_LAUNCH_CHECK();
    QH_CHECK(hipMemcpyAsync(a->host, a->red + kRedBlocks * 2, sizeof(double), hipMemcpyDeviceToHost, a->stream));
    QH_CHECK(hipStreamSynchronize(a->stream));
    *out = 0.5 * a->host[0];
    return QVMHIP_OK;
}

int max_abs_diff(qvmhip_state *a, qvmhip_state *b, double *out) {
    const uint32_t grid = red_grid(a->len);
    dispatch2<DiffLaunch>(a, b, grid);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_max, dim3(1), dim3(64), 0, a->stream, a->red, (int)grid, a->red + kRedBlocks * 2);
    QH_LAUNCH_CHECK();
    QH_CHECK(hipMemcpyAsync(a->host, a->red + kRedBlocks * 2, sizeof(double), hipMemcpyDeviceToHost, a->stream));
    QH_CHECK(hipStreamSynchronize(a->stream));
    *out = a->host[0];
    return QVMHIP_OK;
}

int inner_product(qvmhip_state *a, qvmhip_state *b, double *re, double *im) {
    const uint32_t grid = red_grid(a->len);
    dispatch2<DotLaunch>(a, b, grid);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(kBlock), 0, a->stream, a->red, (int)grid, 2,
                       a->red + kRedBlocks * 2);
    QH_LAUNCH_CHECK();
    QH_CHECK(hipMemcpyAsync(a->host, a->red + kRedBlocks * 2, 2 * sizeof(double), hipMemcpyDeviceToHost,
                            a->stream));
    QH_CHECK(hipStreamSynchronize(a->stream));
    *re = a->host[0];
    *im = a->host[1];
    return QVMHIP_OK;
}

namespace {
template <typename T>
void launch_export(qvmhip_state *s, int mode, uint64_t start, uint64_t count, uint64_t *dev) {
    const vec2_t<T> *psi = (const vec2_t<T> *)s->amps;
    const dim3 grid(sweep_grid(count)), block(kBlock);
    switch (mode) {
        case kExportAmps: hipLaunchKernelGGL((k_export<T, false, false>), grid, block, 0, s->stream, psi, start, count, dev); break;
        case kExportAmpsBE: hipLaunchKernelGGL((k_export<T, false, true>), grid, block, 0, s->stream, psi, start, count, dev); break;
        case kExportProbs: hipLaunchKernelGGL((k_export<T, true, false>), grid, block, 0, s->stream, psi, start, count, dev); break;
        default: hipLaunchKernelGGL((k_export<T, true, true>), grid, block, 0, s->stream, psi, start, count, dev); break;
    }
}
}  // namespace

// Chunked device -> host export.  The pinned stage is split in two halves;
// chunk i is converted on the device and copied into half i%2 while the host
// copies chunk i-1 out of the other half.  FP64 amplitudes in native order
// need no conversion and are copied straight from the state.
int export_range(qvmhip_state *s, uint64_t start, uint64_t count, int mode, void *host_out) {
    const size_t eb = (mode == kExportAmps || mode == kExportAmpsBE) ? 16 : 8;
    const bool direct = mode == kExportAmps && s->dtype == QVMHIP_F64;
    const uint64_t per = (s->stage_bytes / 2) / eb;
    uint64_t *dev[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int rc = QVMHIP_OK;
    auto cleanup = [&]() {
        for (int h = 0; h < 2; ++h) {
            if (dev[h]) (void)hipFree(dev[h]);
            if (ev[h]) (void)hipEventDestroy(ev[h]);
        }
    };
    for (int h = 0; h < 2; ++h) {
        hipError_t e = hipEventCreateWithFlags(&ev[h], hipEventDisableTiming);
        if (e == hipSuccess && !direct && count > (uint64_t)h * per)
            e = hipMalloc((void **)&dev[h], std::min<uint64_t>(per, count) * eb);
        if (e != hipSuccess) {
            cleanup();
            return fail(e == hipErrorOutOfMemory ? QVMHIP_ENOMEM : QVMHIP_EDEVICE,
                        std::string("export buffers: ") + hipGetErrorString(e));
        }
    }
    char *out = (char *)host_out;
    char *half[2] = {(char *)s->stage, (char *)s->stage + per * eb};
    const uint64_t nchunks = (count + per - 1) / per;
    for (uint64_t i = 0; i <= nchunks && rc == QVMHIP_OK; ++i) {
        if (i < nchunks) {
            const int h = (int)(i & 1);
            const uint64_t off = i * per, c = std::min<uint64_t>(per, count - off);
            const void *src;
            if (direct) {
                src = (const double2 *)s->amps + start + off;
            } else {
                if (s->dtype == QVMHIP_F64)
                    launch_export<double>(s, mode, start + off, c, dev[h]);
                else
                    launch_export<float>(s, mode, start + off, c, dev[h]);
                hipError_t e = hipGetLastError();
                if (e != hipSuccess) {
                    rc = fail(QVMHIP_EDEVICE, std::string("export kernel: ") + hipGetErrorString(e));
                    break;
                }
                src = dev[h];
            }
            hipError_t e = hipMemcpyAsync(half[h], src, c * eb, hipMemcpyDeviceToHost, s->stream);
            if (e == hipSuccess) e = hipEventRecord(ev[h], s->stream);
            if (e != hipSuccess) {
                rc = fail(QVMHIP_EDEVICE, std::string("export copy: ") + hipGetErrorString(e));
                break;
            }
        }
        if (i >= 1) {
            const uint64_t j = i - 1;
            const int h = (int)(j & 1);
            const uint64_t off = j * per, c = std::min<uint64_t>(per, count - off);
            hipError_t e = hipEventSynchronize(ev[h]);
            if (e != hipSuccess) {
                rc = fail(QVMHIP_EDEVICE, std::string("export wait: ") + hipGetErrorString(e));
                break;
            }
            std::memcpy(out + off * eb, half[h], c * eb);
        }
    }
    if (rc != QVMHIP_OK) (void)hipStreamSynchronize(s->stream);  // no copy may still target the stage
    cleanup();
    return rc;
}

int probabilities(qvmhip_state *s, uint64_t start, uint64_t count, double *host_out) {
    return export_range(s, start, count, kExportProbs, host_out);
}

int accumulate(qvmhip_state *dst, const qvmhip_state *src) {
    const uint32_t grid = sweep_grid(dst->len);
    if (dst->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_accumulate<double>, dim3(grid), dim3(kBlock), 0, dst->stream, (double2 *)dst->amps,
                           (const double2 *)src->amps, dst->len);
    else
        hipLaunchKernelGGL(k_accumulate<float>, dim3(grid), dim3(kBlock), 0, dst->stream, (float2 *)dst->amps,
                           (const float2 *)src->amps, dst->len);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int density_diagonal(qvmhip_state *s, uint32_t nq, double *host_out) {
    const uint64_t dim = uint64_t(1) << nq;
    double *dev = nullptr;
    QH_CHECK(hipMalloc((void **)&dev, dim * sizeof(double)));
    const uint32_t grid = sweep_grid(dim);
    if (s->dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_diag_gather<double>, dim3(grid), dim3(kBlock), 0, s->stream, (const double2 *)s->amps,
                           dim, dev);
    else
        hipLaunchKernelGGL(k_diag_gather<float>, dim3(grid), dim3(kBlock), 0, s->stream, (const float2 *)s->amps,
                           dim, dev);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(host_out, dev, dim * sizeof(double), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dev);
    if (e != hipSuccess) return fail(QVMHIP_EDEVICE, std::string("density diagonal: ") + hipGetErrorString(e));
    return QVMHIP_OK;
}

uint64_t batch_finish_blocks(uint64_t len) { return (len + kFinishAmps - 1) / kFinishAmps; }

int launch_batch_finish(hipStream_t st, int dtype, const void *buf, const BatchFinishArgs &a) {
    if (a.blocks == 0) return QVMHIP_OK;
    const uint32_t g1 = (uint32_t)std::min<uint64_t>(a.blocks, kMaxGrid);
    if (dtype == QVMHIP_F64)
        hipLaunchKernelGGL(k_bfinish_partial<double>, dim3(g1), dim3(kBlock), 0, st, (const double2 *)buf, a);
    else
        hipLaunchKernelGGL(k_bfinish_partial<float>, dim3(g1), dim3(kBlock), 0, st, (const float2 *)buf, a);
    QH_LAUNCH_CHECK();
    if (a.zz_part) {
        hipLaunchKernelGGL(k_bfinish_zz, dim3(a.count), dim3(kBlock), 0, st, a);
        QH_LAUNCH_CHECK();
    }
    if (a.shots) {
        hipLaunchKernelGGL(k_bfinish_scan, dim3(a.count), dim3(kBlock), 0, st, a);
        QH_LAUNCH_CHECK();
        const uint64_t wpb = kBlock / kWave;
        const uint32_t g = (uint32_t)std::min<uint64_t>(((uint64_t)a.count * a.shots + wpb - 1) / wpb, kMaxGrid);
        if (dtype == QVMHIP_F64)
            hipLaunchKernelGGL(k_bfinish_sample<double>, dim3(g), dim3(kBlock), 0, st, (const double2 *)buf, a);
        else
            hipLaunchKernelGGL(k_bfinish_sample<float>, dim3(g), dim3(kBlock), 0, st, (const float2 *)buf, a);
        QH_LAUNCH_CHECK();
    }
    return QVMHIP_OK;
}

int convert_f32_to_f64(qvmhip_state *s, uint64_t start, uint64_t count, double *dev_out) {
    hipLaunchKernelGGL(k_widen, dim3(sweep_grid(count)), dim3(kBlock), 0, s->stream,
                       (const float2 *)s->amps + start, count, (double2 *)dev_out);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

int convert_f64_to_f32(qvmhip_state *s, uint64_t start, uint64_t count, const double *dev_in) {
    hipLaunchKernelGGL(k_narrow, dim3(sweep_grid(count)), dim3(kBlock), 0, s->stream,
                       (const double2 *)dev_in, count, (float2 *)s->amps + start);
    QH_LAUNCH_CHECK();
    return QVMHIP_OK;
}

}  // namespace qvmhip
