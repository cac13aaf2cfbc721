// qvmhip_internal.h -- shared definitions of the qvmhip engine (CDNA4 / gfx950).
//
// Statevector layout in HBM: 2^n interleaved complex amplitudes, qubit q =
// bit q of the index.  FP64 states are double2 (16 B/amp), FP32 states are
// float2 (8 B/amp).  All index arithmetic is 64-bit so 34-qubit FP32 states
// (2^34 amps, 128 GiB) are addressable.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qvmhip.h"

namespace qvmhip {

constexpr int kBlock = 256;         // threads per workgroup for streaming kernels
constexpr int kMaxK = 6;            // widest dense block (64 x 64)
constexpr int kMaxSorted = 16;      // targets + controls of one gate
constexpr int kRedBlocks = 2048;    // partial sums of one reduction launch
constexpr int kChunkLog2 = 8;       // sampling chunk: 256 amplitudes
// A dispatch holds at most 2^32 work-items; sweep kernels cap their grid at
// 2^22 workgroups and grid-stride over the rest (34-qubit states).
constexpr uint64_t kMaxGrid = uint64_t(1) << 22;

// Workgroups of a dispatch go to the 8 XCDs round robin: workgroup b runs on
// XCD b % 8.  xcd_block(b, nb) renumbers the nb workgroups so that each XCD
// takes one contiguous share of [0, nb), the x-th share for XCD x: each XCD
// then streams through its own part of the state (its own pages for its
// TLB) instead of every eighth 4 KiB piece of all of it.  Measured on a
// 16 GiB in-place sweep, one amplitude per thread: 6.74 vs 6.60 TB/s.  A
// bijection for every nb (the shares differ by at most one workgroup), so
// it changes which workgroup does what, never the result.
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t q = nb >> 3, r = nb & 7u, x = b & 7u;
    return x * q + (x < r ? x : r) + (b >> 3);
}

// The persistent-grid form: items [0, n) are cut into 8 contiguous shares,
// share x for the workgroups on XCD x; a workgroup takes items rank, rank +
// count, ... of its share (rank among the count workgroups of its XCD).  With
// fewer than 8 workgroups every workgroup strides over all items.
struct XcdShare {
    uint64_t lo, hi;
    uint32_t rank, count;
};
__device__ __forceinline__ XcdShare xcd_share(uint32_t b, uint32_t nb, uint64_t n) {
    XcdShare s;
    if (nb < 8) {
        s.lo = 0;
        s.hi = n;
        s.rank = b;
        s.count = nb;
        return s;
    }
    const uint32_t x = b & 7u, r = nb & 7u;
    s.rank = b >> 3;
    s.count = (nb >> 3) + (x < r ? 1u : 0u);
    const uint64_t per = n >> 3, rem = n & 7u;
    s.lo = x * per + (x < rem ? x : rem);
    s.hi = s.lo + per + (x < rem ? 1u : 0u);
    return s;
}

template <typename T> struct Vec2;
template <> struct Vec2<double> { using type = double2; };
template <> struct Vec2<float> { using type = float2; };
template <typename T> using vec2_t = typename Vec2<T>::type;

// Complement iteration: inject zeros at `sorted` (ascending) positions, then
// OR the control mask.  This is map-complement (src/wavefunction.lisp:155-175)
// with the control qubits of a controlled gate pinned to 1.
struct GroupSpec {
    uint32_t nsorted;
    uint32_t sorted[kMaxSorted];
    uint64_t ctrl_mask;
    uint64_t ngroups;
};

__host__ __device__ inline uint64_t insert_zero(uint64_t x, uint32_t q) {
    const uint64_t lo = x & ((uint64_t(1) << q) - 1);
    return ((x >> q) << (q + 1)) | lo;
}

__device__ inline uint64_t group_base(const GroupSpec &g, uint64_t i) {
#pragma unroll 1
    for (uint32_t j = 0; j < g.nsorted; ++j) i = insert_zero(i, g.sorted[j]);
    return i | g.ctrl_mask;
}

// Streaming loads/stores.  NT = true marks them non-temporal (the state is
// swept once per gate and never re-read from cache); measured on MI355X this
// is faster when each contiguous run of the access pattern is >= 128 B, and
// slower for shorter runs.
template <typename V> struct NativeVec;
template <> struct NativeVec<double2> { typedef double type __attribute__((ext_vector_type(2))); };
template <> struct NativeVec<float2> { typedef float type __attribute__((ext_vector_type(2))); };

template <bool NT, typename V>
__device__ inline V ldg(const V *p) {
    if constexpr (NT) {
        typedef typename NativeVec<V>::type N;
        const N v = __builtin_nontemporal_load((const N *)p);
        V r;
        r.x = v.x;
        r.y = v.y;
        return r;
    } else {
        return *p;
    }
}

template <bool NT, typename V>
__device__ inline void stg(V *p, const V &x) {
    if constexpr (NT) {
        typedef typename NativeVec<V>::type N;
        N v;
        v.x = x.x;
        v.y = x.y;
        __builtin_nontemporal_store(v, (N *)p);
    } else {
        *p = x;
    }
}

// Error reporting: thread-local message + status code.
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

// Makes `dev` the calling thread's current device for the duration of one
// C-ABI call, so that what a handle allocates and launches goes to the device
// it was created on whatever device the caller has current, and restores the
// caller's device on return.
struct DeviceScope {
    int prev = -1;
    explicit DeviceScope(int dev) {
        int cur = -1;
        if (hipGetDevice(&cur) == hipSuccess && cur != dev && hipSetDevice(dev) == hipSuccess) prev = cur;
    }
    ~DeviceScope() {
        if (prev >= 0) (void)hipSetDevice(prev);
    }
    DeviceScope(const DeviceScope &) = delete;
    DeviceScope &operator=(const DeviceScope &) = delete;
};

// Orders `waiter` after everything queued so far on `done` (a kernel on one
// handle's stream that reads another handle's state must finish before the
// other handle's next call can change it).
inline hipError_t stream_after(hipStream_t waiter, hipStream_t done) {
    if (waiter == done) return hipSuccess;
    hipEvent_t ev = nullptr;
    hipError_t e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
    if (e != hipSuccess) return e;
    e = hipEventRecord(ev, done);
    if (e == hipSuccess) e = hipStreamWaitEvent(waiter, ev, 0);
    (void)hipEventDestroy(ev);  // released once the recorded work has completed
    return e;
}

}  // namespace qvmhip

// The state handle behind the opaque qvmhip_state of the C ABI.
struct qvmhip_state {
    uint32_t n = 0;
    int dtype = QVMHIP_F64;
    int device = 0;
    uint64_t len = 0;          // 2^n
    size_t amp_bytes = 16;     // 16 (F64) or 8 (F32)
    void *amps = nullptr;      // device statevector
    hipStream_t stream = nullptr;
    bool own_stream = true;
    // scratch
    double *red = nullptr;          // kRedBlocks partials + results
    double *host = nullptr;         // pinned host scalars (64 doubles)
    void *mat = nullptr;            // device matrix slot (2 x 64 KiB)
    double *chunk_sum = nullptr;    // sampling: per-chunk probability sums
    double *chunk_pre = nullptr;    // sampling: exclusive prefix of chunk sums
    double *scan_tmp = nullptr;     // scan block sums
    uint64_t nchunks = 0;
    uint64_t *shot_out = nullptr;   // device shot results
    double *shot_u = nullptr;       // device uniforms
    uint8_t *shot_near = nullptr;   // device near-boundary flags
    uint64_t shot_cap = 0;
    uint64_t *host_shots = nullptr; // pinned
    double *host_u = nullptr;       // pinned
    void *stage = nullptr;          // pinned host staging for read/write
    size_t stage_bytes = 0;
    uint32_t flags = 0;
    // live kernel timing (qvmhip_profile_*): one event pair per launch
    struct ProfRec {
        int kind;
        hipEvent_t a, b;
    };
    bool profiling = false;
    std::vector<ProfRec> prof;
    std::vector<hipEvent_t> event_pool;
    // programs recorded against this state: destroying the state detaches
    // them (their state pointer becomes null), so a program destroyed later
    // frees only its own graph and operands
    std::vector<qvmhip_program *> programs;
};

namespace qvmhip {
// Live per-launch timing.  prof_begin/prof_end bracket one kernel launch
// with HIP events on the state's stream when profiling is on.
inline hipEvent_t prof_event(qvmhip_state *s) {
    if (!s->event_pool.empty()) {
        hipEvent_t e = s->event_pool.back();
        s->event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}
inline void prof_begin(qvmhip_state *s, int kind) {
    if (!s->profiling) return;
    qvmhip_state::ProfRec r{kind, prof_event(s), prof_event(s)};
    if (r.a) (void)hipEventRecord(r.a, s->stream);
    s->prof.push_back(r);
}
inline void prof_end(qvmhip_state *s) {
    if (!s->profiling || s->prof.empty()) return;
    if (s->prof.back().b) (void)hipEventRecord(s->prof.back().b, s->stream);
}

// Launch helpers implemented in the kernel files.
int launch_gate(qvmhip_state *s, const double *U, uint32_t k, const uint32_t *qubits, int force_dense);
int launch_diagonal(qvmhip_state *s, const double *d, uint32_t k, const uint32_t *qubits);
int launch_permutation(qvmhip_state *s, const uint32_t *perm, uint32_t k, const uint32_t *qubits);
int launch_fill_zero_state(qvmhip_state *s);
void program_detach(qvmhip_program *p);  // the state is being destroyed
int launch_scale(qvmhip_state *s, double factor);
int reduce_probability(qvmhip_state *s, int mode, uint32_t q, double *out_dev);
int reduce_norm2(qvmhip_state *s, double *out_dev);
int launch_force_measure(qvmhip_state *s, uint32_t q, int bit, double inv_norm, int compiled_form);
int sample_shots(qvmhip_state *s, const double *u, uint64_t nshots, uint64_t *out, int rule_ge, double eps = 0.0,
                 uint8_t *near_flags = nullptr, uint64_t *n_near = nullptr);
// Sampling on any device statevector: chunk sums and their exclusive prefix
// (nchunks = ceil(len / 256) doubles each, scan_tmp ceil(nchunks / 1024)),
// then one search per draw (all pointers on the device).
int launch_chunk_prefix(hipStream_t st, int dtype, const void *amps, uint64_t len, double *chunk_sum,
                        double *scan_tmp, double *chunk_pre);
int launch_sample_search(hipStream_t st, int dtype, const void *amps, uint64_t len, const double *chunk_pre,
                         const double *u, uint64_t nshots, int rule_ge, uint64_t *out, double eps,
                         uint8_t *near_flags, unsigned long long *near_count);
int collapse_to_basis(qvmhip_state *s, const uint64_t *index_dev);
// ops_on_device = 0: public ops on the host (scheduled here); 1: an already
// scheduled program on the device.
// cut: an already scheduled program holds CUT_PHASE ops (for host ops the
// scheduler's output decides).
int launch_tile_pass(qvmhip_state *s, uint32_t m, const uint32_t *tile_qubits, uint32_t nops,
                     const qvmhip_tile_op *ops_dev_or_host, int ops_on_device, bool cut = false);
bool tile_ops_have_cut(const qvmhip_tile_op *ops, size_t n);
int launch_fused_dense(qvmhip_state *s, const void *mat_dev, uint32_t k, const uint32_t *qubits);
int expectation_zz(qvmhip_state *s, const uint32_t *edges, uint32_t nedges, double *out);
int inner_product(qvmhip_state *a, qvmhip_state *b, double *re, double *im);
int max_abs_diff(qvmhip_state *a, qvmhip_state *b, double *out);
int tv_distance(qvmhip_state *a, qvmhip_state *b, double *out);
int probabilities(qvmhip_state *s, uint64_t start, uint64_t count, double *host_out);
int accumulate(qvmhip_state *dst, const qvmhip_state *src);
int density_diagonal(qvmhip_state *s, uint32_t nq, double *host_out);
// export modes of export_range: 16-byte (re, im) doubles or 8-byte probabilities
enum { kExportAmps = 0, kExportAmpsBE = 1, kExportProbs = 2, kExportProbsBE = 3 };
int export_range(qvmhip_state *s, uint64_t start, uint64_t count, int mode, void *host_out);
int convert_f32_to_f64(qvmhip_state *s, uint64_t start, uint64_t count, double *dev_out);
int convert_f64_to_f32(qvmhip_state *s, uint64_t start, uint64_t count, const double *dev_in);
// Capturable (kernel launches only) forms used by batches and programs.
int expectation_zz_to_device(qvmhip_state *s, const uint32_t *edges_dev, uint32_t nedges, double *out_dev);
// Device tables of a batch's finishing pass (all instances in one set of
// launches): per instance its amplitude offset and length, the prefix over
// instances of sampling chunks (count + 1 entries) and of work blocks
// (count + 1), its edge list (offset into `edges`, count); per work block
// its instance.  zz_part / exp_out are null without MaxCut edges, the
// chunk arrays / u / shot_out without shots.
struct BatchFinishArgs {
    uint32_t count;
    uint64_t blocks;
    const uint64_t *inst_off, *inst_len, *chunk_base, *blk_base;
    const uint32_t *blk_inst, *edges, *edge_off, *nedges;
    // per instance: its edges are stored low endpoint first and grouped as
    // [both endpoints below the chunk bits | one below, one above | both
    // above]; ecount[3 i + 0..2] counts the groups
    const uint32_t *ecount;
    double *zz_part, *exp_out, *chunk_sum, *chunk_pre;
    const double *u;
    uint64_t shots;
    uint64_t *shot_out;
};
uint64_t batch_finish_blocks(uint64_t len);  // work blocks of one instance
int launch_batch_finish(hipStream_t st, int dtype, const void *buf, const BatchFinishArgs &a);
int sample_to_device(qvmhip_state *s, const double *u_dev, uint64_t nshots, uint64_t *out_dev, int rule_ge);
// Tile passes: validation of a public op list (tile-local bit positions) and
// its translation into the slot-space program the tile kernels run.
int check_tile_ops(int dtype, uint32_t m, uint32_t nops, const qvmhip_tile_op *ops);
// GDIAG outside qubits: in range, not tile qubits, distinct
int check_tile_globals(uint32_t n, uint32_t m, const uint32_t *tq, uint32_t nops, const qvmhip_tile_op *ops);
void schedule_tile_ops(int dtype, bool fma, uint32_t m, const qvmhip_tile_op *in, uint32_t nin,
                       std::vector<qvmhip_tile_op> &out);
size_t tile_entry_bytes();
// Fills one batch tile entry; returns its number of work blocks of
// tiles_per_block tiles.  zero_mask: state qubits untouched before this pass
// of a run that starts from |0...0> (blocks with one of them set are zero).
uint64_t build_tile_entry(void *entry, int dtype, bool fma, uint32_t n, uint32_t m, const uint32_t *tq,
                          uint64_t amp_offset, uint64_t block_start, uint64_t op_off, uint32_t nops,
                          uint32_t tiles_per_block, bool from_zero, uint64_t zero_mask);
// op_buf: the device buffer the entries' op offsets point into
// cut: some entry of the launch holds CUT_PHASE ops
// zero: some entry starts from |0...0> or has a nonzero zero_mask
int launch_tile_batch(hipStream_t st, int dtype, bool fma, bool cut, bool zero, uint32_t m, void *base,
                      const void *entries_dev, const uint32_t *blk_ent_dev, uint64_t total_blocks,
                      uint32_t tiles_per_block, const void *op_buf);
}  // namespace qvmhip

#define QH_CHECK(expr)                                                                    \
    do {                                                                                  \
        hipError_t _e = (expr);                                                           \
        if (_e != hipSuccess)                                                             \
            return ::qvmhip::fail(QVMHIP_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

#define QH_LAUNCH_CHECK()                                                                 \
    do {                                                                                  \
        hipError_t _e = hipGetLastError();                                                \
        if (_e != hipSuccess)                                                             \
            return ::qvmhip::fail(QVMHIP_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(_e)); \
    } while (0)
