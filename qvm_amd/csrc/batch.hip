// batch.hip -- many small states run as one captured graph.
//
// The QAOA-style workload is hundreds of independent circuits on 10..24
// qubits.  Launching each instance's kernels separately leaves the GPU idle
// (a 2^10-amplitude tile pass is one workgroup), so a batch keeps every state
// in one HBM buffer and runs level-synchronously: level L is the L-th planned
// item of every instance, and all tile passes of a level with the same tile
// width go out as ONE batched kernel (k_tile_batch) whose workgroups are
// distributed over the instances by a prefix of their tile counts.  After the
// last level, each instance's cost expectation and samples are computed.  The
// whole sequence -- reset to |0...0>, levels, expectations, sampling -- is
// captured once into a hipGraph and replayed with one graph launch.
#include <algorithm>
#include <cstring>
#include <map>
#include <vector>

#include "qvmhip_internal.h"

using namespace qvmhip;

namespace {

enum { kItemTile = 1, kItemGate = 2 };

struct BItem {
    int kind = 0;
    uint32_t k = 0;            // tile m or gate width
    uint32_t qubits[16] = {0};
    size_t op_off = 0;         // tile ops in blob
    uint32_t nops = 0;
    bool cut = false;          // tile ops hold CUT_PHASE ops
    std::vector<double> U;     // gate matrix (k <= 3)
};

// Tiles per work block of the batched tile kernel: enough to amortise the
// entry lookup, few enough to spread small states over many workgroups.
constexpr uint32_t kTilesPerBlock = 4;

struct BLaunch {
    int kind;
    uint32_t m;         // tile width
    size_t ent_off;     // bytes into the entry buffer
    uint32_t nent;
    uint64_t blocks;    // work blocks
    size_t map_off;     // first element of this launch's block -> entry map
    uint32_t inst;      // gate launches
    const BItem *item;  // gate launches
    bool cut;           // tile launches: an entry holds CUT_PHASE ops
    bool zero;          // tile launches: an entry starts from |0...0> or has all-zero blocks
};

template <typename T>
__global__ void k_batch_reset(vec2_t<T> *__restrict__ buf, uint64_t total) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += stride) {
        vec2_t<T> z;
        z.x = T(0);
        z.y = T(0);
        buf[i] = z;
    }
}

template <typename T>
__global__ void k_batch_ones(vec2_t<T> *__restrict__ buf, const uint64_t *__restrict__ starts, uint32_t count) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) {
        vec2_t<T> one;
        one.x = T(1);
        one.y = T(0);
        buf[starts[i]] = one;
    }
}

}  // namespace

struct qvmhip_batch {
    uint32_t count = 0;
    int dtype = QVMHIP_F64;
    uint32_t flags = 0;  // QVMHIP_FLAG_* (FMA-contracted tile passes)
    size_t amp_bytes = 16;
    std::vector<uint32_t> nq;
    std::vector<uint64_t> offset;  // amplitude offset of each state
    uint64_t total = 0;
    void *buf = nullptr;
    hipStream_t stream = nullptr;
    // scratch shared by all views (work on one stream is serialised)
    double *red = nullptr, *chunk_sum = nullptr, *chunk_pre = nullptr, *scan_tmp = nullptr;
    void *mat = nullptr;
    std::vector<qvmhip_state> views;
    // programs
    std::vector<std::vector<BItem>> prog;
    std::vector<char> blob;
    void *blob_dev = nullptr;
    std::vector<std::vector<uint32_t>> edges;
    uint32_t *edges_dev = nullptr;
    std::vector<size_t> edge_off;
    double *exp_dev = nullptr;
    uint64_t shots = 0;
    std::vector<double> u_host;
    double *u_dev = nullptr;
    uint64_t *out_dev = nullptr;
    uint64_t *starts_dev = nullptr;
    // finishing pass (<H_C> and shots of every instance in four launches)
    BatchFinishArgs fin{};
    bool first_pass_resets = false;  // every program starts with a tile pass (no reset sweep)
    std::vector<void *> fin_dev;  // its device tables, freed with the batch
    std::vector<BLaunch> launches;
    std::vector<char> entries;
    void *entries_dev = nullptr;
    std::vector<uint32_t> blkmap;  // per tile launch: work block -> entry (launch-local index)
    uint32_t *blkmap_dev = nullptr;
    bool finalized = false, use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;  // the caller's current device at creation
};

namespace {

void free_batch(qvmhip_batch *b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (void *p : {(void *)b->buf, (void *)b->red, (void *)b->chunk_sum, (void *)b->chunk_pre, (void *)b->scan_tmp,
                    b->mat, b->blob_dev, (void *)b->edges_dev, (void *)b->exp_dev, (void *)b->u_dev,
                    (void *)b->out_dev, (void *)b->starts_dev, b->entries_dev, (void *)b->blkmap_dev})
        if (p) (void)hipFree(p);
    for (void *p : b->fin_dev)
        if (p) (void)hipFree(p);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int check_inst(const qvmhip_batch *b, uint32_t inst) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    if (b->finalized) return fail(QVMHIP_EINVAL, "batch already finalized");
    if (inst >= b->count) return fail(QVMHIP_EINVAL, "instance index out of range");
    return QVMHIP_OK;
}

size_t push(std::vector<char> &blob, const void *p, size_t bytes) {
    const size_t off = (blob.size() + 255) & ~size_t(255);
    blob.resize(off + bytes);
    std::memcpy(blob.data() + off, p, bytes);
    return off;
}

// Enqueue the whole batch on its stream (kernel launches only, capturable).
int enqueue(qvmhip_batch *b) {
    const uint32_t rg = (uint32_t)std::min<uint64_t>((b->total + 255) / 256, 256 * 64);
    if (b->first_pass_resets) {
        // every instance starts with a tile pass that builds |0...0> itself
    } else if (b->dtype == QVMHIP_F64) {
        hipLaunchKernelGGL(k_batch_reset<double>, dim3(rg), dim3(256), 0, b->stream, (double2 *)b->buf, b->total);
        hipLaunchKernelGGL(k_batch_ones<double>, dim3((b->count + 255) / 256), dim3(256), 0, b->stream,
                           (double2 *)b->buf, b->starts_dev, b->count);
    } else {
        hipLaunchKernelGGL(k_batch_reset<float>, dim3(rg), dim3(256), 0, b->stream, (float2 *)b->buf, b->total);
        hipLaunchKernelGGL(k_batch_ones<float>, dim3((b->count + 255) / 256), dim3(256), 0, b->stream,
                           (float2 *)b->buf, b->starts_dev, b->count);
    }
    QH_LAUNCH_CHECK();
    for (const BLaunch &L : b->launches) {
        int rc;
        if (L.kind == kItemTile)
            rc = launch_tile_batch(b->stream, b->dtype, (b->flags & QVMHIP_FLAG_FMA_TILES) != 0, L.cut, L.zero, L.m,
                                   b->buf, (const char *)b->entries_dev + L.ent_off, b->blkmap_dev + L.map_off,
                                   L.blocks, kTilesPerBlock, b->blob_dev);
        else
            rc = launch_gate(&b->views[L.inst], L.item->U.data(), L.item->k, L.item->qubits, 0);
        if (rc) return rc;
    }
    if (b->fin.zz_part || b->fin.shots) return launch_batch_finish(b->stream, b->dtype, b->buf, b->fin);
    return QVMHIP_OK;
}

// Device tables of the finishing pass; called once by finalize.
int build_finish(qvmhip_batch *b) {
    BatchFinishArgs &f = b->fin;
    f.count = b->count;
    std::vector<uint64_t> len(b->count), chunk_base(b->count + 1, 0), blk_base(b->count + 1, 0);
    std::vector<uint32_t> blk_inst, edge_off(b->count), nedges(b->count), ecount(3 * (size_t)b->count, 0);
    std::vector<uint32_t> edges;  // per instance: low endpoint first, grouped below / across / above the chunk bits
    bool any_edges = false;
    for (uint32_t i = 0; i < b->count; ++i) {
        len[i] = uint64_t(1) << b->nq[i];
        chunk_base[i + 1] = chunk_base[i] + std::max<uint64_t>(1, len[i] >> kChunkLog2);
        const uint64_t nb = batch_finish_blocks(len[i]);
        blk_base[i + 1] = blk_base[i] + nb;
        blk_inst.insert(blk_inst.end(), (size_t)nb, i);
        edge_off[i] = (uint32_t)edges.size();
        nedges[i] = (uint32_t)(b->edges[i].size() / 2);
        any_edges = any_edges || nedges[i] != 0;
        const uint32_t lowb = (uint32_t)kChunkLog2;
        for (int group = 0; group < 3; ++group) {
            for (uint32_t e = 0; e < nedges[i]; ++e) {
                uint32_t x = b->edges[i][2 * e], y = b->edges[i][2 * e + 1];
                if (x > y) std::swap(x, y);
                const int g = (y < lowb) ? 0 : (x < lowb ? 1 : 2);
                if (g != group) continue;
                edges.push_back(x);
                edges.push_back(y);
                ++ecount[3 * (size_t)i + g];
            }
        }
    }
    f.blocks = blk_base[b->count];
    if (!any_edges && !b->shots) return QVMHIP_OK;
    auto up = [&](const void *host, size_t bytes, void **dev) -> int {
        QH_CHECK(hipMalloc(dev, std::max<size_t>(bytes, 8)));
        b->fin_dev.push_back(*dev);
        if (host) QH_CHECK(hipMemcpy(*dev, host, bytes, hipMemcpyHostToDevice));
        return QVMHIP_OK;
    };
    void *d;
    int rc;
    if ((rc = up(b->offset.data(), 8 * b->count, &d))) return rc;
    f.inst_off = (const uint64_t *)d;
    if ((rc = up(len.data(), 8 * b->count, &d))) return rc;
    f.inst_len = (const uint64_t *)d;
    if ((rc = up(chunk_base.data(), 8 * chunk_base.size(), &d))) return rc;
    f.chunk_base = (const uint64_t *)d;
    if ((rc = up(blk_base.data(), 8 * blk_base.size(), &d))) return rc;
    f.blk_base = (const uint64_t *)d;
    if ((rc = up(blk_inst.data(), 4 * blk_inst.size(), &d))) return rc;
    f.blk_inst = (const uint32_t *)d;
    if ((rc = up(edge_off.data(), 4 * edge_off.size(), &d))) return rc;
    f.edge_off = (const uint32_t *)d;
    if ((rc = up(nedges.data(), 4 * nedges.size(), &d))) return rc;
    f.nedges = (const uint32_t *)d;
    if ((rc = up(edges.data(), 4 * edges.size(), &d))) return rc;
    f.edges = (const uint32_t *)d;
    if ((rc = up(ecount.data(), 4 * ecount.size(), &d))) return rc;
    f.ecount = (const uint32_t *)d;
    if (any_edges) {
        if ((rc = up(nullptr, 8 * f.blocks, &d))) return rc;
        f.zz_part = (double *)d;
        f.exp_out = b->exp_dev;
    }
    if (b->shots) {
        if ((rc = up(nullptr, 8 * chunk_base[b->count], &d))) return rc;
        f.chunk_sum = (double *)d;
        if ((rc = up(nullptr, 8 * chunk_base[b->count], &d))) return rc;
        f.chunk_pre = (double *)d;
        f.u = b->u_dev;
        f.shots = b->shots;
        f.shot_out = b->out_dev;
    }
    return QVMHIP_OK;
}

}  // namespace

extern "C" {

int qvmhip_batch_create(uint32_t count, const uint32_t *nqubits, int dtype, qvmhip_batch **out) {
    if (!out || !nqubits || count == 0) return fail(QVMHIP_EINVAL, "bad argument");
    if (dtype != QVMHIP_F64 && dtype != QVMHIP_F32) return fail(QVMHIP_EINVAL, "dtype must be F64 or F32");
    *out = nullptr;
    int dev = 0;
    QH_CHECK(hipGetDevice(&dev));
    qvmhip_batch *b = new qvmhip_batch();
    b->device = dev;
    b->count = count;
    b->dtype = dtype;
    b->amp_bytes = dtype == QVMHIP_F64 ? 16 : 8;
    uint32_t nmax = 0;
    for (uint32_t i = 0; i < count; ++i) {
        if (nqubits[i] < 1 || nqubits[i] > 30) {
            free_batch(b);
            return fail(QVMHIP_EINVAL, "batch states have 1..30 qubits");
        }
        b->nq.push_back(nqubits[i]);
        b->offset.push_back(b->total);
        b->total += uint64_t(1) << nqubits[i];
        nmax = std::max(nmax, nqubits[i]);
    }
    const uint64_t nch = std::max<uint64_t>(1, (uint64_t(1) << nmax) >> kChunkLog2);
    const uint64_t nseg = (nch + 1023) / 1024;
    hipError_t e = hipSuccess;
    auto bail = [&](const char *what) {
        free_batch(b);
        return fail(e == hipErrorOutOfMemory ? QVMHIP_ENOMEM : QVMHIP_EDEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
    };
    if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) return bail("stream");
    if ((e = hipMalloc(&b->buf, b->total * b->amp_bytes)) != hipSuccess) return bail("hipMalloc(states)");
    if ((e = hipMalloc((void **)&b->red, sizeof(double) * (kRedBlocks * 2 + 16))) != hipSuccess) return bail("red");
    if ((e = hipMalloc((void **)&b->chunk_sum, sizeof(double) * nch)) != hipSuccess) return bail("chunks");
    if ((e = hipMalloc((void **)&b->chunk_pre, sizeof(double) * nch)) != hipSuccess) return bail("prefix");
    if ((e = hipMalloc((void **)&b->scan_tmp, sizeof(double) * nseg)) != hipSuccess) return bail("scan");
    if ((e = hipMalloc(&b->mat, 128 * 1024)) != hipSuccess) return bail("matrix slot");
    if ((e = hipMalloc((void **)&b->exp_dev, sizeof(double) * count)) != hipSuccess) return bail("expectations");
    if ((e = hipMemsetAsync(b->exp_dev, 0, sizeof(double) * count, b->stream)) != hipSuccess) return bail("memset");
    if ((e = hipMalloc((void **)&b->starts_dev, sizeof(uint64_t) * count)) != hipSuccess) return bail("starts");
    if ((e = hipMemcpyAsync(b->starts_dev, b->offset.data(), sizeof(uint64_t) * count, hipMemcpyHostToDevice,
                            b->stream)) != hipSuccess)
        return bail("starts upload");
    if ((e = hipStreamSynchronize(b->stream)) != hipSuccess) return bail("upload");
    b->views.resize(count);
    for (uint32_t i = 0; i < count; ++i) {
        qvmhip_state &v = b->views[i];
        v.n = b->nq[i];
        v.dtype = dtype;
        v.device = dev;
        v.len = uint64_t(1) << v.n;
        v.amp_bytes = b->amp_bytes;
        v.amps = (char *)b->buf + b->offset[i] * b->amp_bytes;
        v.stream = b->stream;
        v.own_stream = false;
        v.red = b->red;
        v.mat = b->mat;
        v.chunk_sum = b->chunk_sum;
        v.chunk_pre = b->chunk_pre;
        v.scan_tmp = b->scan_tmp;
        v.nchunks = std::max<uint64_t>(1, v.len >> kChunkLog2);
    }
    b->prog.resize(count);
    b->edges.resize(count);
    *out = b;
    return QVMHIP_OK;
}

int qvmhip_batch_set_flags(qvmhip_batch *b, uint32_t flags) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    if (flags & ~QVMHIP_FLAG_FMA_TILES) return fail(QVMHIP_EINVAL, "unknown batch flags");
    for (const auto &p : b->prog)
        if (!p.empty()) return fail(QVMHIP_EINVAL, "batch flags must be set before items are added");
    b->flags = flags;
    return QVMHIP_OK;
}

int qvmhip_batch_add_tile_pass(qvmhip_batch *b, uint32_t inst, uint32_t m, const uint32_t *tile_qubits,
                               uint32_t nops, const qvmhip_tile_op *ops) {
    int rc = check_inst(b, inst);
    if (rc) return rc;
    const uint32_t n = b->nq[inst];
    if (m < 8 || m > n || m > 13 || (m == 13 && b->dtype == QVMHIP_F64))
        return fail(QVMHIP_EINVAL, "tile size must be 8..min(n,12) (F64) or ..13 (F32)");
    if (!tile_qubits || (nops && !ops)) return fail(QVMHIP_EINVAL, "null argument");
    for (uint32_t i = 0; i < m; ++i) {
        if (tile_qubits[i] >= n) return fail(QVMHIP_EINVAL, "tile qubit out of range");
        if (i && tile_qubits[i] <= tile_qubits[i - 1]) return fail(QVMHIP_EINVAL, "tile qubits must ascend");
    }
    rc = check_tile_ops(b->dtype, m, nops, ops);
    if (!rc) rc = check_tile_globals(n, m, tile_qubits, nops, ops);
    if (rc) return rc;
    if (!nops) return QVMHIP_OK;
    BItem it;
    it.kind = kItemTile;
    it.k = m;
    std::copy(tile_qubits, tile_qubits + m, it.qubits);
    std::vector<qvmhip_tile_op> sched;
    schedule_tile_ops(b->dtype, (b->flags & QVMHIP_FLAG_FMA_TILES) != 0, m, ops, nops, sched);
    it.nops = (uint32_t)sched.size();
    it.cut = tile_ops_have_cut(sched.data(), sched.size());
    it.op_off = push(b->blob, sched.data(), sizeof(qvmhip_tile_op) * sched.size());
    b->prog[inst].push_back(std::move(it));
    return QVMHIP_OK;
}

int qvmhip_batch_add_matrix(qvmhip_batch *b, uint32_t inst, const double *U, uint32_t k, const uint32_t *qubits) {
    int rc = check_inst(b, inst);
    if (rc) return rc;
    if (!U || !qubits || k < 1 || k > 3 || k > b->nq[inst])
        return fail(QVMHIP_EINVAL, "batch gates are 1..3 qubits (wider gates go into tile passes or programs)");
    for (uint32_t i = 0; i < k; ++i) {
        if (qubits[i] >= b->nq[inst]) return fail(QVMHIP_EINVAL, "qubit out of range");
        for (uint32_t j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return fail(QVMHIP_EINVAL, "repeated qubit in gate");
    }
    BItem it;
    it.kind = kItemGate;
    it.k = k;
    std::copy(qubits, qubits + k, it.qubits);
    it.U.assign(U, U + (size_t(2) << (2 * k)));
    b->prog[inst].push_back(std::move(it));
    return QVMHIP_OK;
}

int qvmhip_batch_set_zz(qvmhip_batch *b, uint32_t inst, const uint32_t *edges, uint32_t nedges) {
    int rc = check_inst(b, inst);
    if (rc) return rc;
    if (nedges > 256 || (nedges && !edges)) return fail(QVMHIP_EINVAL, "at most 256 edges");
    for (uint32_t i = 0; i < 2 * nedges; ++i)
        if (edges[i] >= b->nq[inst]) return fail(QVMHIP_EINVAL, "edge qubit out of range");
    b->edges[inst].assign(edges, edges + 2 * nedges);
    return QVMHIP_OK;
}

int qvmhip_batch_set_shots(qvmhip_batch *b, uint64_t shots, const double *u) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    if (b->finalized) return fail(QVMHIP_EINVAL, "batch already finalized");
    if (shots && !u) return fail(QVMHIP_EINVAL, "null uniforms");
    for (uint64_t i = 0; i < shots * b->count; ++i)
        if (!(u[i] >= 0.0 && u[i] <= 1.0)) return fail(QVMHIP_ERANGE, "uniform draw outside [0,1]");
    b->shots = shots;
    b->u_host.assign(u, u + shots * b->count);
    return QVMHIP_OK;
}

int qvmhip_batch_finalize(qvmhip_batch *b, int use_graph) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    DeviceScope ds(b->device);
    if (b->finalized) return fail(QVMHIP_EINVAL, "batch already finalized");
    hipError_t e;
    if (!b->blob.empty()) {
        if ((e = hipMalloc(&b->blob_dev, b->blob.size())) != hipSuccess) return fail(QVMHIP_ENOMEM, "blob");
        QH_CHECK(hipMemcpy(b->blob_dev, b->blob.data(), b->blob.size(), hipMemcpyHostToDevice));
    }
    std::vector<uint32_t> eh;
    b->edge_off.assign(b->count, 0);
    for (uint32_t i = 0; i < b->count; ++i) {
        b->edge_off[i] = eh.size() * sizeof(uint32_t);
        eh.insert(eh.end(), b->edges[i].begin(), b->edges[i].end());
    }
    if (!eh.empty()) {
        QH_CHECK(hipMalloc((void **)&b->edges_dev, eh.size() * sizeof(uint32_t)));
        QH_CHECK(hipMemcpy(b->edges_dev, eh.data(), eh.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    if (b->shots) {
        QH_CHECK(hipMalloc((void **)&b->u_dev, b->u_host.size() * sizeof(double)));
        QH_CHECK(hipMemcpy(b->u_dev, b->u_host.data(), b->u_host.size() * sizeof(double), hipMemcpyHostToDevice));
        QH_CHECK(hipMalloc((void **)&b->out_dev, b->u_host.size() * sizeof(uint64_t)));
    }
    {
        int rc = build_finish(b);
        if (rc) return rc;
    }
    // when every instance's first item is a tile pass, that pass builds the
    // initial |0...0> in registers and the reset sweep of the buffer goes
    b->first_pass_resets = true;
    for (uint32_t i = 0; i < b->count; ++i)
        if (b->prog[i].empty() || b->prog[i][0].kind != kItemTile) b->first_pass_resets = false;
    // qubits touched before each item of a run (which starts from |0...0>):
    // amplitudes with a bit set outside them are still zero
    std::vector<std::vector<uint64_t>> touched(b->count);
    for (uint32_t i = 0; i < b->count; ++i) {
        uint64_t S = 0;
        for (const BItem &it : b->prog[i]) {
            touched[i].push_back(S);
            for (uint32_t j = 0; j < it.k; ++j) S |= uint64_t(1) << it.qubits[j];
        }
    }
    // levels: the L-th item of every instance; tile passes grouped by width
    size_t depth = 0;
    for (auto &p : b->prog) depth = std::max(depth, p.size());
    const size_t esz = tile_entry_bytes();
    std::vector<size_t> ent_host_off;  // per launch: offset into b->entries
    for (size_t L = 0; L < depth; ++L) {
        std::map<uint32_t, std::vector<uint32_t>> by_m;
        for (uint32_t i = 0; i < b->count; ++i) {
            if (L >= b->prog[i].size()) continue;
            const BItem &it = b->prog[i][L];
            if (it.kind == kItemTile)
                by_m[it.k].push_back(i);
            else
                b->launches.push_back({kItemGate, 0, 0, 0, 0, 0, i, &it, false, false});
        }
        for (auto &kv : by_m) {
            BLaunch bl{kItemTile, kv.first, b->entries.size(), (uint32_t)kv.second.size(), 0, b->blkmap.size(), 0,
                       nullptr, false, false};
            uint64_t blocks = 0;
            uint32_t local = 0;
            for (uint32_t i : kv.second) {
                const BItem &it = b->prog[i][L];
                bl.cut = bl.cut || it.cut;
                std::vector<char> ent(esz);
                uint64_t tmask = 0;
                for (uint32_t j = 0; j < it.k; ++j) tmask |= uint64_t(1) << it.qubits[j];
                const uint64_t all = (uint64_t(1) << b->nq[i]) - 1;
                const bool from_zero = L == 0 && b->first_pass_resets;
                const uint64_t zero_mask = all & ~touched[i][L] & ~tmask;
                bl.zero = bl.zero || from_zero || zero_mask != 0;
                const uint64_t nb = build_tile_entry(
                    ent.data(), b->dtype, (b->flags & QVMHIP_FLAG_FMA_TILES) != 0, b->nq[i], it.k, it.qubits,
                    b->offset[i], blocks, it.op_off, it.nops, kTilesPerBlock, from_zero, zero_mask);
                b->entries.insert(b->entries.end(), ent.begin(), ent.end());
                b->blkmap.insert(b->blkmap.end(), (size_t)nb, local);
                blocks += nb;
                ++local;
            }
            bl.blocks = blocks;
            b->launches.push_back(bl);
        }
    }
    if (!b->entries.empty()) {
        QH_CHECK(hipMalloc(&b->entries_dev, b->entries.size()));
        QH_CHECK(hipMemcpy(b->entries_dev, b->entries.data(), b->entries.size(), hipMemcpyHostToDevice));
        QH_CHECK(hipMalloc((void **)&b->blkmap_dev, b->blkmap.size() * sizeof(uint32_t)));
        QH_CHECK(hipMemcpy(b->blkmap_dev, b->blkmap.data(), b->blkmap.size() * sizeof(uint32_t),
                           hipMemcpyHostToDevice));
    }
    b->use_graph = use_graph != 0;
    if (b->use_graph) {
        QH_CHECK(hipStreamSynchronize(b->stream));
        QH_CHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        int rc = enqueue(b);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(b->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(QVMHIP_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        b->graph = g;
        QH_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    b->finalized = true;
    return QVMHIP_OK;
}

int qvmhip_batch_run(qvmhip_batch *b) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    DeviceScope ds(b->device);
    if (!b->finalized) return fail(QVMHIP_EINVAL, "batch not finalized");
    if (b->use_graph) {
        QH_CHECK(hipGraphLaunch(b->exec, b->stream));
        return QVMHIP_OK;
    }
    return enqueue(b);
}

int qvmhip_batch_sync(qvmhip_batch *b) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    DeviceScope ds(b->device);
    QH_CHECK(hipStreamSynchronize(b->stream));
    return QVMHIP_OK;
}

int qvmhip_batch_results(qvmhip_batch *b, double *expectations, uint64_t *shots) {
    if (!b) return fail(QVMHIP_EINVAL, "null batch");
    DeviceScope ds(b->device);
    QH_CHECK(hipStreamSynchronize(b->stream));
    if (expectations) QH_CHECK(hipMemcpy(expectations, b->exp_dev, sizeof(double) * b->count, hipMemcpyDeviceToHost));
    if (shots && b->shots)
        QH_CHECK(hipMemcpy(shots, b->out_dev, sizeof(uint64_t) * b->shots * b->count, hipMemcpyDeviceToHost));
    return QVMHIP_OK;
}

int qvmhip_batch_read_amplitudes(qvmhip_batch *b, uint32_t inst, uint64_t start, uint64_t count, double *out) {
    if (!b || inst >= b->count || (count && !out)) return fail(QVMHIP_EINVAL, "bad argument");
    DeviceScope ds(b->device);
    const uint64_t len = uint64_t(1) << b->nq[inst];
    if (start > len || count > len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    QH_CHECK(hipStreamSynchronize(b->stream));
    if (b->dtype == QVMHIP_F64) {
        QH_CHECK(hipMemcpy(out, (const double2 *)b->buf + b->offset[inst] + start, count * 16, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> tmp(2 * count);
        QH_CHECK(hipMemcpy(tmp.data(), (const float2 *)b->buf + b->offset[inst] + start, count * 8,
                           hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < 2 * count; ++i) out[i] = tmp[i];
    }
    return QVMHIP_OK;
}

int qvmhip_batch_destroy(qvmhip_batch *b) {
    if (!b) return QVMHIP_OK;
    DeviceScope ds(b->device);
    (void)hipStreamSynchronize(b->stream);
    free_batch(b);
    return QVMHIP_OK;
}

}  // extern "C"
