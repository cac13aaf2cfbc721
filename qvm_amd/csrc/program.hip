// program.hip -- compiled programs: a recorded list of gates, tile passes and
// fused blocks whose device operands are uploaded once, replayed eagerly or
// as one captured hipGraph.  This is the engine side of
// compile-loaded-program / compile-instruction (src/qvm.lisp:166-175,
// src/compile-gate.lisp:461-526): the host compiles once, every run replays.
#include <algorithm>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

using namespace qvmhip;

namespace {

enum ItemKind { kGate = 1, kFused = 2, kTile = 3 };

struct Item {
    int kind = 0;
    uint32_t k = 0;             // gate / fused width, tile m
    uint32_t qubits[16] = {0};  // gate qubits (nat-tuple order) or tile qubits
    size_t host_off = 0;        // operand bytes in blob
    size_t bytes = 0;
    uint32_t nops = 0;
    bool cut = false;           // tile ops hold CUT_PHASE ops
    std::vector<double> U;      // small gate matrices (launched by value)
};

}  // namespace

struct qvmhip_program {
    qvmhip_state *s = nullptr;
    std::vector<Item> items;
    std::vector<char> blob;
    void *dev = nullptr;
    bool finalized = false;
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

size_t push_blob(qvmhip_program *p, const void *data, size_t bytes) {
    const size_t off = (p->blob.size() + 255) & ~size_t(255);
    p->blob.resize(off + bytes);
    std::memcpy(p->blob.data() + off, data, bytes);
    return off;
}

int replay(qvmhip_program *p) {
    qvmhip_state *s = p->s;
    for (const Item &it : p->items) {
        int rc = QVMHIP_OK;
        const char *dev = (const char *)p->dev + it.host_off;
        switch (it.kind) {
            case kGate:
                rc = launch_gate(s, it.U.data(), it.k, it.qubits, 0);
                break;
            case kFused:
                rc = launch_fused_dense(s, dev, it.k, it.qubits);
                break;
            case kTile:
                rc = launch_tile_pass(s, it.k, it.qubits, it.nops, (const qvmhip_tile_op *)dev, 1, it.cut);
                break;
        }
        if (rc) return rc;
    }
    return QVMHIP_OK;
}

int dense_width_of(const double *U, uint32_t k) {
    // width of the dense part after control detection would need the full
    // split; small gates (k <= 3) always launch by value, wider ones go
    // through the fused kernel with the matrix in the program arena.
    (void)U;
    return (int)k;
}

}  // namespace

constexpr const char *kStateGone = "the program's state was destroyed";

void qvmhip::program_detach(qvmhip_program *p) { p->s = nullptr; }

extern "C" {

int qvmhip_program_create(qvmhip_state *s, qvmhip_program **out) {
    if (!s || !out) return fail(QVMHIP_EINVAL, "null argument");
    *out = new qvmhip_program();
    (*out)->s = s;
    s->programs.push_back(*out);
    return QVMHIP_OK;
}

int qvmhip_program_add_matrix(qvmhip_program *p, const double *U, uint32_t k, const uint32_t *qubits) {
    if (!p || !U || !qubits) return fail(QVMHIP_EINVAL, "null argument");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    if (p->finalized) return fail(QVMHIP_EINVAL, "program already finalized");
    if (k < 1 || k > 6 || k > p->s->n) return fail(QVMHIP_EINVAL, "gate arity must be 1..6");
    for (uint32_t i = 0; i < k; ++i) {
        if (qubits[i] >= p->s->n) return fail(QVMHIP_EINVAL, "qubit out of range");
        for (uint32_t j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return fail(QVMHIP_EINVAL, "repeated qubit in gate");
    }
    if (dense_width_of(U, k) > 3) return qvmhip_program_add_fused(p, U, k, qubits);
    Item it;
    it.kind = kGate;
    it.k = k;
    std::copy(qubits, qubits + k, it.qubits);
    it.U.assign(U, U + (size_t(2) << (2 * k)));
    p->items.push_back(std::move(it));
    return QVMHIP_OK;
}

int qvmhip_program_add_fused(qvmhip_program *p, const double *U, uint32_t k, const uint32_t *qubits) {
    if (!p || !U || !qubits) return fail(QVMHIP_EINVAL, "null argument");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    if (p->finalized) return fail(QVMHIP_EINVAL, "program already finalized");
    if (k < 3 || k > 6 || k > p->s->n) return fail(QVMHIP_EINVAL, "fused blocks are 3..6 qubits");
    for (uint32_t i = 0; i < k; ++i) {
        if (qubits[i] >= p->s->n) return fail(QVMHIP_EINVAL, "qubit out of range");
        for (uint32_t j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return fail(QVMHIP_EINVAL, "repeated qubit in gate");
    }
    Item it;
    it.kind = kFused;
    it.k = k;
    std::copy(qubits, qubits + k, it.qubits);
    it.bytes = sizeof(double) * (size_t(2) << (2 * k));
    it.host_off = push_blob(p, U, it.bytes);
    p->items.push_back(std::move(it));
    return QVMHIP_OK;
}

int qvmhip_program_add_tile_pass(qvmhip_program *p, uint32_t m, const uint32_t *tile_qubits, uint32_t nops,
                                 const qvmhip_tile_op *ops) {
    if (!p || !tile_qubits || (nops && !ops)) return fail(QVMHIP_EINVAL, "null argument");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    if (p->finalized) return fail(QVMHIP_EINVAL, "program already finalized");
    if (m > p->s->n || m > 16) return fail(QVMHIP_EINVAL, "tile wider than the state");
    for (uint32_t i = 0; i < m; ++i) {
        if (tile_qubits[i] >= p->s->n) return fail(QVMHIP_EINVAL, "tile qubit out of range");
        if (i && tile_qubits[i] <= tile_qubits[i - 1]) return fail(QVMHIP_EINVAL, "tile qubits must ascend");
    }
    int rc = check_tile_ops(p->s->dtype, m, nops, ops);
    if (!rc) rc = check_tile_globals(p->s->n, m, tile_qubits, nops, ops);
    if (rc) return rc;
    if (nops == 0) return QVMHIP_OK;
    Item it;
    it.kind = kTile;
    it.k = m;
    std::copy(tile_qubits, tile_qubits + m, it.qubits);
    std::vector<qvmhip_tile_op> sched;
    schedule_tile_ops(p->s->dtype, (p->s->flags & QVMHIP_FLAG_FMA_TILES) != 0, m, ops, nops, sched);
    it.nops = (uint32_t)sched.size();
    it.cut = tile_ops_have_cut(sched.data(), sched.size());
    it.bytes = sizeof(qvmhip_tile_op) * sched.size();
    it.host_off = push_blob(p, sched.data(), it.bytes);
    p->items.push_back(std::move(it));
    return QVMHIP_OK;
}

int qvmhip_program_finalize(qvmhip_program *p, int use_graph) {
    if (!p) return fail(QVMHIP_EINVAL, "null program");
    if (p->finalized) return fail(QVMHIP_EINVAL, "program already finalized");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    qvmhip_state *s = p->s;
    DeviceScope ds(s->device);
    if (!p->blob.empty()) {
        QH_CHECK(hipMalloc(&p->dev, p->blob.size()));
        QH_CHECK(hipMemcpyAsync(p->dev, p->blob.data(), p->blob.size(), hipMemcpyHostToDevice, s->stream));
        QH_CHECK(hipStreamSynchronize(s->stream));
    }
    p->use_graph = use_graph != 0;
    if (p->use_graph) {
        QH_CHECK(hipStreamSynchronize(s->stream));
        const bool was_profiling = s->profiling;
        s->profiling = false;  // no event records inside the captured graph
        QH_CHECK(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = replay(p);
        s->profiling = was_profiling;
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return fail(QVMHIP_EDEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        p->graph = g;
        QH_CHECK(hipGraphInstantiate(&p->exec, p->graph, nullptr, nullptr, 0));
    }
    p->finalized = true;
    return QVMHIP_OK;
}

int qvmhip_program_run(qvmhip_program *p) {
    if (!p) return fail(QVMHIP_EINVAL, "null program");
    if (!p->finalized) return fail(QVMHIP_EINVAL, "program not finalized");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    DeviceScope ds(p->s->device);
    if (p->use_graph) {
        QH_CHECK(hipGraphLaunch(p->exec, p->s->stream));
        return QVMHIP_OK;
    }
    return replay(p);
}

int qvmhip_program_destroy(qvmhip_program *p) {
    if (!p) return QVMHIP_OK;
    if (p->s) {
        DeviceScope ds(p->s->device);
        (void)hipStreamSynchronize(p->s->stream);
        auto &v = p->s->programs;
        v.erase(std::remove(v.begin(), v.end(), p), v.end());
    }
    if (p->exec) (void)hipGraphExecDestroy(p->exec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    if (p->dev) (void)hipFree(p->dev);
    delete p;
    return QVMHIP_OK;
}

int qvmhip_time_program(qvmhip_program *p, int reps, double *ms) {
    if (!p || !ms || reps < 1) return fail(QVMHIP_EINVAL, "bad argument");
    if (!p->s) return fail(QVMHIP_EINVAL, kStateGone);
    qvmhip_state *s = p->s;
    DeviceScope ds(s->device);
    struct Events {  // destroyed on every return
        hipEvent_t a = nullptr, b = nullptr;
        ~Events() {
            if (a) (void)hipEventDestroy(a);
            if (b) (void)hipEventDestroy(b);
        }
    } ev;
    QH_CHECK(hipEventCreate(&ev.a));
    QH_CHECK(hipEventCreate(&ev.b));
    QH_CHECK(hipEventRecord(ev.a, s->stream));
    for (int r = 0; r < reps; ++r) {
        int rc = qvmhip_program_run(p);
        if (rc) {
            (void)hipStreamSynchronize(s->stream);  // no pending record of the events
            return rc;
        }
    }
    QH_CHECK(hipEventRecord(ev.b, s->stream));
    QH_CHECK(hipEventSynchronize(ev.b));
    float t = 0;
    QH_CHECK(hipEventElapsedTime(&t, ev.a, ev.b));
    *ms = (double)t / reps;
    return QVMHIP_OK;
}

}  // extern "C"
