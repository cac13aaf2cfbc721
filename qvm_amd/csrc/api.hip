// api.hip -- the C ABI of include/qvmhip.h: state lifetime, argument checks,
// measurement protocol and compiled programs.  Kernels live in gates.hip,
// reduce.hip, tile.hip and fused.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "qvmhip_internal.h"

namespace qvmhip {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int fail(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

}  // namespace qvmhip

using namespace qvmhip;

namespace {

constexpr size_t kMatSlot = 128 * 1024;
constexpr size_t kStageBytes = size_t(64) << 20;

uint64_t scratch_bytes(uint32_t n) {
    const uint64_t len = uint64_t(1) << n;
    const uint64_t nch = std::max<uint64_t>(1, len >> kChunkLog2);
    const uint64_t nseg = (nch + 1023) / 1024;
    return sizeof(double) * (kRedBlocks * 2 + 16) + kMatSlot + sizeof(double) * (2 * nch + nseg);
}

int check_qubits(const qvmhip_state *s, uint32_t k, const uint32_t *qubits) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    if (k < 1 || k > (uint32_t)kMaxK) return fail(QVMHIP_EINVAL, "gate arity must be 1..6");
    if (k > s->n) return fail(QVMHIP_EINVAL, "gate wider than the state");
    if (!qubits) return fail(QVMHIP_EINVAL, "null qubit list");
    for (uint32_t i = 0; i < k; ++i) {
        if (qubits[i] >= s->n)
            return fail(QVMHIP_EINVAL, "qubit " + std::to_string(qubits[i]) + " out of range for a " +
                                           std::to_string(s->n) + "-qubit state");
        for (uint32_t j = 0; j < i; ++j)
            if (qubits[i] == qubits[j]) return fail(QVMHIP_EINVAL, "repeated qubit in gate");
    }
    return QVMHIP_OK;
}

bool finite_matrix(const double *U, uint32_t k) {
    const size_t n = size_t(2) << (2 * k);
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(U[i])) return false;
    return true;
}

int scalar_to_host(qvmhip_state *s, double *dev, double *out) {
    QH_CHECK(hipMemcpyAsync(s->host, dev, sizeof(double), hipMemcpyDeviceToHost, s->stream));
    QH_CHECK(hipStreamSynchronize(s->stream));
    *out = s->host[0];
    return QVMHIP_OK;
}

void free_state(qvmhip_state *s) {
    if (!s) return;
    for (auto &r : s->prof) {
        if (r.a) (void)hipEventDestroy(r.a);
        if (r.b) (void)hipEventDestroy(r.b);
    }
    for (auto e : s->event_pool) (void)hipEventDestroy(e);
    if (s->amps) (void)hipFree(s->amps);
    if (s->red) (void)hipFree(s->red);
    if (s->mat) (void)hipFree(s->mat);
    if (s->chunk_sum) (void)hipFree(s->chunk_sum);
    if (s->chunk_pre) (void)hipFree(s->chunk_pre);
    if (s->scan_tmp) (void)hipFree(s->scan_tmp);
    if (s->shot_out) (void)hipFree(s->shot_out);
    if (s->shot_u) (void)hipFree(s->shot_u);
    if (s->shot_near) (void)hipFree(s->shot_near);
    if (s->host) (void)hipHostFree(s->host);
    if (s->stage) (void)hipHostFree(s->stage);
    if (s->stream && s->own_stream) (void)hipStreamDestroy(s->stream);
    delete s;
}

}  // namespace

extern "C" {

const char *qvmhip_last_error(void) { return g_last_error.c_str(); }

const char *qvmhip_version(void) { return "qvmhip 0.1.0 (gfx950)"; }

int qvmhip_device_count(int *out) {
    if (!out) return fail(QVMHIP_EINVAL, "null out");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    *out = n;
    return QVMHIP_OK;
}

int qvmhip_state_bytes(uint32_t n, int dtype, uint64_t *out) {
    if (!out) return fail(QVMHIP_EINVAL, "null out");
    if (n < 1 || n > 40) return fail(QVMHIP_EINVAL, "qubit count must be 1..40");
    if (dtype != QVMHIP_F64 && dtype != QVMHIP_F32) return fail(QVMHIP_EINVAL, "dtype must be F64 or F32");
    const uint64_t amp = dtype == QVMHIP_F64 ? 16 : 8;
    *out = (uint64_t(1) << n) * amp + scratch_bytes(n);
    return QVMHIP_OK;
}

int qvmhip_create(uint32_t n, int dtype, const qvmhip_config *cfg, qvmhip_state **out) {
    if (!out) return fail(QVMHIP_EINVAL, "null out");
    *out = nullptr;
    uint64_t need = 0;
    int rc = qvmhip_state_bytes(n, dtype, &need);
    if (rc) return rc;
    if (cfg && cfg->struct_size != sizeof(qvmhip_config))
        return fail(QVMHIP_EINVAL, "qvmhip_config.struct_size mismatch");
    if (cfg && (cfg->flags & ~QVMHIP_FLAG_FMA_TILES)) return fail(QVMHIP_EINVAL, "unknown config flags");
    if (cfg && cfg->memory_cap_bytes && need > cfg->memory_cap_bytes)
        return fail(QVMHIP_ENOMEM, "state needs " + std::to_string(need) + " bytes, cap is " +
                                       std::to_string(cfg->memory_cap_bytes));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(QVMHIP_EDEVICE, "no HIP device available");
    int dev = (cfg && cfg->device >= 0) ? cfg->device : -1;
    if (dev >= ndev) return fail(QVMHIP_EINVAL, "device ordinal out of range");
    if (dev < 0) QH_CHECK(hipGetDevice(&dev));
    DeviceScope ds(dev);  // the caller's current device is left as it was
    qvmhip_state *s = new qvmhip_state();
    s->n = n;
    s->dtype = dtype;
    s->device = dev;
    s->flags = cfg ? cfg->flags : 0u;
    s->len = uint64_t(1) << n;
    s->amp_bytes = dtype == QVMHIP_F64 ? 16 : 8;
    s->nchunks = std::max<uint64_t>(1, s->len >> kChunkLog2);
    const uint64_t nseg = (s->nchunks + 1023) / 1024;
    auto bail = [&](hipError_t e, const char *what) {
        free_state(s);
        return fail(e == hipErrorOutOfMemory ? QVMHIP_ENOMEM : QVMHIP_EDEVICE,
                    std::string(what) + ": " + hipGetErrorString(e));
    };
    hipError_t e;
    if ((e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(e, "hipStreamCreate");
    if ((e = hipMalloc(&s->amps, s->len * s->amp_bytes)) != hipSuccess) return bail(e, "hipMalloc(state)");
    if ((e = hipMalloc((void **)&s->red, sizeof(double) * (kRedBlocks * 2 + 16))) != hipSuccess)
        return bail(e, "hipMalloc(reduction)");
    if ((e = hipMalloc(&s->mat, kMatSlot)) != hipSuccess) return bail(e, "hipMalloc(matrix)");
    if ((e = hipMalloc((void **)&s->chunk_sum, sizeof(double) * s->nchunks)) != hipSuccess)
        return bail(e, "hipMalloc(chunks)");
    if ((e = hipMalloc((void **)&s->chunk_pre, sizeof(double) * s->nchunks)) != hipSuccess)
        return bail(e, "hipMalloc(prefix)");
    if ((e = hipMalloc((void **)&s->scan_tmp, sizeof(double) * nseg)) != hipSuccess)
        return bail(e, "hipMalloc(scan)");
    if ((e = hipHostMalloc((void **)&s->host, sizeof(double) * 64, hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc");
    s->stage_bytes = kStageBytes;
    if ((e = hipHostMalloc(&s->stage, s->stage_bytes, hipHostMallocDefault)) != hipSuccess)
        return bail(e, "hipHostMalloc(stage)");
    rc = launch_fill_zero_state(s);
    if (rc) {
        free_state(s);
        return rc;
    }
    if ((e = hipStreamSynchronize(s->stream)) != hipSuccess) return bail(e, "initial fill");
    *out = s;
    return QVMHIP_OK;
}

int qvmhip_destroy(qvmhip_state *s) {
    if (!s) return QVMHIP_OK;
    DeviceScope ds(s->device);
    (void)hipStreamSynchronize(s->stream);
    for (qvmhip_program *p : s->programs) program_detach(p);
    free_state(s);
    return QVMHIP_OK;
}

int qvmhip_num_qubits(const qvmhip_state *s, uint32_t *n) {
    if (!s || !n) return fail(QVMHIP_EINVAL, "null argument");
    *n = s->n;
    return QVMHIP_OK;
}

int qvmhip_dtype(const qvmhip_state *s, int *dtype) {
    if (!s || !dtype) return fail(QVMHIP_EINVAL, "null argument");
    *dtype = s->dtype;
    return QVMHIP_OK;
}

int qvmhip_reset_zero(qvmhip_state *s) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    return launch_fill_zero_state(s);
}

int qvmhip_sync(qvmhip_state *s) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    QH_CHECK(hipStreamSynchronize(s->stream));
    return QVMHIP_OK;
}

int qvmhip_apply_matrix(qvmhip_state *s, const double *U, uint32_t k, const uint32_t *qubits) {
    int rc = check_qubits(s, k, qubits);
    if (rc) return rc;
    DeviceScope ds(s->device);
    if (!U || !finite_matrix(U, k)) return fail(QVMHIP_EINVAL, "matrix is null or not finite");
    return launch_gate(s, U, k, qubits, 0);
}

int qvmhip_apply_matrix_dense(qvmhip_state *s, const double *U, uint32_t k, const uint32_t *qubits) {
    int rc = check_qubits(s, k, qubits);
    if (rc) return rc;
    DeviceScope ds(s->device);
    if (!U || !finite_matrix(U, k)) return fail(QVMHIP_EINVAL, "matrix is null or not finite");
    return launch_gate(s, U, k, qubits, 1);
}

int qvmhip_apply_diagonal(qvmhip_state *s, const double *d, uint32_t k, const uint32_t *qubits) {
    int rc = check_qubits(s, k, qubits);
    if (rc) return rc;
    DeviceScope ds(s->device);
    if (!d) return fail(QVMHIP_EINVAL, "null diagonal");
    for (uint32_t i = 0; i < (2u << k); ++i)
        if (!std::isfinite(d[i])) return fail(QVMHIP_EINVAL, "diagonal not finite");
    return launch_diagonal(s, d, k, qubits);
}

int qvmhip_apply_permutation(qvmhip_state *s, const uint32_t *perm, uint32_t k, const uint32_t *qubits) {
    int rc = check_qubits(s, k, qubits);
    if (rc) return rc;
    DeviceScope ds(s->device);
    if (!perm) return fail(QVMHIP_EINVAL, "null permutation");
    return launch_permutation(s, perm, k, qubits);
}

int qvmhip_apply_fused(qvmhip_state *s, const double *U, uint32_t k, const uint32_t *qubits) {
    int rc = check_qubits(s, k, qubits);
    if (rc) return rc;
    DeviceScope ds(s->device);
    if (k < 3) return fail(QVMHIP_EINVAL, "fused blocks are 3..6 qubits");
    if (!U || !finite_matrix(U, k)) return fail(QVMHIP_EINVAL, "matrix is null or not finite");
    // upload to the matrix slot's second half (first half: ops / dense)
    const size_t bytes = sizeof(double) * 2 * (size_t(1) << (2 * k));
    void *dst = (char *)s->mat + kMatSlot / 2;
    QH_CHECK(hipMemcpyAsync(dst, U, bytes, hipMemcpyHostToDevice, s->stream));
    return launch_fused_dense(s, dst, k, qubits);
}

int qvmhip_apply_tile_pass(qvmhip_state *s, uint32_t m, const uint32_t *tile_qubits, uint32_t nops,
                           const qvmhip_tile_op *ops) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    if (m > s->n) return fail(QVMHIP_EINVAL, "tile wider than the state");
    if (!tile_qubits || (nops && !ops)) return fail(QVMHIP_EINVAL, "null tile arguments");
    for (uint32_t i = 0; i < m; ++i) {
        if (tile_qubits[i] >= s->n) return fail(QVMHIP_EINVAL, "tile qubit out of range");
        if (i && tile_qubits[i] <= tile_qubits[i - 1]) return fail(QVMHIP_EINVAL, "tile qubits must ascend");
    }
    int rc = check_tile_ops(s->dtype, m, nops, ops);
    if (!rc) rc = check_tile_globals(s->n, m, tile_qubits, nops, ops);
    if (rc) return rc;
    if (nops == 0) return QVMHIP_OK;
    return launch_tile_pass(s, m, tile_qubits, nops, ops, 0);
}

int qvmhip_tile_schedule(int dtype, uint32_t flags, uint32_t m, uint32_t nops, const qvmhip_tile_op *ops,
                         qvmhip_tile_op *out, uint32_t cap, uint32_t *nout) {
    if (!nout || (nops && !ops) || (cap && !out)) return fail(QVMHIP_EINVAL, "null argument");
    if (dtype != QVMHIP_F64 && dtype != QVMHIP_F32) return fail(QVMHIP_EINVAL, "unknown dtype");
    if (flags & ~QVMHIP_FLAG_FMA_TILES) return fail(QVMHIP_EINVAL, "unknown flags");
    int rc = check_tile_ops(dtype, m, nops, ops);
    if (rc) return rc;
    std::vector<qvmhip_tile_op> sched;
    schedule_tile_ops(dtype, (flags & QVMHIP_FLAG_FMA_TILES) != 0, m, ops, nops, sched);
    *nout = (uint32_t)sched.size();
    if (sched.size() > cap) return fail(QVMHIP_ERANGE, "output capacity too small");
    std::copy(sched.begin(), sched.end(), out);
    return QVMHIP_OK;
}

int qvmhip_prob_excited(qvmhip_state *s, uint32_t q, double *p1) {
    if (!s || !p1) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (q >= s->n) return fail(QVMHIP_EINVAL, "qubit out of range");
    int rc = reduce_probability(s, 1, q, s->red + kRedBlocks * 2);
    if (rc) return rc;
    return scalar_to_host(s, s->red + kRedBlocks * 2, p1);
}

int qvmhip_prob_ground(qvmhip_state *s, uint32_t q, double *p0) {
    if (!s || !p0) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (q >= s->n) return fail(QVMHIP_EINVAL, "qubit out of range");
    int rc = reduce_probability(s, 2, q, s->red + kRedBlocks * 2);
    if (rc) return rc;
    return scalar_to_host(s, s->red + kRedBlocks * 2, p0);
}

int qvmhip_norm(qvmhip_state *s, double *out) {
    if (!s || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    double n2 = 0;
    int rc = reduce_norm2(s, s->red + kRedBlocks * 2);
    if (rc) return rc;
    rc = scalar_to_host(s, s->red + kRedBlocks * 2, &n2);
    if (rc) return rc;
    *out = std::sqrt(n2);
    return QVMHIP_OK;
}

int qvmhip_total_probability(qvmhip_state *s, double *out) {
    if (!s || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    int rc = reduce_norm2(s, s->red + kRedBlocks * 2);
    if (rc) return rc;
    return scalar_to_host(s, s->red + kRedBlocks * 2, out);
}

int qvmhip_normalize(qvmhip_state *s, double length) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    double len = length;
    if (!(len > 0)) {
        int rc = qvmhip_norm(s, &len);
        if (rc) return rc;
    }
    if (!(len > 0) || !std::isfinite(len)) return fail(QVMHIP_ERANGE, "cannot normalize a zero/non-finite state");
    return launch_scale(s, 1.0 / len);
}

int qvmhip_probabilities(qvmhip_state *s, uint64_t start, uint64_t count, double *out) {
    if (!s || (count && !out)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (start > s->len || count > s->len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    if (!count) return QVMHIP_OK;
    return probabilities(s, start, count, out);
}

int qvmhip_expectation_zz(qvmhip_state *s, const uint32_t *edges, uint32_t nedges, double *out) {
    if (!s || !out || (nedges && !edges)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    for (uint32_t i = 0; i < 2 * nedges; ++i)
        if (edges[i] >= s->n) return fail(QVMHIP_EINVAL, "edge qubit out of range");
    return expectation_zz(s, edges, nedges, out);
}

int qvmhip_inner_product(qvmhip_state *a, qvmhip_state *b, double *re, double *im) {
    if (!a || !b || !re || !im) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(a->device);
    if (a->n != b->n) return fail(QVMHIP_EINVAL, "states differ in size");
    if (a != b) {
        // b's queued work must finish before a's stream reads it
        QH_CHECK(hipStreamSynchronize(b->stream));
    }
    return inner_product(a, b, re, im);
}

int qvmhip_max_abs_diff(qvmhip_state *a, qvmhip_state *b, double *out) {
    if (!a || !b || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(a->device);
    if (a->n != b->n) return fail(QVMHIP_EINVAL, "states differ in size");
    if (a != b) QH_CHECK(hipStreamSynchronize(b->stream));
    return max_abs_diff(a, b, out);
}

int qvmhip_tv_distance(qvmhip_state *a, qvmhip_state *b, double *out) {
    if (!a || !b || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(a->device);
    if (a->n != b->n) return fail(QVMHIP_EINVAL, "states differ in size");
    if (a != b) QH_CHECK(hipStreamSynchronize(b->stream));
    return tv_distance(a, b, out);
}

int qvmhip_profile_enable(qvmhip_state *s, int on) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    QH_CHECK(hipStreamSynchronize(s->stream));
    for (auto &r : s->prof) {
        if (r.a) s->event_pool.push_back(r.a);
        if (r.b) s->event_pool.push_back(r.b);
    }
    s->prof.clear();
    s->profiling = on != 0;
    return QVMHIP_OK;
}

int qvmhip_profile_summary(qvmhip_state *s, int kind, uint64_t *count, double *total_ms) {
    if (!s || !count || !total_ms) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    QH_CHECK(hipStreamSynchronize(s->stream));
    uint64_t c = 0;
    double t = 0.0;
    for (auto &r : s->prof) {
        if (r.kind != kind || !r.a || !r.b) continue;
        float ms = 0.f;
        QH_CHECK(hipEventElapsedTime(&ms, r.a, r.b));
        t += ms;
        ++c;
    }
    *count = c;
    *total_ms = t;
    return QVMHIP_OK;
}

int qvmhip_force_measure(qvmhip_state *s, uint32_t q, int bit, double p1) {
    if (!s) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(s->device);
    if (q >= s->n) return fail(QVMHIP_EINVAL, "qubit out of range");
    if (bit != 0 && bit != 1) return fail(QVMHIP_EINVAL, "bit must be 0 or 1");
    if (!(p1 >= 0.0 && p1 <= 1.0)) return fail(QVMHIP_ERANGE, "excited probability outside [0,1]");
    // src/measurement.lisp:18-21: 1/sqrt(p1) for bit 1, 1/sqrt(1 - p1) for bit 0
    const double inv_norm = bit == 1 ? 1.0 / std::sqrt(p1) : 1.0 / std::sqrt(1.0 - p1);
    // the reference's (/ (sqrt 0)) signals division-by-zero there
    if (!std::isfinite(inv_norm))
        return fail(QVMHIP_ERANGE, "cannot force qubit " + std::to_string(q) + " to " + std::to_string(bit) +
                                       ": that outcome has probability 0");
    return launch_force_measure(s, q, bit, inv_norm, 0);
}

int qvmhip_measure(qvmhip_state *s, uint32_t q, double u, int compiled, int *bit, int *consumed) {
    if (!s || !bit || !consumed) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (q >= s->n) return fail(QVMHIP_EINVAL, "Trying to measure qubit " + std::to_string(q) + " on a QVM with only " +
                                                 std::to_string(s->n) + " qubits");
    if (!(u >= 0.0 && u <= 1.0)) return fail(QVMHIP_ERANGE, "uniform draw outside [0,1]");
    if (!compiled) {
        double p1 = 0;
        int rc = qvmhip_prob_excited(s, q, &p1);
        if (rc) return rc;
        int b;
        if (p1 == 0.0) {
            b = 0;
            *consumed = 0;
        } else {
            *consumed = 1;
            b = (u <= p1) ? 1 : 0;
        }
        const double pc = std::min(std::max(p1, 0.0), 1.0);
        const double inv_norm = b == 1 ? 1.0 / std::sqrt(pc) : 1.0 / std::sqrt(1.0 - pc);
        rc = launch_force_measure(s, q, b, inv_norm, 0);
        if (rc) return rc;
        *bit = b;
        return QVMHIP_OK;
    }
    double p0 = 0;
    int rc = qvmhip_prob_ground(s, q, &p0);
    if (rc) return rc;
    *consumed = 1;
    // the reference draws u in [0, 1), so u < p0 always holds when p0 >= 1;
    // keeping that for u = 1 avoids a 1/sqrt(1 - p0) of 0 or less
    const bool keep_zero = u < p0 || p0 >= 1.0;
    const int b = keep_zero ? 0 : 1;
    const double inv_norm = keep_zero ? 1.0 / std::sqrt(p0) : 1.0 / std::sqrt(1.0 - p0);
    rc = launch_force_measure(s, q, b, inv_norm, 1);
    if (rc) return rc;
    *bit = b;
    return QVMHIP_OK;
}

int qvmhip_measure_all(qvmhip_state *s, double u, uint64_t *index) {
    if (!s || !index) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (!(u >= 0.0 && u <= 1.0)) return fail(QVMHIP_ERANGE, "uniform draw outside [0,1]");
    int rc = sample_shots(s, &u, 1, nullptr, 0);
    if (rc) return rc;
    rc = collapse_to_basis(s, s->shot_out);
    if (rc) return rc;
    QH_CHECK(hipMemcpyAsync(s->host, s->shot_out, sizeof(uint64_t), hipMemcpyDeviceToHost, s->stream));
    QH_CHECK(hipStreamSynchronize(s->stream));
    std::memcpy(index, s->host, sizeof(uint64_t));
    return QVMHIP_OK;
}

static int check_draws(const double *u, uint64_t nshots) {
    for (uint64_t i = 0; i < nshots; ++i)
        if (!(u[i] >= 0.0 && u[i] <= 1.0)) return fail(QVMHIP_ERANGE, "uniform draw outside [0,1]");
    return QVMHIP_OK;
}

int qvmhip_sample_measure_all(qvmhip_state *s, const double *u, uint64_t nshots, uint64_t *out) {
    if (!s || (nshots && (!u || !out))) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    int rc = check_draws(u, nshots);
    if (rc) return rc;
    return sample_shots(s, u, nshots, out, 0);
}

int qvmhip_sample(qvmhip_state *s, const double *u, uint64_t nshots, uint64_t *out) {
    if (!s || (nshots && (!u || !out))) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    int rc = check_draws(u, nshots);
    if (rc) return rc;
    return sample_shots(s, u, nshots, out, 1);
}

int qvmhip_sample_ex(qvmhip_state *s, const double *u, uint64_t nshots, int rule, double eps, uint64_t *out,
                     uint8_t *near_flags, uint64_t *n_near) {
    if (!s || (nshots && (!u || !out))) return fail(QVMHIP_EINVAL, "null argument");
    if (rule != QVMHIP_RULE_GT && rule != QVMHIP_RULE_GE) return fail(QVMHIP_EINVAL, "rule must be GT or GE");
    if (!(eps >= 0.0) || !std::isfinite(eps)) return fail(QVMHIP_EINVAL, "eps must be finite and >= 0");
    DeviceScope ds(s->device);
    int rc = check_draws(u, nshots);
    if (rc) return rc;
    return sample_shots(s, u, nshots, out, rule == QVMHIP_RULE_GE ? 1 : 0, eps, near_flags, n_near);
}

int qvmhip_read_amplitudes(qvmhip_state *s, uint64_t start, uint64_t count, double *out) {
    if (!s || (count && !out)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (start > s->len || count > s->len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    if (!count) return QVMHIP_OK;
    return export_range(s, start, count, kExportAmps, out);
}

int qvmhip_read_wavefunction_be(qvmhip_state *s, uint64_t start, uint64_t count, uint8_t *out) {
    if (!s || (count && !out)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (start > s->len || count > s->len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    if (!count) return QVMHIP_OK;
    return export_range(s, start, count, kExportAmpsBE, out);
}

int qvmhip_read_probabilities_be(qvmhip_state *s, uint64_t start, uint64_t count, uint8_t *out) {
    if (!s || (count && !out)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (start > s->len || count > s->len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    if (!count) return QVMHIP_OK;
    return export_range(s, start, count, kExportProbsBE, out);
}

int qvmhip_write_amplitudes(qvmhip_state *s, uint64_t start, uint64_t count, const double *in) {
    if (!s || (count && !in)) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (start > s->len || count > s->len - start) return fail(QVMHIP_EINVAL, "range out of bounds");
    if (!count) return QVMHIP_OK;
    const uint64_t per = s->stage_bytes / 16;
    struct Wide {
        double *p = nullptr;
        hipStream_t st = nullptr;
        ~Wide() {
            if (p) {
                (void)hipStreamSynchronize(st);
                (void)hipFree(p);
            }
        }
    } wide_buf;
    wide_buf.st = s->stream;
    if (s->dtype == QVMHIP_F32) QH_CHECK(hipMalloc((void **)&wide_buf.p, per * 16));
    double *wide = wide_buf.p;
    for (uint64_t off = 0; off < count; off += per) {
        const uint64_t c = std::min<uint64_t>(per, count - off);
        QH_CHECK(hipStreamSynchronize(s->stream));  // stage is free again
        std::memcpy(s->stage, in + 2 * off, c * 16);
        if (s->dtype == QVMHIP_F64) {
            QH_CHECK(hipMemcpyAsync((double2 *)s->amps + start + off, s->stage, c * 16, hipMemcpyHostToDevice,
                                    s->stream));
        } else {
            QH_CHECK(hipMemcpyAsync(wide, s->stage, c * 16, hipMemcpyHostToDevice, s->stream));
            int rc = convert_f64_to_f32(s, start + off, c, wide);
            if (rc) return rc;
        }
    }
    QH_CHECK(hipStreamSynchronize(s->stream));
    return QVMHIP_OK;
}

int qvmhip_accumulate(qvmhip_state *dst, const qvmhip_state *src) {
    if (!dst || !src) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(dst->device);
    if (dst->n != src->n || dst->dtype != src->dtype) return fail(QVMHIP_EINVAL, "states differ in size or dtype");
    if (dst == src) return accumulate(dst, src);
    QH_CHECK(stream_after(dst->stream, src->stream));  // src's queued work first
    int rc = accumulate(dst, src);
    if (rc) return rc;
    QH_CHECK(stream_after(src->stream, dst->stream));  // src's next call after the read
    return QVMHIP_OK;
}

int qvmhip_density_diagonal(qvmhip_state *s, uint32_t nq, double *out) {
    if (!s || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(s->device);
    if (2 * nq != s->n) return fail(QVMHIP_EINVAL, "a density state of nq qubits has 2*nq qubits");
    return density_diagonal(s, nq, out);
}

int qvmhip_copy(qvmhip_state *dst, const qvmhip_state *src) {
    if (!dst || !src) return fail(QVMHIP_EINVAL, "null state");
    DeviceScope ds(dst->device);
    if (dst->n != src->n || dst->dtype != src->dtype) return fail(QVMHIP_EINVAL, "states differ in size or dtype");
    if (dst == src) return QVMHIP_OK;
    QH_CHECK(stream_after(dst->stream, src->stream));  // src's queued work first
    QH_CHECK(hipMemcpyAsync(dst->amps, src->amps, dst->len * dst->amp_bytes, hipMemcpyDeviceToDevice, dst->stream));
    QH_CHECK(stream_after(src->stream, dst->stream));  // src's next call after the copy
    return QVMHIP_OK;
}

}  // extern "C"
