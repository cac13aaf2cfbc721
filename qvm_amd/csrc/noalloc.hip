// noalloc.hip -- the sampler of qvmhip_sample_ex for callers that cannot
// allocate device memory (code under HIP graph capture, serving loops that
// forbid hipMalloc): it runs on a statevector the caller owns, takes the
// draws and returns the indices in the caller's device memory, and keeps
// its chunk sums, chunk prefix and scan totals in `scratch`, a static device
// array of fixed size that each GPU holds for the life of the process.
//
// Every call on one GPU shares that GPU's array and leaves the last call's
// contents in it; calls that overlap on one GPU (two streams, two threads)
// race.  Nothing in it must start at zero: each call rewrites the chunk sums,
// the segment totals and the prefix it reads.  The caller's near-boundary
// count is cleared by a fill kernel (no hipMemsetAsync, so the call may be
// captured into a graph).
#include <mutex>

#include "qvmhip_internal.h"

using namespace qvmhip;

// Largest state: 2^26 amplitudes -> 2^18 chunks of 256.
constexpr uint32_t kNoAllocMaxQubits = 26;
constexpr uint64_t kNoAllocChunks = uint64_t(1) << (kNoAllocMaxQubits - kChunkLog2);
constexpr uint64_t kNoAllocSegs = (kNoAllocChunks + 1023) / 1024;
// chunk sums | chunk prefix | segment totals: 4 MiB + 2 KiB per GPU
constexpr uint64_t kScratchDoubles = 2 * kNoAllocChunks + kNoAllocSegs;

static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mu;
double *g_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
int scratch_on_current_device(double **out) {
    int dev = -1;
    QH_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) return fail(QVMHIP_EDEVICE, "device ordinal out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_ptr[dev]) {
        void *p = nullptr;
        QH_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(QVMHIP_EDEVICE, "scratch symbol has no device address");
        g_ptr[dev] = (double *)p;
    }
    *out = g_ptr[dev];
    return QVMHIP_OK;
}

__global__ void k_clear_count(unsigned long long *c) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *c = 0ull;
}

}  // namespace

extern "C" int qvmhip_sample_noalloc(const void *amps, uint32_t n, int dtype, const double *u, uint64_t nshots,
                                     int rule, double eps, uint64_t *out, uint64_t *n_near, void *stream) {
    if (n < 1 || n > kNoAllocMaxQubits)
        return fail(QVMHIP_EINVAL, "qvmhip_sample_noalloc: qubit count must be 1.." +
                                       std::to_string(kNoAllocMaxQubits) + " (static scratch of " +
                                       std::to_string(kScratchDoubles * 8) + " bytes)");
    if (dtype != QVMHIP_F64 && dtype != QVMHIP_F32) return fail(QVMHIP_EINVAL, "dtype must be F64 or F32");
    if (rule != QVMHIP_RULE_GT && rule != QVMHIP_RULE_GE) return fail(QVMHIP_EINVAL, "rule must be GT or GE");
    if (!(eps >= 0.0) || !(eps < 1.0)) return fail(QVMHIP_EINVAL, "eps must lie in [0, 1)");
    if (!amps || (nshots && (!u || !out))) return fail(QVMHIP_EINVAL, "null argument");
    double *sc = nullptr;
    int rc = scratch_on_current_device(&sc);
    if (rc) return rc;
    hipStream_t st = (hipStream_t)stream;
    const uint64_t len = uint64_t(1) << n;
    double *chunk_sum = sc, *chunk_pre = sc + kNoAllocChunks, *seg = sc + 2 * kNoAllocChunks;
    if (n_near) {
        hipLaunchKernelGGL(k_clear_count, dim3(1), dim3(64), 0, st, (unsigned long long *)n_near);
        QH_LAUNCH_CHECK();
    }
    if (nshots == 0) return QVMHIP_OK;
    rc = launch_chunk_prefix(st, dtype, amps, len, chunk_sum, seg, chunk_pre);
    if (rc) return rc;
    return launch_sample_search(st, dtype, amps, len, chunk_pre, u, nshots, rule == QVMHIP_RULE_GE ? 1 : 0, out, eps,
                                nullptr, (unsigned long long *)n_near);
}
