// mlp.hip -- the fusion cost model: a small MLP (din -> h -> h -> 1, ReLU)
// trained and evaluated on the GPU by these kernels.  It predicts log launch
// time of a planned launch (fused block, tile pass or plain gate) from
// features of the launch; the planner sums predictions to pick a plan.
//
// Training is full-batch Adam on mean-squared error.  Per-sample gradients
// are written to a [batch x nparams] buffer and summed per parameter in a
// fixed order (deterministic; no atomics).  Everything is FP32.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qvmhip_internal.h"

using namespace qvmhip;

namespace {

constexpr int kMaxH = 64;
constexpr int kMaxIn = 32;

struct Shape {
    int din, h;
    // parameter offsets in one flat vector
    int w1, b1, w2, b2, w3, b3, np;
};

Shape make_shape(int din, int h) {
    Shape s;
    s.din = din;
    s.h = h;
    s.w1 = 0;
    s.b1 = s.w1 + din * h;
    s.w2 = s.b1 + h;
    s.b2 = s.w2 + h * h;
    s.w3 = s.b2 + h;
    s.b3 = s.w3 + h;
    s.np = s.b3 + 1;
    return s;
}

// forward for one sample; returns output, fills activations
__device__ inline float fwd(const Shape &S, const float *__restrict__ p, const float *__restrict__ x, float *a1,
                            float *a2) {
    for (int j = 0; j < S.h; ++j) {
        float z = p[S.b1 + j];
        for (int i = 0; i < S.din; ++i) z += p[S.w1 + j * S.din + i] * x[i];
        a1[j] = z > 0.f ? z : 0.f;
    }
    for (int j = 0; j < S.h; ++j) {
        float z = p[S.b2 + j];
        for (int i = 0; i < S.h; ++i) z += p[S.w2 + j * S.h + i] * a1[i];
        a2[j] = z > 0.f ? z : 0.f;
    }
    float y = p[S.b3];
    for (int i = 0; i < S.h; ++i) y += p[S.w3 + i] * a2[i];
    return y;
}

__global__ void k_mlp_forward(Shape S, const float *__restrict__ p, const float *__restrict__ X, int nb,
                              float *__restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    float a1[kMaxH], a2[kMaxH];
    out[b] = fwd(S, p, X + (size_t)b * S.din, a1, a2);
}

// per-sample gradient of 0.5*(y - t)^2 / nb into G[b * np + param]
__global__ void k_mlp_sample_grads(Shape S, const float *__restrict__ p, const float *__restrict__ X,
                                   const float *__restrict__ T, int nb, float *__restrict__ G,
                                   float *__restrict__ loss) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    const float *x = X + (size_t)b * S.din;
    float a1[kMaxH], a2[kMaxH], d2[kMaxH], d1[kMaxH];
    const float y = fwd(S, p, x, a1, a2);
    const float e = (y - T[b]) / (float)nb;
    loss[b] = 0.5f * (y - T[b]) * (y - T[b]) / (float)nb;
    float *g = G + (size_t)b * S.np;
    g[S.b3] = e;
    for (int i = 0; i < S.h; ++i) {
        g[S.w3 + i] = e * a2[i];
        d2[i] = a2[i] > 0.f ? e * p[S.w3 + i] : 0.f;
    }
    for (int j = 0; j < S.h; ++j) {
        g[S.b2 + j] = d2[j];
        for (int i = 0; i < S.h; ++i) g[S.w2 + j * S.h + i] = d2[j] * a1[i];
    }
    for (int i = 0; i < S.h; ++i) {
        float acc = 0.f;
        for (int j = 0; j < S.h; ++j) acc += p[S.w2 + j * S.h + i] * d2[j];
        d1[i] = a1[i] > 0.f ? acc : 0.f;
    }
    for (int j = 0; j < S.h; ++j) {
        g[S.b1 + j] = d1[j];
        for (int i = 0; i < S.din; ++i) g[S.w1 + j * S.din + i] = d1[j] * x[i];
    }
}

// sum per-sample gradients over the batch, one thread per parameter, in
// sample order; also sums the loss into out_loss (thread 0).
__global__ void k_mlp_reduce(Shape S, const float *__restrict__ G, const float *__restrict__ loss, int nb,
                             float *__restrict__ grad, float *__restrict__ out_loss) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < S.np) {
        float acc = 0.f;
        for (int b = 0; b < nb; ++b) acc += G[(size_t)b * S.np + i];
        grad[i] = acc;
    }
    if (i == 0) {
        float l = 0.f;
        for (int b = 0; b < nb; ++b) l += loss[b];
        out_loss[0] = l;
    }
}

__global__ void k_adam(int np, float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                       float *__restrict__ v, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const float gi = g[i];
    m[i] = b1 * m[i] + (1.f - b1) * gi;
    v[i] = b2 * v[i] + (1.f - b2) * gi * gi;
    const float mh = m[i] / c1, vh = v[i] / c2;
    p[i] -= lr * mh / (sqrtf(vh) + eps);
}

}  // namespace

struct qvmhip_mlp {
    Shape S;
    hipStream_t stream = nullptr;
    float *p = nullptr, *m = nullptr, *v = nullptr, *grad = nullptr, *loss1 = nullptr;
    float *X = nullptr, *T = nullptr, *G = nullptr, *lossb = nullptr, *out = nullptr;
    int cap = 0;
    int step = 0;
    int device = 0;  // the caller's current device at creation
};

namespace {
int ensure_batch(qvmhip_mlp *h, int nb) {
    if (nb <= h->cap) return QVMHIP_OK;
    for (float *q : {h->X, h->T, h->G, h->lossb, h->out})
        if (q) (void)hipFree(q);
    QH_CHECK(hipMalloc((void **)&h->X, sizeof(float) * nb * h->S.din));
    QH_CHECK(hipMalloc((void **)&h->T, sizeof(float) * nb));
    QH_CHECK(hipMalloc((void **)&h->G, sizeof(float) * (size_t)nb * h->S.np));
    QH_CHECK(hipMalloc((void **)&h->lossb, sizeof(float) * nb));
    QH_CHECK(hipMalloc((void **)&h->out, sizeof(float) * nb));
    h->cap = nb;
    return QVMHIP_OK;
}
}  // namespace

extern "C" {

int qvmhip_mlp_create(int din, int hidden, const float *init_params, qvmhip_mlp **out) {
    if (!out || din < 1 || din > kMaxIn || hidden < 1 || hidden > kMaxH)
        return fail(QVMHIP_EINVAL, "mlp: din must be 1..32 and hidden 1..64");
    *out = nullptr;
    qvmhip_mlp *h = new qvmhip_mlp();
    h->S = make_shape(din, hidden);
    hipError_t e = hipGetDevice(&h->device);
    const size_t bytes = sizeof(float) * h->S.np;
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking);
    for (float **q : {&h->p, &h->m, &h->v, &h->grad})
        if (e == hipSuccess) e = hipMalloc((void **)q, bytes);
    if (e == hipSuccess) e = hipMalloc((void **)&h->loss1, sizeof(float));
    if (e == hipSuccess)
        e = init_params ? hipMemcpyAsync(h->p, init_params, bytes, hipMemcpyHostToDevice, h->stream)
                        : hipMemsetAsync(h->p, 0, bytes, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->m, 0, bytes, h->stream);
    if (e == hipSuccess) e = hipMemsetAsync(h->v, 0, bytes, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        qvmhip_mlp_destroy(h);
        return fail(e == hipErrorOutOfMemory ? QVMHIP_ENOMEM : QVMHIP_EDEVICE,
                    std::string("mlp create: ") + hipGetErrorString(e));
    }
    *out = h;
    return QVMHIP_OK;
}

int qvmhip_mlp_num_params(const qvmhip_mlp *h, int *np) {
    if (!h || !np) return fail(QVMHIP_EINVAL, "null argument");
    *np = h->S.np;
    return QVMHIP_OK;
}

int qvmhip_mlp_get_params(qvmhip_mlp *h, float *out) {
    if (!h || !out) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(h->device);
    QH_CHECK(hipMemcpyAsync(out, h->p, sizeof(float) * h->S.np, hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    return QVMHIP_OK;
}

int qvmhip_mlp_set_params(qvmhip_mlp *h, const float *in) {
    if (!h || !in) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(h->device);
    QH_CHECK(hipMemcpyAsync(h->p, in, sizeof(float) * h->S.np, hipMemcpyHostToDevice, h->stream));
    h->step = 0;
    QH_CHECK(hipMemsetAsync(h->m, 0, sizeof(float) * h->S.np, h->stream));
    QH_CHECK(hipMemsetAsync(h->v, 0, sizeof(float) * h->S.np, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    return QVMHIP_OK;
}

int qvmhip_mlp_predict(qvmhip_mlp *h, const float *X, int nb, float *out) {
    if (!h || nb < 0 || (nb && (!X || !out))) return fail(QVMHIP_EINVAL, "null argument");
    DeviceScope ds(h->device);
    if (!nb) return QVMHIP_OK;
    int rc = ensure_batch(h, nb);
    if (rc) return rc;
    QH_CHECK(hipMemcpyAsync(h->X, X, sizeof(float) * nb * h->S.din, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_mlp_forward, dim3((nb + 127) / 128), dim3(128), 0, h->stream, h->S, h->p, h->X, nb, h->out);
    QH_LAUNCH_CHECK();
    QH_CHECK(hipMemcpyAsync(out, h->out, sizeof(float) * nb, hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    return QVMHIP_OK;
}

// Gradient of the mean of 0.5*(y - t)^2 at the current parameters.
int qvmhip_mlp_gradient(qvmhip_mlp *h, const float *X, const float *T, int nb, float *grad_out, float *loss_out) {
    if (!h || nb < 1 || !X || !T || !grad_out) return fail(QVMHIP_EINVAL, "bad argument");
    DeviceScope ds(h->device);
    int rc = ensure_batch(h, nb);
    if (rc) return rc;
    QH_CHECK(hipMemcpyAsync(h->X, X, sizeof(float) * nb * h->S.din, hipMemcpyHostToDevice, h->stream));
    QH_CHECK(hipMemcpyAsync(h->T, T, sizeof(float) * nb, hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_mlp_sample_grads, dim3((nb + 127) / 128), dim3(128), 0, h->stream, h->S, h->p, h->X, h->T,
                       nb, h->G, h->lossb);
    QH_LAUNCH_CHECK();
    hipLaunchKernelGGL(k_mlp_reduce, dim3((h->S.np + 255) / 256), dim3(256), 0, h->stream, h->S, h->G, h->lossb, nb,
                       h->grad, h->loss1);
    QH_LAUNCH_CHECK();
    QH_CHECK(hipMemcpyAsync(grad_out, h->grad, sizeof(float) * h->S.np, hipMemcpyDeviceToHost, h->stream));
    float l = 0.f;
    QH_CHECK(hipMemcpyAsync(&l, h->loss1, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    if (loss_out) *loss_out = l;
    return QVMHIP_OK;
}

// `epochs` full-batch Adam steps; *final_loss = loss before the last step.
int qvmhip_mlp_train(qvmhip_mlp *h, const float *X, const float *T, int nb, int epochs, float lr,
                     float *final_loss) {
    if (!h || nb < 1 || !X || !T || epochs < 1 || !(lr > 0)) return fail(QVMHIP_EINVAL, "bad argument");
    DeviceScope ds(h->device);
    int rc = ensure_batch(h, nb);
    if (rc) return rc;
    QH_CHECK(hipMemcpyAsync(h->X, X, sizeof(float) * nb * h->S.din, hipMemcpyHostToDevice, h->stream));
    QH_CHECK(hipMemcpyAsync(h->T, T, sizeof(float) * nb, hipMemcpyHostToDevice, h->stream));
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int e = 0; e < epochs; ++e) {
        hipLaunchKernelGGL(k_mlp_sample_grads, dim3((nb + 127) / 128), dim3(128), 0, h->stream, h->S, h->p, h->X,
                           h->T, nb, h->G, h->lossb);
        hipLaunchKernelGGL(k_mlp_reduce, dim3((h->S.np + 255) / 256), dim3(256), 0, h->stream, h->S, h->G,
                           h->lossb, nb, h->grad, h->loss1);
        ++h->step;
        const float c1 = 1.f - std::pow(b1, (float)h->step), c2 = 1.f - std::pow(b2, (float)h->step);
        hipLaunchKernelGGL(k_adam, dim3((h->S.np + 255) / 256), dim3(256), 0, h->stream, h->S.np, h->p, h->grad,
                           h->m, h->v, lr, b1, b2, eps, c1, c2);
    }
    QH_LAUNCH_CHECK();
    float l = 0.f;
    QH_CHECK(hipMemcpyAsync(&l, h->loss1, sizeof(float), hipMemcpyDeviceToHost, h->stream));
    QH_CHECK(hipStreamSynchronize(h->stream));
    if (final_loss) *final_loss = l;
    return QVMHIP_OK;
}

int qvmhip_mlp_destroy(qvmhip_mlp *h) {
    if (!h) return QVMHIP_OK;
    DeviceScope ds(h->device);
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    for (float *q : {h->p, h->m, h->v, h->grad, h->loss1, h->X, h->T, h->G, h->lossb, h->out})
        if (q) (void)hipFree(q);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    delete h;
    return QVMHIP_OK;
}

}  // extern "C"
