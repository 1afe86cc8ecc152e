// Multilayer perceptron of the learned initial density (no counterpart in
// the reference): tanh or SiLU hidden layers, linear output, mean-squared-error loss
// and Adam, all in FP32 HIP kernels for gfx950.  The layer sizes are small
// (tens of units) and the batch is large (thousands of monomers), so each
// dense layer is a lane-per-output kernel over a batch tile; weight
// gradients reduce over the batch with one workgroup per weight.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <algorithm>
#include <mutex>
#include <new>
#include <random>
#include <string>
#include <vector>
#include "../../include/mmd.h"

namespace mmd {
void set_last_error(const char* msg);   // the library's mmd_last_error() text
std::recursive_mutex& device_alloc_mutex();   // held around allocations and graph captures
}

namespace {

int mlp_fail(int code, const char* fmt, ...);

template <typename T>
struct Buf {
    T* p = nullptr;
    size_t n = 0;
    Buf() = default;
    Buf(const Buf&) = delete;
    Buf& operator=(const Buf&) = delete;
    Buf(Buf&& o) noexcept : p(o.p), n(o.n)
    {
        o.p = nullptr;
        o.n = 0;
    }
    ~Buf()
    {
        if (p) {
            std::lock_guard<std::recursive_mutex> lk(mmd::device_alloc_mutex());
            (void)hipFree(p);
        }
    }
    hipError_t reserve(size_t count)
    {
        if (count <= n && p) return hipSuccess;
        std::lock_guard<std::recursive_mutex> lk(mmd::device_alloc_mutex());
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) n = count;
        return e;
    }
};

#define MLP_TRY(x)                                                                              \
    do {                                                                                        \
        hipError_t e_ = (x);                                                                    \
        if (e_ != hipSuccess) return mlp_fail(MMD_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// -- kernels ------------------------------------------------------------------

__global__ void mlp_gather(const float* __restrict__ X, const int* __restrict__ perm, int B, int d,
                           float* __restrict__ out)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * d) return;
    const int i = t / d, j = t - i * d;
    out[t] = X[(size_t)perm[i] * d + j];
}

// Activations of the hidden layers
constexpr int ACT_LINEAR = 0, ACT_TANH = 1, ACT_SILU = 2;

__device__ __forceinline__ float sigmoidf(float z) { return 1.f / (1.f + expf(-z)); }

// out[i, j] = act(z), z = bias[j] + sum_k H[i, k] W[k, j]; zout (optional)
// keeps z for the backward pass of SiLU
__global__ void mlp_dense_fwd(const float* __restrict__ H, const float* __restrict__ W,
                              const float* __restrict__ bias, int B, int a, int b, int act, float* __restrict__ out,
                              float* __restrict__ zout)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * b) return;
    const int i = t / b, j = t - i * b;
    float s = bias[j];
    const float* h = H + (size_t)i * a;
    for (int k = 0; k < a; ++k) s = fmaf(h[k], W[(size_t)k * b + j], s);
    if (zout) zout[t] = s;
    out[t] = act == ACT_TANH ? tanhf(s) : (act == ACT_SILU ? s * sigmoidf(s) : s);
}

// dY = scale * 2 (Y - T); loss += scale * sum (Y - T)^2
__global__ void mlp_out_grad(const float* __restrict__ Y, const float* __restrict__ T, int n, float scale,
                             float* __restrict__ dY, float* loss)
{
    __shared__ float red[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    float l = 0.f;
    if (t < n) {
        const float r = Y[t] - T[t];
        dY[t] = 2.f * scale * r;
        l = scale * r * r;
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, red[0]);
}

// dW[k, j] = sum_i H[i, k] dZ[i, j] (blocks 0 .. a*b-1); db[j] = sum_i dZ[i, j]
// (blocks a*b .. a*b+b-1).  One 256-thread workgroup per output.
__global__ void mlp_dense_bwd_w(const float* __restrict__ H, const float* __restrict__ dZ, int B, int a, int b,
                                float* __restrict__ dW, float* __restrict__ db)
{
    __shared__ float red[256];
    const int o = blockIdx.x;
    float s = 0.f;
    if (o < a * b) {
        const int k = o / b, j = o - k * b;
        for (int i = threadIdx.x; i < B; i += blockDim.x) s = fmaf(H[(size_t)i * a + k], dZ[(size_t)i * b + j], s);
    } else {
        const int j = o - a * b;
        for (int i = threadIdx.x; i < B; i += blockDim.x) s += dZ[(size_t)i * b + j];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = blockDim.x / 2; st > 0; st >>= 1) {
        if ((int)threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if (o < a * b) dW[o] = red[0];
        else db[o - a * b] = red[0];
    }
}

// dZprev[i, k] = (sum_j dZ[i, j] W[k, j]) * act'(z[i, k]): 1 - H^2 for tanh
// (H its output), sigma(z) (1 + z (1 - sigma(z))) for SiLU (Z its input)
__global__ void mlp_dense_bwd_x(const float* __restrict__ dZ, const float* __restrict__ W,
                                const float* __restrict__ H, const float* __restrict__ Z, int act, int B, int a,
                                int b, float* __restrict__ dprev)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * a) return;
    const int i = t / a, k = t - i * a;
    float s = 0.f;
    const float* dz = dZ + (size_t)i * b;
    const float* w = W + (size_t)k * b;
    for (int j = 0; j < b; ++j) s = fmaf(dz[j], w[j], s);
    if (act == ACT_SILU) {
        const float z = Z[t], sg = sigmoidf(z);
        dprev[t] = s * sg * (1.f + z * (1.f - sg));
    } else {
        const float h = H[t];
        dprev[t] = s * (1.f - h * h);
    }
}

__global__ void mlp_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                         float* __restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const float gi = g[t];
    const float mi = b1 * m[t] + (1.f - b1) * gi;
    const float vi = b2 * v[t] + (1.f - b2) * gi * gi;
    m[t] = mi;
    v[t] = vi;
    p[t] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

__global__ void mlp_zero(float* x, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[t] = 0.f;
}

inline unsigned blocks(size_t n) { return (unsigned)((n + 255) / 256); }

int mlp_fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    mmd::set_last_error(buf);
    return code;
}

}  // namespace

struct mmd_mlp {
    int device = 0;
    hipStream_t st = nullptr;
    std::vector<int> dims;       // nlayer + 1
    std::vector<size_t> woff;    // offset of W_l in the flat parameter vector
    std::vector<size_t> boff;    // offset of b_l
    size_t nparam = 0;
    Buf<float> param, grad, m, v, loss;
    Buf<float> X, Y;             // training data on the device
    Buf<int> perm;
    std::vector<Buf<float>> act; // activations per layer (act[0] = input batch)
    std::vector<Buf<float>> pre; // SiLU: pre-activations of the hidden layers
    std::vector<Buf<float>> dz;  // gradients w.r.t. pre-activations per layer
    int adam_t = 0;
    int hidden_act = ACT_TANH;   // activation of the hidden layers
};

namespace {

int mlp_reserve(mmd_mlp* h, int B)
{
    const int L = (int)h->dims.size() - 1;
    for (int l = 0; l <= L; ++l) MLP_TRY(h->act[l].reserve((size_t)B * h->dims[l]));
    for (int l = 1; l <= L; ++l) MLP_TRY(h->dz[l].reserve((size_t)B * h->dims[l]));
    if (h->hidden_act == ACT_SILU)
        for (int l = 1; l < L; ++l) MLP_TRY(h->pre[l].reserve((size_t)B * h->dims[l]));
    return MMD_OK;
}

// forward of a batch already in act[0]
int mlp_forward_batch(mmd_mlp* h, int B)
{
    const int L = (int)h->dims.size() - 1;
    for (int l = 1; l <= L; ++l) {
        const int a = h->dims[l - 1], b = h->dims[l];
        const bool hidden = l < L;
        hipLaunchKernelGGL(mlp_dense_fwd, dim3(blocks((size_t)B * b)), dim3(256), 0, h->st, h->act[l - 1].p,
                           h->param.p + h->woff[l - 1], h->param.p + h->boff[l - 1], B, a, b,
                           hidden ? h->hidden_act : ACT_LINEAR, h->act[l].p,
                           hidden && h->hidden_act == ACT_SILU ? h->pre[l].p : nullptr);
        MLP_TRY(hipGetLastError());
    }
    return MMD_OK;
}

// loss and gradients of a batch in act[0] against targets T (device);
// gradients are written to h->grad, the loss added to h->loss[0]
int mlp_backward_batch(mmd_mlp* h, int B, const float* T, float scale)
{
    const int L = (int)h->dims.size() - 1;
    const int dl = h->dims[L];
    hipLaunchKernelGGL(mlp_out_grad, dim3(blocks((size_t)B * dl)), dim3(256), 0, h->st, h->act[L].p, T, B * dl,
                       scale, h->dz[L].p, h->loss.p);
    MLP_TRY(hipGetLastError());
    for (int l = L; l >= 1; --l) {
        const int a = h->dims[l - 1], b = h->dims[l];
        hipLaunchKernelGGL(mlp_dense_bwd_w, dim3(a * b + b), dim3(256), 0, h->st, h->act[l - 1].p, h->dz[l].p, B, a,
                           b, h->grad.p + h->woff[l - 1], h->grad.p + h->boff[l - 1]);
        MLP_TRY(hipGetLastError());
        if (l > 1) {
            hipLaunchKernelGGL(mlp_dense_bwd_x, dim3(blocks((size_t)B * a)), dim3(256), 0, h->st, h->dz[l].p,
                               h->param.p + h->woff[l - 1], h->act[l - 1].p,
                               h->hidden_act == ACT_SILU ? h->pre[l - 1].p : nullptr, h->hidden_act, B, a, b,
                               h->dz[l - 1].p);
            MLP_TRY(hipGetLastError());
        }
    }
    return MMD_OK;
}

int mlp_check(mmd_mlp* h)
{
    if (!h) return mlp_fail(MMD_ERR_ARG, "null perceptron");
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) return mlp_fail(MMD_ERR_DEVICE, "hipSetDevice: %s", hipGetErrorString(e));
    return MMD_OK;
}

}  // namespace

extern "C" {

int mmd_mlp_create(mmd_mlp** out, int nlayer, const int32_t* dims, int device_id)
{
    if (!out || nlayer < 1 || nlayer > 8 || !dims) return mlp_fail(MMD_ERR_ARG, "invalid perceptron shape");
    for (int l = 0; l <= nlayer; ++l)
        if (dims[l] < 1 || dims[l] > 4096) return mlp_fail(MMD_ERR_ARG, "layer width %d out of range", dims[l]);
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return mlp_fail(MMD_ERR_DEVICE, "no HIP device available");
    if (device_id < 0 || device_id >= n) return mlp_fail(MMD_ERR_ARG, "device %d out of range", device_id);
    MLP_TRY(hipSetDevice(device_id));
    mmd_mlp* h = new (std::nothrow) mmd_mlp();
    if (!h) return mlp_fail(MMD_ERR_NOMEM, "perceptron allocation failed");
    h->device = device_id;
    h->dims.assign(dims, dims + nlayer + 1);
    size_t off = 0;
    for (int l = 0; l < nlayer; ++l) {
        h->woff.push_back(off);
        off += (size_t)dims[l] * dims[l + 1];
        h->boff.push_back(off);
        off += dims[l + 1];
    }
    h->nparam = off;
    h->act.resize(nlayer + 1);
    h->pre.resize(nlayer + 1);
    h->dz.resize(nlayer + 1);
    hipError_t e = hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking);
    if (e == hipSuccess) e = h->param.reserve(off);
    if (e == hipSuccess) e = h->grad.reserve(off);
    if (e == hipSuccess) e = h->m.reserve(off);
    if (e == hipSuccess) e = h->v.reserve(off);
    if (e == hipSuccess) e = h->loss.reserve(1);
    if (e != hipSuccess) {
        mmd_mlp_destroy(h);
        return mlp_fail(MMD_ERR_DEVICE, "perceptron setup: %s", hipGetErrorString(e));
    }
    *out = h;
    return MMD_OK;
}

void mmd_mlp_destroy(mmd_mlp* h)
{
    if (!h) return;
    (void)hipSetDevice(h->device);
    hipStream_t st = h->st;
    if (st) (void)hipStreamSynchronize(st);
    delete h;
    if (st) (void)hipStreamDestroy(st);
}

int64_t mmd_mlp_nparam(mmd_mlp* h) { return h ? (int64_t)h->nparam : (int64_t)MMD_ERR_ARG; }

int mmd_mlp_set_activation(mmd_mlp* h, int activation)
{
    if (!h) return mlp_fail(MMD_ERR_ARG, "null perceptron");
    if (activation != 0 && activation != 1) return mlp_fail(MMD_ERR_ARG, "activation must be 0 (tanh) or 1 (SiLU)");
    h->hidden_act = activation == 1 ? ACT_SILU : ACT_TANH;
    return MMD_OK;
}

int mmd_mlp_set_params(mmd_mlp* h, const float* params)
{
    int rc = mlp_check(h);
    if (rc) return rc;
    if (!params) return mlp_fail(MMD_ERR_ARG, "null pointer");
    MLP_TRY(hipMemcpyAsync(h->param.p, params, sizeof(float) * h->nparam, hipMemcpyHostToDevice, h->st));
    hipLaunchKernelGGL(mlp_zero, dim3(blocks(h->nparam)), dim3(256), 0, h->st, h->m.p, (int)h->nparam);
    hipLaunchKernelGGL(mlp_zero, dim3(blocks(h->nparam)), dim3(256), 0, h->st, h->v.p, (int)h->nparam);
    MLP_TRY(hipGetLastError());
    MLP_TRY(hipStreamSynchronize(h->st));
    h->adam_t = 0;
    return MMD_OK;
}

int mmd_mlp_get_params(mmd_mlp* h, float* params)
{
    int rc = mlp_check(h);
    if (rc) return rc;
    if (!params) return mlp_fail(MMD_ERR_ARG, "null pointer");
    MLP_TRY(hipMemcpyAsync(params, h->param.p, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->st));
    MLP_TRY(hipStreamSynchronize(h->st));
    return MMD_OK;
}

int mmd_mlp_forward(mmd_mlp* h, int n, const float* X, float* Y)
{
    int rc = mlp_check(h);
    if (rc) return rc;
    if (n < 0 || (n > 0 && (!X || !Y))) return mlp_fail(MMD_ERR_ARG, "invalid forward arguments");
    const int L = (int)h->dims.size() - 1;
    const int CH = 8192;
    rc = mlp_reserve(h, std::min(n, CH));
    if (rc) return rc;
    for (int i0 = 0; i0 < n; i0 += CH) {
        const int B = std::min(CH, n - i0);
        MLP_TRY(hipMemcpyAsync(h->act[0].p, X + (size_t)i0 * h->dims[0], sizeof(float) * B * h->dims[0],
                               hipMemcpyHostToDevice, h->st));
        rc = mlp_forward_batch(h, B);
        if (rc) return rc;
        MLP_TRY(hipMemcpyAsync(Y + (size_t)i0 * h->dims[L], h->act[L].p, sizeof(float) * B * h->dims[L],
                               hipMemcpyDeviceToHost, h->st));
    }
    MLP_TRY(hipStreamSynchronize(h->st));
    return MMD_OK;
}

int mmd_mlp_loss_grad(mmd_mlp* h, int n, const float* X, const float* Yt, float* loss, float* grad)
{
    int rc = mlp_check(h);
    if (rc) return rc;
    if (n < 1 || !X || !Yt) return mlp_fail(MMD_ERR_ARG, "invalid gradient arguments");
    const int L = (int)h->dims.size() - 1;
    rc = mlp_reserve(h, n);
    if (rc) return rc;
    Buf<float> T;
    MLP_TRY(T.reserve((size_t)n * h->dims[L]));
    MLP_TRY(hipMemcpyAsync(h->act[0].p, X, sizeof(float) * n * h->dims[0], hipMemcpyHostToDevice, h->st));
    MLP_TRY(hipMemcpyAsync(T.p, Yt, sizeof(float) * n * h->dims[L], hipMemcpyHostToDevice, h->st));
    hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(256), 0, h->st, h->loss.p, 1);
    rc = mlp_forward_batch(h, n);
    if (rc) return rc;
    rc = mlp_backward_batch(h, n, T.p, 1.0f / ((float)n * h->dims[L]));
    if (rc) return rc;
    if (loss) MLP_TRY(hipMemcpyAsync(loss, h->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->st));
    if (grad) MLP_TRY(hipMemcpyAsync(grad, h->grad.p, sizeof(float) * h->nparam, hipMemcpyDeviceToHost, h->st));
    MLP_TRY(hipStreamSynchronize(h->st));
    return MMD_OK;
}

int mmd_mlp_train(mmd_mlp* h, int n, const float* X, const float* Yt, int epochs, int batch, float lr,
                  uint64_t seed, float* epoch_loss)
{
    int rc = mlp_check(h);
    if (rc) return rc;
    if (n < 1 || !X || !Yt || epochs < 0 || batch < 1 || !(lr > 0.f)) return mlp_fail(MMD_ERR_ARG, "invalid training arguments");
    const int L = (int)h->dims.size() - 1;
    const int d0 = h->dims[0], dl = h->dims[L];
    batch = std::min(batch, n);
    MLP_TRY(h->X.reserve((size_t)n * d0));
    MLP_TRY(h->Y.reserve((size_t)n * dl));
    MLP_TRY(h->perm.reserve(n));
    Buf<float> T;
    MLP_TRY(T.reserve((size_t)batch * dl));
    rc = mlp_reserve(h, batch);
    if (rc) return rc;
    MLP_TRY(hipMemcpyAsync(h->X.p, X, sizeof(float) * n * d0, hipMemcpyHostToDevice, h->st));
    MLP_TRY(hipMemcpyAsync(h->Y.p, Yt, sizeof(float) * n * dl, hipMemcpyHostToDevice, h->st));
    std::mt19937_64 rng(seed);
    std::vector<int> perm(n);
    for (int i = 0; i < n; ++i) perm[i] = i;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    for (int ep = 0; ep < epochs; ++ep) {
        std::shuffle(perm.begin(), perm.end(), rng);
        MLP_TRY(hipMemcpyAsync(h->perm.p, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, h->st));
        hipLaunchKernelGGL(mlp_zero, dim3(1), dim3(256), 0, h->st, h->loss.p, 1);
        const int nb = n / batch;   // full batches only
        for (int k = 0; k < nb; ++k) {
            hipLaunchKernelGGL(mlp_gather, dim3(blocks((size_t)batch * d0)), dim3(256), 0, h->st, h->X.p,
                               h->perm.p + (size_t)k * batch, batch, d0, h->act[0].p);
            hipLaunchKernelGGL(mlp_gather, dim3(blocks((size_t)batch * dl)), dim3(256), 0, h->st, h->Y.p,
                               h->perm.p + (size_t)k * batch, batch, dl, T.p);
            rc = mlp_forward_batch(h, batch);
            if (rc) return rc;
            rc = mlp_backward_batch(h, batch, T.p, 1.0f / ((float)batch * dl));
            if (rc) return rc;
            ++h->adam_t;
            const float c1 = 1.f / (1.f - powf(b1, (float)h->adam_t));
            const float c2 = 1.f / (1.f - powf(b2, (float)h->adam_t));
            hipLaunchKernelGGL(mlp_adam, dim3(blocks(h->nparam)), dim3(256), 0, h->st, h->param.p, h->grad.p, h->m.p,
                               h->v.p, (int)h->nparam, lr, b1, b2, eps, c1, c2);
            MLP_TRY(hipGetLastError());
        }
        if (epoch_loss) {
            float l = 0.f;
            MLP_TRY(hipMemcpyAsync(&l, h->loss.p, sizeof(float), hipMemcpyDeviceToHost, h->st));
            MLP_TRY(hipStreamSynchronize(h->st));
            epoch_loss[ep] = nb > 0 ? l / nb : 0.f;
        }
    }
    MLP_TRY(hipStreamSynchronize(h->st));
    return MMD_OK;
}

}  // extern "C"
