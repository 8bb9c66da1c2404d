// Small fully connected network for the learned continuation predictor:
// forward (inference and training), backward and AdamW on gfx950.
//
// Layout: L linear layers, layer l maps n_l inputs to n_{l+1} outputs,
// Z = A_prev W' + b with W row-major (out x in); tanh on hidden layers,
// identity on the output layer.  The master parameters are float; the
// forward pass reads bf16 copies of the weights (round-to-nearest-even of the
// float master, refreshed after every update) and accumulates in float.  The
// backward pass differentiates the computation actually performed (bf16
// weights, float activations); gradients and the optimiser state are float.
//
// The batch is small (hundreds of samples, 128 hidden units), so every
// kernel maps one thread to one output element with a sequential float
// reduction: deterministic, and far from any bandwidth or MFMA limit.
//
// No reference counterpart: the reference continuation uses the secant
// predictor only (SSR_indirect_continuation.m:119), which stays the fallback.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "ssr_common.h"

namespace {

// stream-ordered copy on the library stream, completed before returning
hipError_t ssr_copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind)
{
    hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, ssr_stream());
    if (e != hipSuccess) return e;
    return hipStreamSynchronize(ssr_stream());
}

__host__ __device__ inline uint16_t f2bf(float f)
{
    uint32_t u;
    memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even (finite inputs)
    return (uint16_t)(u >> 16);
}

__host__ __device__ inline float bf2f(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    memcpy(&f, &u, 4);
    return f;
}

__global__ void k_round_bf16(int64_t n, const float *__restrict__ w, uint16_t *__restrict__ wb)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) wb[i] = f2bf(w[i]);
}

// Z[s,o] = b[o] + sum_i X[s,i] W[o,i];  A = tanh(Z) (act) or Z
__global__ void k_linear_fwd(int B, int nin, int nout, const float *__restrict__ X, const uint16_t *__restrict__ Wb,
                             const float *__restrict__ b, float *__restrict__ Z, float *__restrict__ A, int act)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * nout) return;
    const int s = (int)(t / nout), o = (int)(t % nout);
    const float *x = X + (int64_t)s * nin;
    const uint16_t *w = Wb + (int64_t)o * nin;
    float z = b[o];
    for (int i = 0; i < nin; ++i) z += x[i] * bf2f(w[i]);
    if (Z) Z[t] = z;
    A[t] = act ? tanhf(z) : z;
}

// dZ = dA * (1 - A^2) for tanh layers (in place on dA)
__global__ void k_tanh_bwd(int64_t n, const float *__restrict__ A, float *__restrict__ dA)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dA[i] *= 1.0f - A[i] * A[i];
}

// dX[s,i] = sum_o dZ[s,o] W[o,i]
__global__ void k_linear_bwd_data(int B, int nin, int nout, const float *__restrict__ dZ,
                                  const uint16_t *__restrict__ Wb, float *__restrict__ dX)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)B * nin) return;
    const int s = (int)(t / nin), i = (int)(t % nin);
    const float *dz = dZ + (int64_t)s * nout;
    float a = 0.0f;
    for (int o = 0; o < nout; ++o) a += dz[o] * bf2f(Wb[(int64_t)o * nin + i]);
    dX[t] = a;
}

// dW[o,i] = sum_s dZ[s,o] X[s,i];  db[o] = sum_s dZ[s,o] (threads i == nin)
__global__ void k_linear_bwd_param(int B, int nin, int nout, const float *__restrict__ dZ,
                                   const float *__restrict__ X, float *__restrict__ dW, float *__restrict__ db)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)nout * (nin + 1)) return;
    const int o = (int)(t / (nin + 1)), i = (int)(t % (nin + 1));
    float a = 0.0f;
    if (i < nin) {
        for (int s = 0; s < B; ++s) a += dZ[(int64_t)s * nout + o] * X[(int64_t)s * nin + i];
        dW[(int64_t)o * nin + i] = a;
    } else {
        for (int s = 0; s < B; ++s) a += dZ[(int64_t)s * nout + o];
        db[o] = a;
    }
}

// Predictor loss per sample (n_out = nc + 1): coefficients delta = Y[0:nc]
// with L_U = delta' G delta - 2 delta' g + h (G nc x nc row-major), and the
// lambda correction e = Y[nc] - lam_t with L_l = e^2.  loss[s] = L_U + w L_l;
// dY = d(mean loss)/dY.
__global__ void k_predictor_loss(int B, int nc, const float *__restrict__ Y, const float *__restrict__ G,
                                 const float *__restrict__ g, const float *__restrict__ h,
                                 const float *__restrict__ lam_t, float w_lam, float *__restrict__ dY,
                                 float *__restrict__ loss)
{
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= B) return;
    const int no = nc + 1;
    const float *y = Y + (int64_t)s * no;
    const float *Gs = G + (int64_t)s * nc * nc;
    const float *gs = g + (int64_t)s * nc;
    float L = h[s];
    for (int i = 0; i < nc; ++i) {
        float Gd = 0.0f;
        for (int j = 0; j < nc; ++j) Gd += Gs[i * nc + j] * y[j];
        L += y[i] * Gd - 2.0f * y[i] * gs[i];
        if (dY) dY[(int64_t)s * no + i] = 2.0f * (Gd - gs[i]) / (float)B;
    }
    const float e = y[nc] - lam_t[s];
    L += w_lam * e * e;
    if (dY) dY[(int64_t)s * no + nc] = 2.0f * w_lam * e / (float)B;
    loss[s] = L;
}

__global__ void k_adamw(int64_t n, float *__restrict__ w, const float *__restrict__ gr, float *__restrict__ m,
                        float *__restrict__ v, float lr, float b1, float b2, float eps, float bc1, float bc2,
                        float wd)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float g = gr[i];
    const float mi = b1 * m[i] + (1.0f - b1) * g;
    const float vi = b2 * v[i] + (1.0f - b2) * g * g;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * ((mi / bc1) / (sqrtf(vi / bc2) + eps) + wd * w[i]);
}

uint64_t splitmix(uint64_t &s)
{
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

}  // namespace

struct ssr_mlp {
    std::vector<int> width;      // n_layers + 1 widths
    std::vector<int64_t> woff;   // offset of W_l in the parameter vector (b_l follows W_l)
    int64_t n_params = 0;
    float *p = nullptr, *grad = nullptr, *m = nullptr, *v = nullptr;
    uint16_t *pb = nullptr;  // bf16 copy (weights and biases; biases are read as float from p)
    int step = 0;
    int cap_B = 0;
    std::vector<float *> act;  // activations A_0 (= input copy not stored) .. A_L, size B x width
    float *dA = nullptr, *dB = nullptr, *loss = nullptr;
    ~ssr_mlp()
    {
        (void)hipFree(p);
        (void)hipFree(grad);
        (void)hipFree(m);
        (void)hipFree(v);
        (void)hipFree(pb);
        for (auto a : act) (void)hipFree(a);
        (void)hipFree(dA);
        (void)hipFree(dB);
        (void)hipFree(loss);
    }
    int layers() const { return (int)width.size() - 1; }
    int maxw() const
    {
        int w = 0;
        for (int x : width) w = x > w ? x : w;
        return w;
    }
    int ensure(int B)
    {
        if (B <= cap_B) return SSR_OK;
        for (auto a : act) (void)hipFree(a);
        act.assign(layers() + 1, nullptr);
        for (int l = 1; l <= layers(); ++l)
            SSR_HIP_CHECK(hipMalloc((void **)&act[l], sizeof(float) * (size_t)B * width[l]));
        (void)hipFree(dA);
        (void)hipFree(dB);
        (void)hipFree(loss);
        SSR_HIP_CHECK(hipMalloc((void **)&dA, sizeof(float) * (size_t)B * maxw()));
        SSR_HIP_CHECK(hipMalloc((void **)&dB, sizeof(float) * (size_t)B * maxw()));
        SSR_HIP_CHECK(hipMalloc((void **)&loss, sizeof(float) * (size_t)B));
        cap_B = B;
        return SSR_OK;
    }
    int refresh_bf16()
    {
        hipLaunchKernelGGL(k_round_bf16, dim3(ssr_blocks(n_params, 256)), dim3(256), 0, ssr_stream(), n_params, p, pb);
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
    // forward over the batch; activations of every layer kept in act[]
    int forward(int B, const float *X)
    {
        int rc = ensure(B);
        if (rc) return rc;
        const float *in = X;
        for (int l = 0; l < layers(); ++l) {
            const int nin = width[l], nout = width[l + 1];
            const int64_t t = (int64_t)B * nout;
            hipLaunchKernelGGL(k_linear_fwd, dim3(ssr_blocks(t, 256)), dim3(256), 0, ssr_stream(), B, nin, nout, in,
                               pb + woff[l], p + woff[l] + (int64_t)nout * nin, (float *)nullptr, act[l + 1],
                               l + 1 < layers() ? 1 : 0);
            in = act[l + 1];
        }
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
    // gradients of sum_s dY[s]' Y[s] w.r.t. all parameters (after forward(B, X))
    int backward(int B, const float *X, const float *dY)
    {
        float *cur = dA, *nxt = dB;
        SSR_HIP_CHECK(hipMemcpyAsync(cur, dY, sizeof(float) * (size_t)B * width.back(), hipMemcpyDeviceToDevice,
                                     ssr_stream()));
        for (int l = layers() - 1; l >= 0; --l) {
            const int nin = width[l], nout = width[l + 1];
            if (l + 1 < layers()) {
                const int64_t nz = (int64_t)B * nout;
                hipLaunchKernelGGL(k_tanh_bwd, dim3(ssr_blocks(nz, 256)), dim3(256), 0, ssr_stream(), nz, act[l + 1],
                                   cur);
            }
            const float *Xin = l == 0 ? X : act[l];
            const int64_t tp = (int64_t)nout * (nin + 1);
            hipLaunchKernelGGL(k_linear_bwd_param, dim3(ssr_blocks(tp, 256)), dim3(256), 0, ssr_stream(), B, nin, nout,
                               cur, Xin, grad + woff[l], grad + woff[l] + (int64_t)nout * nin);
            if (l > 0) {
                const int64_t td = (int64_t)B * nin;
                hipLaunchKernelGGL(k_linear_bwd_data, dim3(ssr_blocks(td, 256)), dim3(256), 0, ssr_stream(), B, nin,
                                   nout, cur, pb + woff[l], nxt);
                float *tmp = cur;
                cur = nxt;
                nxt = tmp;
            }
        }
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
};

extern "C" {

// Network with widths {n_in, n_hidden x (n_layers-1), n_out}; Xavier-uniform
// weights from `seed`, zero biases, the output layer scaled by out_scale (a
// small value starts the predictor at its baseline).
int ssr_mlp_create(int n_in, int n_hidden, int n_out, int n_layers, uint64_t seed, double out_scale, ssr_mlp **out)
{
    SSR_REQUIRE(out && n_in > 0 && n_hidden > 0 && n_out > 0 && n_layers >= 1 && n_layers <= 16,
                "mlp_create: bad sizes");
    SSR_REQUIRE(n_in <= 4096 && n_hidden <= 4096 && n_out <= 4096, "mlp_create: width above 4096");
    ssr_mlp *m = new ssr_mlp();
    m->width.push_back(n_in);
    for (int l = 0; l + 1 < n_layers; ++l) m->width.push_back(n_hidden);
    m->width.push_back(n_out);
    int64_t off = 0;
    for (int l = 0; l < m->layers(); ++l) {
        m->woff.push_back(off);
        off += (int64_t)m->width[l + 1] * m->width[l] + m->width[l + 1];
    }
    m->n_params = off;
    std::vector<float> h((size_t)off, 0.0f);
    uint64_t st = seed;
    for (int l = 0; l < m->layers(); ++l) {
        const int nin = m->width[l], nout = m->width[l + 1];
        const double a = sqrt(6.0 / (nin + nout)) * (l + 1 == m->layers() ? out_scale : 1.0);
        for (int64_t k = 0; k < (int64_t)nout * nin; ++k) {
            const double u = (double)(splitmix(st) >> 11) * (1.0 / 9007199254740992.0);
            h[m->woff[l] + k] = (float)((2.0 * u - 1.0) * a);
        }
    }
    auto fail = [&](hipError_t e) {
        ssr_set_error("mlp_create: %s", hipGetErrorString(e));
        delete m;
        return SSR_ERR_HIP;
    };
    hipError_t e;
    if ((e = hipMalloc((void **)&m->p, sizeof(float) * off)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->grad, sizeof(float) * off)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->m, sizeof(float) * off)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->v, sizeof(float) * off)) != hipSuccess) return fail(e);
    if ((e = hipMalloc((void **)&m->pb, sizeof(uint16_t) * off)) != hipSuccess) return fail(e);
    if ((e = ssr_copy_sync(m->p, h.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    std::vector<float> z((size_t)off, 0.0f);
    if ((e = ssr_copy_sync(m->m, z.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = ssr_copy_sync(m->v, z.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    if ((e = ssr_copy_sync(m->grad, z.data(), sizeof(float) * off, hipMemcpyHostToDevice)) != hipSuccess) return fail(e);
    int rc = m->refresh_bf16();
    if (rc) {
        delete m;
        return rc;
    }
    *out = m;
    return SSR_OK;
}

int ssr_mlp_destroy(ssr_mlp *m)
{
    delete m;
    return SSR_OK;
}

// info[0] = n_params, info[1] = n_layers, info[2..] = widths (n_layers + 1)
int ssr_mlp_info(const ssr_mlp *m, int64_t *info, int cap)
{
    SSR_REQUIRE(m && info && cap >= 2 + (int)m->width.size(), "mlp_info: buffer too small");
    info[0] = m->n_params;
    info[1] = m->layers();
    for (size_t i = 0; i < m->width.size(); ++i) info[2 + i] = m->width[i];
    return SSR_OK;
}

// Float master parameters, per layer W (out x in, row-major) then b (host).
int ssr_mlp_get_params(const ssr_mlp *m, float *host)
{
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    SSR_HIP_CHECK(ssr_copy_sync(host, m->p, sizeof(float) * m->n_params, hipMemcpyDeviceToHost));
    return SSR_OK;
}

int ssr_mlp_set_params(ssr_mlp *m, const float *host)
{
    SSR_HIP_CHECK(ssr_copy_sync(m->p, host, sizeof(float) * m->n_params, hipMemcpyHostToDevice));
    m->step = 0;
    std::vector<float> z((size_t)m->n_params, 0.0f);
    SSR_HIP_CHECK(ssr_copy_sync(m->m, z.data(), sizeof(float) * m->n_params, hipMemcpyHostToDevice));
    SSR_HIP_CHECK(ssr_copy_sync(m->v, z.data(), sizeof(float) * m->n_params, hipMemcpyHostToDevice));
    return m->refresh_bf16();
}

// Y (device, B x n_out) = net(X) (device, B x n_in), bf16 weights.
int ssr_mlp_forward(ssr_mlp *m, int B, const float *X, float *Y)
{
    SSR_REQUIRE(m && X && Y && B > 0, "mlp_forward: bad arguments");
    int rc = m->forward(B, X);
    if (rc) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(Y, m->act.back(), sizeof(float) * (size_t)B * m->width.back(),
                                 hipMemcpyDeviceToDevice, ssr_stream()));
    return SSR_OK;
}

// grad (device, n_params) of sum_s dY[s]'Y[s] at X (runs the forward pass).
int ssr_mlp_backward(ssr_mlp *m, int B, const float *X, const float *dY, float *grad)
{
    SSR_REQUIRE(m && X && dY && grad && B > 0, "mlp_backward: bad arguments");
    int rc = m->forward(B, X);
    if (rc || (rc = m->backward(B, X, dY))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(grad, m->grad, sizeof(float) * m->n_params, hipMemcpyDeviceToDevice, ssr_stream()));
    return SSR_OK;
}

// Predictor loss and its gradient w.r.t. the network output (see k_predictor_loss).
int ssr_mlp_predictor_loss(int B, int nc, const float *Y, const float *G, const float *g, const float *h,
                           const float *lam_t, float w_lam, float *dY, float *loss)
{
    SSR_REQUIRE(B > 0 && nc > 0 && Y && G && g && h && lam_t && loss, "mlp_predictor_loss: bad arguments");
    hipLaunchKernelGGL(k_predictor_loss, dim3(ssr_blocks(B, 256)), dim3(256), 0, ssr_stream(), B, nc, Y, G, g, h,
                       lam_t, w_lam, dY, loss);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// One AdamW step on the mean predictor loss over the batch; mean loss
// (before the step) to *loss_host.
int ssr_mlp_train_step(ssr_mlp *m, int B, const float *X, const float *G, const float *g, const float *h,
                       const float *lam_t, float w_lam, float lr, float wd, float *loss_host)
{
    SSR_REQUIRE(m && B > 0 && X && G && g && h && lam_t, "mlp_train_step: bad arguments");
    const int nc = m->width.back() - 1;
    SSR_REQUIRE(nc >= 1, "mlp_train_step: the output must hold nc coefficients and one lambda correction");
    int rc = m->forward(B, X);
    if (rc) return rc;
    float *dY = m->dB;  // B x n_out, consumed by backward's first copy
    hipLaunchKernelGGL(k_predictor_loss, dim3(ssr_blocks(B, 256)), dim3(256), 0, ssr_stream(), B, nc,
                       (const float *)m->act.back(), G, g, h, lam_t, w_lam, dY, m->loss);
    SSR_HIP_CHECK(hipGetLastError());
    // backward copies dY into dA first, then ping-pongs dA/dB
    if ((rc = m->backward(B, X, dY))) return rc;
    m->step++;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.0f - powf(b1, (float)m->step), bc2 = 1.0f - powf(b2, (float)m->step);
    hipLaunchKernelGGL(k_adamw, dim3(ssr_blocks(m->n_params, 256)), dim3(256), 0, ssr_stream(), m->n_params, m->p,
                       m->grad, m->m, m->v, lr, b1, b2, 1e-8f, bc1, bc2, wd);
    SSR_HIP_CHECK(hipGetLastError());
    if ((rc = m->refresh_bf16())) return rc;
    if (loss_host) {
        std::vector<float> l((size_t)B);
        SSR_HIP_CHECK(hipMemcpyAsync(l.data(), m->loss, sizeof(float) * B, hipMemcpyDeviceToHost, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        double s = 0.0;
        for (float x : l) s += x;
        *loss_host = (float)(s / B);
    }
    return SSR_OK;
}

}  // extern "C"
