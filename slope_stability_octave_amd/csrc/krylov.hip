// Sparse and dense linear algebra for the Krylov solvers on gfx950:
//   node-block CSR SpMV (dim x dim blocks), masked BLAS-1, tall-skinny
//   multi-dot V'u and update u += V h, node-block Jacobi preconditioner.
//
// Reductions are two-stage with a fixed grid (deterministic: no atomics), a
// 64-lane wavefront reduction by cross-lane shuffles then LDS across waves.
#include "ssr_common.h"

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedBlocks = 1024;

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Block-wide sum of kRedThreads values; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sh)
{
    v = wave_sum(v);
    const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
    if (l == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kRedThreads / 64; ++k) s += sh[k];
    __syncthreads();
    return s;
}

// ---------------------------------------------------------------------------
// SpMV y = K x on node-block CSR.  A group of G lanes handles one block row:
// lanes stride over the row's blocks, each forming dim partial outputs, then
// the group reduces with shuffles.  Block values are row-major dim x dim.
// ---------------------------------------------------------------------------
template <int DIM, int G>
__global__ __launch_bounds__(256) void k_spmv_bsr(int64_t n_rows, const int32_t *__restrict__ row_ptr,
                                                  const int32_t *__restrict__ col,
                                                  const double *__restrict__ val,
                                                  const double *__restrict__ x, double *__restrict__ y,
                                                  double alpha, double beta)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    double acc[DIM];
#pragma unroll
    for (int i = 0; i < DIM; ++i) acc[i] = 0.0;
    if (r < n_rows) {
        const int32_t b0 = row_ptr[r], b1 = row_ptr[r + 1];
        for (int32_t b = b0 + l; b < b1; b += G) {
            const int64_t c = col[b];
            double xv[DIM];
#pragma unroll
            for (int j = 0; j < DIM; ++j) xv[j] = x[c * DIM + j];
            const double *v = val + (int64_t)b * DIM * DIM;
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) acc[i] += v[i * DIM + j] * xv[j];
        }
    }
#pragma unroll
    for (int i = 0; i < DIM; ++i)
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) acc[i] += __shfl_xor(acc[i], o, G);
    if (r < n_rows && l == 0) {
#pragma unroll
        for (int i = 0; i < DIM; ++i) {
            const double o = alpha * acc[i];
            y[r * DIM + i] = beta == 0.0 ? o : o + beta * y[r * DIM + i];
        }
    }
}

// partial[b] = sum over the block's share of a[i]*b[i] (times mask if given)
__global__ __launch_bounds__(kRedThreads) void k_dot_partial(int64_t n, const double *__restrict__ a,
                                                             const double *__restrict__ b,
                                                             const double *__restrict__ m,
                                                             double *__restrict__ part)
{
    __shared__ double sh[kRedThreads / 64];
    double s = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const double t = a[i] * b[i];
        s += m ? m[i] * t : t;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out[c] = sum_b part[c*nb + b] for c < k (one block per output)
__global__ __launch_bounds__(kRedThreads) void k_reduce_partials(int nb, int k, const double *__restrict__ part,
                                                                 double *__restrict__ out)
{
    __shared__ double sh[kRedThreads / 64];
    const int c = blockIdx.x;
    if (c >= k) return;
    double s = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) s += part[(int64_t)c * nb + i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) out[c] = s;
}

// Multi-dot h = V(:,0:k)' u, columns stride ld.  Columns in chunks of 8.
__global__ __launch_bounds__(kRedThreads) void k_multidot_partial(int64_t n, int k, const double *__restrict__ V,
                                                                  int64_t ld, const double *__restrict__ u,
                                                                  double *__restrict__ part)
{
    __shared__ double sh[kRedThreads / 64];
    for (int c0 = 0; c0 < k; c0 += 8) {
        double s[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        const int nc = k - c0 < 8 ? k - c0 : 8;
        for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
            const double ui = u[i];
#pragma unroll
            for (int c = 0; c < 8; ++c)
                if (c < nc) s[c] += V[(int64_t)(c0 + c) * ld + i] * ui;
        }
        for (int c = 0; c < nc; ++c) {
            const double t = block_sum(s[c], sh);
            if (threadIdx.x == 0) part[(int64_t)(c0 + c) * gridDim.x + blockIdx.x] = t;
        }
    }
}

// u = alpha*u + beta * V(:,0:k) h   (h on device)
__global__ __launch_bounds__(256) void k_gemv_update(int64_t n, int k, const double *__restrict__ V, int64_t ld,
                                                     const double *__restrict__ h, double alpha, double beta,
                                                     double *__restrict__ u)
{
    __shared__ double sh[128];
    for (int c = threadIdx.x; c < k && c < 128; c += blockDim.x) sh[c] = h[c];
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int c = 0; c < k; ++c) s += V[(int64_t)c * ld + i] * (c < 128 ? sh[c] : h[c]);
    u[i] = alpha * u[i] + beta * s;
}

// y = a*x + b*y
__global__ void k_axpby(int64_t n, double a, const double *__restrict__ x, double b, double *__restrict__ y)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + (b == 0.0 ? 0.0 : b * y[i]);
}

// z = a*x + b*y (z may alias neither or either)
__global__ void k_lincomb(int64_t n, double a, const double *__restrict__ x, double b, const double *__restrict__ y,
                          double *__restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}

// x = v (write-only; also the graph-safe way to clear a buffer)
__global__ void k_fill(int64_t n, double v, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

__global__ void k_mul(int64_t n, const double *__restrict__ m, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] *= m[i];
}

// Node-block Jacobi: invert each diagonal block (masked rows/cols give identity).
template <int DIM>
__global__ void k_block_inverse(int64_t n_n, const int32_t *__restrict__ diag, const double *__restrict__ val,
                                double *__restrict__ inv)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_n) return;
    const double *a = val + (int64_t)diag[r] * DIM * DIM;
    double *o = inv + r * DIM * DIM;
    if (DIM == 2) {
        const double det = a[0] * a[3] - a[1] * a[2];
        o[0] = a[3] / det; o[1] = -a[1] / det; o[2] = -a[2] / det; o[3] = a[0] / det;
    } else {
        const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
        const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
        const double id = 1.0 / det;
        o[0] = c00 * id; o[1] = (a[2] * a[7] - a[1] * a[8]) * id; o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        o[3] = c01 * id; o[4] = (a[0] * a[8] - a[2] * a[6]) * id; o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        o[6] = c02 * id; o[7] = (a[1] * a[6] - a[0] * a[7]) * id; o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    }
}

template <int DIM>
__global__ void k_block_apply(int64_t n_n, const double *__restrict__ inv, const double *__restrict__ x,
                              const double *__restrict__ mask, double *__restrict__ y)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n_n) return;
    const double *o = inv + r * DIM * DIM;
    double xv[DIM];
#pragma unroll
    for (int j = 0; j < DIM; ++j) xv[j] = x[r * DIM + j];
#pragma unroll
    for (int i = 0; i < DIM; ++i) {
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) s += o[i * DIM + j] * xv[j];
        y[r * DIM + i] = mask ? mask[r * DIM + i] * s : s;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI (device pointers)
// ---------------------------------------------------------------------------
static double *g_part = nullptr;   // kRedBlocks * 128 partial sums
static double *g_red = nullptr;    // 128 reduced values

static int ensure_scratch()
{
    if (!g_part) SSR_HIP_CHECK(hipMalloc((void **)&g_part, sizeof(double) * kRedBlocks * 128));
    if (!g_red) SSR_HIP_CHECK(hipMalloc((void **)&g_red, sizeof(double) * 128));
    return SSR_OK;
}

static unsigned red_blocks(int64_t n)
{
    int64_t b = (n + kRedThreads - 1) / kRedThreads;
    return (unsigned)(b < 1 ? 1 : (b > kRedBlocks ? kRedBlocks : b));
}

extern "C" {

int ssr_dev_spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const double *val,
                     const double *x, double *y, double alpha, double beta)
{
    if (n_rows == 0) return SSR_OK;
    const int G = ssr_tuning(SSR_TUNE_SPMV_GROUP);
    const int64_t threads = n_rows * G;
    const dim3 grid(ssr_blocks(threads, 256)), blk(256);
#define SPMV(D_, G_) hipLaunchKernelGGL((k_spmv_bsr<D_, G_>), grid, blk, 0, ssr_stream(), n_rows, row_ptr, col, val, \
                                        x, y, alpha, beta)
    if (dim == 3) {
        if (G == 4) SPMV(3, 4);
        else if (G == 16) SPMV(3, 16);
        else SPMV(3, 8);
    } else {
        if (G == 4) SPMV(2, 4);
        else if (G == 16) SPMV(2, 16);
        else SPMV(2, 8);
    }
#undef SPMV
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// result[0] = sum_i a_i b_i (m_i) written to DEVICE memory
int ssr_dev_dot(int64_t n, const double *a, const double *b, const double *mask, double *result)
{
    int rc = ensure_scratch();
    if (rc) return rc;
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_dot_partial, dim3(nb), dim3(kRedThreads), 0, ssr_stream(), n, a, b, mask, g_part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kRedThreads), 0, ssr_stream(), (int)nb, 1, g_part, result);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// host-returning dot (synchronises)
int ssr_dot(int64_t n, const double *a, const double *b, const double *mask, double *host_result)
{
    int rc = ensure_scratch();
    if (rc) return rc;
    if ((rc = ssr_dev_dot(n, a, b, mask, g_red))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(host_result, g_red, sizeof(double), hipMemcpyDeviceToHost, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    return SSR_OK;
}

// h (device, k values) = V(:,0:k)' u
int ssr_dev_multidot(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h)
{
    SSR_REQUIRE(k >= 0 && k <= 128, "multidot: k must be in [0,128]");
    if (k == 0) return SSR_OK;
    int rc = ensure_scratch();
    if (rc) return rc;
    const unsigned nb = red_blocks(n);
    hipLaunchKernelGGL(k_multidot_partial, dim3(nb), dim3(kRedThreads), 0, ssr_stream(), n, k, V, ld, u, g_part);
    hipLaunchKernelGGL(k_reduce_partials, dim3(k), dim3(kRedThreads), 0, ssr_stream(), (int)nb, k, g_part, h);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_gemv_update(int64_t n, int k, const double *V, int64_t ld, const double *h, double alpha, double beta,
                        double *u)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_gemv_update, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, k, V, ld, h, alpha,
                       beta, u);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_axpby(int64_t n, double a, const double *x, double b, double *y)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_axpby, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, a, x, b, y);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_lincomb(int64_t n, double a, const double *x, double b, const double *y, double *z)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_lincomb, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, a, x, b, y, z);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_fill(int64_t n, double v, double *x)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_fill, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, v, x);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_mul(int64_t n, const double *m, double *x)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_mul, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, m, x);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, double *inv)
{
    if (n_n == 0) return SSR_OK;
    if (dim == 3)
        hipLaunchKernelGGL(k_block_inverse<3>, dim3(ssr_blocks(n_n, 256)), dim3(256), 0, ssr_stream(), n_n, diag, val, inv);
    else
        hipLaunchKernelGGL(k_block_inverse<2>, dim3(ssr_blocks(n_n, 256)), dim3(256), 0, ssr_stream(), n_n, diag, val, inv);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_block_jacobi_apply(int dim, int64_t n_n, const double *inv, const double *x, const double *mask,
                               double *y)
{
    if (n_n == 0) return SSR_OK;
    if (dim == 3)
        hipLaunchKernelGGL(k_block_apply<3>, dim3(ssr_blocks(n_n, 256)), dim3(256), 0, ssr_stream(), n_n, inv, x, mask, y);
    else
        hipLaunchKernelGGL(k_block_apply<2>, dim3(ssr_blocks(n_n, 256)), dim3(256), 0, ssr_stream(), n_n, inv, x, mask, y);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

}  // extern "C"
