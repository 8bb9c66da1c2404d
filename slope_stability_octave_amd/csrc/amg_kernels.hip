// Kernels of the smoothed-aggregation V-cycle: scalar CSR SpMV (coarse
// levels, prolongation, restriction), element-wise scaling and the dense
// coarsest-level solve.
#include "ssr_common.h"

namespace {

// y = alpha*A x + beta*y; G lanes per row, rows with ~10-200 entries.
template <int G>
__global__ __launch_bounds__(256) void k_csr_spmv(int64_t n_rows, const int64_t *__restrict__ rp,
                                                  const int32_t *__restrict__ ci, const double *__restrict__ v,
                                                  const double *__restrict__ x, double *__restrict__ y,
                                                  double alpha, double beta)
{
    const int64_t gid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = gid / G;
    const int l = (int)(gid % G);
    double s = 0.0;
    if (r < n_rows)
        for (int64_t k = rp[r] + l; k < rp[r + 1]; k += G) s += v[k] * x[ci[k]];
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, G);
    if (r < n_rows && l == 0) y[r] = alpha * s + (beta == 0.0 ? 0.0 : beta * y[r]);
}

// z = a .* b
__global__ void k_ewmul(int64_t n, const double *__restrict__ a, const double *__restrict__ b, double *__restrict__ z)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) z[i] = a[i] * b[i];
}

// x += c1 * (d .* r) ; d_out = c2*d_prev + c3*(dinv .* r) ; fused Chebyshev step:
//   dnew = c2*d + c3*(dinv.*r);  x += dnew;  d = dnew
__global__ void k_cheb_step(int64_t n, double c2, double c3, const double *__restrict__ dinv,
                            const double *__restrict__ r, double *__restrict__ d, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double dn = c2 * d[i] + c3 * dinv[i] * r[i];
    d[i] = dn;
    x[i] += dn;
}

// y = M x for a dense n x n row-major matrix (one row per wavefront)
__global__ __launch_bounds__(256) void k_dense_gemv(int n, const double *__restrict__ M, const double *__restrict__ x,
                                                    double *__restrict__ y)
{
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int l = threadIdx.x & 63;
    double s = 0.0;
    if (w < n)
        for (int j = l; j < n; j += 64) s += M[(int64_t)w * n + j] * x[j];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
    if (w < n && l == 0) y[w] = s;
}

}  // namespace

extern "C" {

int ssr_dev_csr_spmv(int64_t n_rows, const int64_t *rp, const int32_t *ci, const double *v, const double *x,
                     double *y, double alpha, double beta)
{
    if (n_rows == 0) return SSR_OK;
    constexpr int G = 8;
    hipLaunchKernelGGL((k_csr_spmv<G>), dim3(ssr_blocks(n_rows * G, 256)), dim3(256), 0, ssr_stream(), n_rows, rp,
                       ci, v, x, y, alpha, beta);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_ewmul(int64_t n, const double *a, const double *b, double *z)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_ewmul, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, a, b, z);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_cheb_step(int64_t n, double c2, double c3, const double *dinv, const double *r, double *d, double *x)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_cheb_step, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, c2, c3, dinv, r, d, x);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_dense_gemv(int n, const double *M, const double *x, double *y)
{
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_dense_gemv, dim3(ssr_blocks((int64_t)n * 64, 256)), dim3(256), 0, ssr_stream(), n, M, x, y);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

}  // extern "C"
