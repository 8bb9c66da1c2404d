// Typed device linear algebra shared by the Krylov solvers and the AMG
// V-cycle (instantiated for double and float; every reduction accumulates in
// double).  All launches go to ssr_stream(); nothing here synchronises.
#pragma once
#include <stdint.h>

namespace ssr {

// y = alpha*K*x + beta*z on the node-block CSR (dim x dim row-major blocks);
// z == nullptr means z = y.  With diag (diagonal block per row) and tblk
// (transposed block per block) given, an operator whose mirrored blocks are
// transposes bit for bit may be read from its blocks on and above the
// diagonal only (SSR_TUNE_SPMV_MIRROR): the same result bits.
template <class T>
int spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const T *val, const T *x,
             T *y, double alpha, double beta, const T *z = nullptr, const int32_t *diag = nullptr,
             const int32_t *tblk = nullptr);
// Y(:, s) = K X(:, s) for s < nv <= 8 (columns at strides ldx, ldy)
template <class T>
int spmm_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const T *val, const T *X,
             int64_t ldx, T *Y, int64_t ldy, int nv);
// G (k x k, column-major) = X' Y for X, Y n x k column-major (ld n); work
// holds gram_work_len(n, k) doubles of per-chunk partial sums, added in a
// fixed order (deterministic)
int64_t gram_work_len(int64_t n, int k);
// (symmetric: X'Y is known to be symmetric, so only the tiles on and above
// the diagonal are formed and the rest mirrored)
int gram(int64_t n, int k, const double *X, const double *Y, double *G, double *work, int64_t work_len,
         int symmetric = 0);
// Y (n x m) = X (n x k) C (k x m), all column-major (ld n, k, n); qstart
// (device, one int per 64-column tile of C, or null) gives the first row of
// C that is nonzero in that tile
// (Y32, if given, receives the same product rounded to float)
int gemm_tall(int64_t n, int k, int m, const double *X, const double *C, double *Y, const int *qstart = nullptr,
              float *Y32 = nullptr);
// y = alpha*A*x + beta*z on scalar CSR (64-bit row pointers); nnz (when
// known) picks the lanes per row
template <class T>
int spmv_csr(int64_t n_rows, const int64_t *rp, const int32_t *ci, const T *v, const T *x, T *y, double alpha,
             double beta, const T *z = nullptr, int64_t nnz = -1);
// y = alpha*A*x + beta*z on block CSR with dense rb x cb blocks (row-major;
// supported shapes: brc_supported); nblk blocks in all picks the lanes per row
template <class T>
int spmv_brc(int rb, int cb, int64_t n_rows, const int32_t *rp, const int32_t *col, const T *val, const T *x, T *y,
             double alpha, double beta, const T *z, int64_t nblk);
bool brc_supported(int rb, int cb);
// result_dev[0] = sum a_i b_i (m_i)
template <class T>
int dot(int64_t n, const T *a, const T *b, const double *mask, double *result_dev);
// h_dev[0:k] = V(:,0:k)' u, k <= 128
template <class T>
int multidot(int64_t n, int k, const T *V, int64_t ld, const T *u, double *h_dev);
// u = alpha*u + beta*V(:,0:k) h_dev, k <= 128
template <class TV, class TU>
int gemv_update(int64_t n, int k, const TV *V, int64_t ld, const double *h_dev, double alpha, double beta, TU *u);
// y = a*x + b*y
template <class T>
int axpby(int64_t n, double a, const T *x, double b, T *y);
// z = a*x + b*y
template <class T>
int lincomb(int64_t n, double a, const T *x, double b, const T *y, T *z);
// x *= m
template <class T>
int mul(int64_t n, const double *m, T *x);
// x = v
template <class T>
int fill(int64_t n, double v, T *x);
// y = a*x (precision conversion)
template <class TI, class TO>
int convert(int64_t n, double a, const TI *x, TO *y);
// y += a*x (mixed precision accumulate)
template <class TI, class TO>
int axpy(int64_t n, double a, const TI *x, TO *y);
// u *= 1/sqrt(nrm2_dev[0]) when sqrt(nrm2) > tiny, else u = 0
template <class T>
int scale_inv_sqrt(int64_t n, const double *nrm2_dev, double tiny, T *u);
// node-block Jacobi
template <class T>
int block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, T *inv);
template <class T>
int block_jacobi_apply(int dim, int64_t n_n, const T *inv, const T *x, const double *mask, T *y);
// Chebyshev steps on D^-1 A: first (x := d = c3 dinv r) or general
// (d = c2 d + c3 dinv r; x += d)
template <class T>
int cheb_first(int64_t n, double c3, const T *dinv, const T *r, T *d, T *x);
template <class T>
int cheb_step(int64_t n, double c2, double c3, const T *dinv, const T *r, T *d, T *x);
// Chebyshev step preconditioned by node-block inverses (dim x dim per node):
// d = c2 d + c3 Binv r; x = d (first) or x += d
template <class T>
int cheb_block_step(int dim, int64_t n_n, double c2, double c3, const T *Binv, const T *r, T *d, T *x, int first);
// y = M x, dense row-major n x n
template <class T>
int dense_gemv(int n, const T *M, const T *x, T *y);

}  // namespace ssr
