// Deflated Krylov solvers driven from the host over device-resident data.
//
// Reference behaviour:
//   DFGMRES.solve / dfgmres_solver.m:1-181  deflated flexible GMRES, no restart,
//       x0 = W W'b, projector P(x) = x - W W'A x, relative-residual stop
//   dcg_solver.m:1-87                        deflated preconditioned CG
//   A_orthogonalize.m:1-58                   A-orthonormalisation of the basis
//                                            in reverse order, drop |w'Aw| <= tol
//   DFGMRES.expand_deflation_basis / copy    (DFGMRES.m:163-221)
//
// Differences kept deliberately (same mathematics):
//   * Arnoldi uses classical Gram-Schmidt with one re-orthogonalisation
//     (two fused multi-dot/update passes) instead of modified Gram-Schmidt
//     (one reduction per basis vector); the Hessenberg least-squares problem
//     is updated with Givens rotations instead of a fresh QR solve, giving the
//     same iterate and residual estimate.
//   * The Arnoldi coefficients and the new basis vector's norm stay on the
//     device (the vector is normalised there); the host receives each
//     Hessenberg column asynchronously and processes it one iteration behind
//     the device, so no iteration waits for a host round trip.  When column j
//     meets the tolerance the already-enqueued iteration j+1 is discarded and
//     the solution is formed from j+1 columns: the iterate is the one the
//     sequential algorithm returns.
//   * A happy breakdown (H(j+1,j) < 1e-14) returns the exact Krylov solution
//     with j columns (dfgmres_solver.m:136-138 would reuse y of step j-1).
//   * Precision 32: the Arnoldi process, the preconditioner and the deflation
//     projector run in float on float copies of the operator and the basis;
//     x0, the true residual r = b - A x and the accumulation x += Z y stay in
//     double, and the float solve is repeated on the new double residual
//     (iterative refinement) until the relative residual meets the tolerance.
// Vectors span all dim*n_n DOFs; Dirichlet entries are kept at zero by the
// mask after each preconditioner application, so every inner product equals
// the reference's inner product over the free DOFs Q.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "linalg.h"
#include "ssr_common.h"

struct ssr_pattern;
const int32_t *ssr_pattern_d_row_ptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_col(const ssr_pattern *p);
const int32_t *ssr_pattern_d_diag(const ssr_pattern *p);
const int32_t *ssr_pattern_d_tblk(const ssr_pattern *p);
const double *ssr_pattern_d_mask(const ssr_pattern *p);
void ssr_pattern_sizes(const ssr_pattern *p, int *dim, int64_t *n_n, int64_t *n_p, int64_t *n_e,
                       int64_t *nnzb, int64_t *nnz_free);

extern "C" int ssr_dot(int64_t n, const double *a, const double *b, const double *mask, double *host_result);

template <class T>
struct DevBuf {
    T *p = nullptr;
    int64_t n = 0;
    ~DevBuf() { ssr_pool_free(p); }
    int alloc(int64_t len)
    {
        if (p && n >= len) return SSR_OK;
        ssr_pool_free(p);
        p = nullptr;
        n = 0;
        const int rc = ssr_pool_malloc((void **)&p, sizeof(T) * (len > 0 ? len : 1));
        if (rc) return rc;
        n = len;
        return SSR_OK;
    }
};

template <class T>
struct HostBuf {  // page-locked host memory (asynchronous device-to-host copies)
    T *p = nullptr;
    int64_t n = 0;
    ~HostBuf() { (void)hipHostFree(p); }
    int alloc(int64_t len)
    {
        if (p && n >= len) return SSR_OK;
        (void)hipHostFree(p);
        p = nullptr;
        n = len;
        SSR_HIP_CHECK(hipHostMalloc((void **)&p, sizeof(T) * (len > 0 ? len : 1), hipHostMallocDefault));
        return SSR_OK;
    }
};

struct BlockJacobi : ssr_precond {
    int dim;
    int64_t n_n;
    const int32_t *diag;
    const double *mask;
    DevBuf<double> inv;
    DevBuf<float> inv32;
    BlockJacobi(const ssr_pattern *P)
    {
        int64_t n_p, n_e, nnzb, nnzf;
        ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
        diag = ssr_pattern_d_diag(P);
        mask = ssr_pattern_d_mask(P);
    }
    int setup(const double *K, const float *K32) override
    {
        int rc;
        if ((rc = inv.alloc(n_n * dim * dim))) return rc;
        if ((rc = ssr::block_jacobi_setup<double>(dim, n_n, diag, K, inv.p))) return rc;
        if (K32) {
            if ((rc = inv32.alloc(n_n * dim * dim))) return rc;
            if ((rc = ssr::block_jacobi_setup<float>(dim, n_n, diag, K, inv32.p))) return rc;
        }
        return SSR_OK;
    }
    int apply(const double *x, double *y) override { return ssr::block_jacobi_apply<double>(dim, n_n, inv.p, x, mask, y); }
    int apply(const float *x, float *y) override
    {
        if (!inv32.p) {
            ssr_set_error("block Jacobi: float apply before a mixed-precision setup");
            return SSR_ERR_STATE;
        }
        return ssr::block_jacobi_apply<float>(dim, n_n, inv32.p, x, mask, y);
    }
};

struct ssr_solver {
    const ssr_pattern *P = nullptr;
    int dim = 0;
    int64_t n_n = 0, n = 0, nnzb = 0;
    const int32_t *row_ptr = nullptr, *col = nullptr;
    const double *mask = nullptr;
    const double *K = nullptr;  // operator values (borrowed)
    int maxit = 100;
    double tol = 1e-1, defl_tol = 1e-3;
    int precision = 64;    // 64: all double; 32: float Krylov/preconditioner + double refinement
    int max_refine = 8;    // float solves per call in precision 32
    std::shared_ptr<ssr_precond> Msp;  // may be shared with copies of this solver
    ssr_precond *M = nullptr;          // = Msp.get()
    // deflation basis (columns of length n) and its float mirror
    DevBuf<double> W;
    int kW = 0, capW = 0;
    DevBuf<float> W32;
    int kW32 = -1;  // columns of W mirrored in W32 (-1: stale)
    DevBuf<float> K32;
    bool K32_valid = false;
    // Krylov storage
    DevBuf<double> V, Z, t1, t2, r;
    DevBuf<float> V32, Z32, t2f;
    DevBuf<double> hcol, hW, ydev, scal;  // Arnoldi columns, deflation coefficients, LSQ solution, scalars
    DevBuf<double> AW, W2, AW2, Gd, Td, gwork;  // re-orthogonalisation: A W, new basis and its A W, Gram, coefficients
    DevBuf<int> Tq;  // first nonzero row of the coefficients per 64-column tile
    // A W of the current basis as left by the re-orthogonalisation (valid
    // while kAW == kW: set_matrix, expand, clear and copy invalidate it).  For
    // a symmetric operator W'A x = (A W)' x, so the projector skips its SpMV.
    int kAW = -1;
    bool symmetric = false;
    DevBuf<float> AW32;
    int kAW32 = -1;
    HostBuf<double> hhost, yhost;
    std::vector<hipEvent_t> ev;
    // stats
    int last_iters = 0, last_refine = 0;
    double last_relres = 0.0;
    int64_t total_iters = 0, n_solves = 0, n_prec = 0, n_spmv = 0;
    ~ssr_solver()
    {
        for (auto e : ev) (void)hipEventDestroy(e);
    }
};

namespace {

template <class T> const T *op_vals(const ssr_solver *s);
template <> const double *op_vals<double>(const ssr_solver *s) { return s->K; }
template <> const float *op_vals<float>(const ssr_solver *s) { return s->K32.p; }
template <class T> const T *defl_basis(const ssr_solver *s);
template <> const double *defl_basis<double>(const ssr_solver *s) { return s->W.p; }
template <> const float *defl_basis<float>(const ssr_solver *s) { return s->W32.p; }

template <class T>
int spmv(ssr_solver *s, const T *x, T *y, double alpha = 1.0, double beta = 0.0, const T *z = nullptr)
{
    s->n_spmv++;
    return ssr::spmv_bsr<T>(s->dim, s->n_n, s->row_ptr, s->col, op_vals<T>(s), x, y, alpha, beta, z,
                            s->symmetric ? ssr_pattern_d_diag(s->P) : nullptr,
                            s->symmetric ? ssr_pattern_d_tblk(s->P) : nullptr);
}

double hnorm(ssr_solver *s, const double *x, int *rc)
{
    double v = 0.0;
    *rc = ssr_dot(s->n, x, x, nullptr, &v);
    return sqrt(v);
}

// u = alpha*u + beta*W (H' x) with the basis of precision T, H = W unless
// given; coefficients in s->hW
template <class T>
int defl_apply(ssr_solver *s, const T *x, double alpha, double beta, T *u, const T *H = nullptr)
{
    int rc;
    const T *W = defl_basis<T>(s);
    if (!H) H = W;
    for (int c0 = 0; c0 < s->kW; c0 += 128) {
        const int k = s->kW - c0 < 128 ? s->kW - c0 : 128;
        if ((rc = ssr::multidot<T>(s->n, k, H + (int64_t)c0 * s->n, s->n, x, s->hW.p + c0))) return rc;
    }
    for (int c0 = 0; c0 < s->kW; c0 += 128) {
        const int k = s->kW - c0 < 128 ? s->kW - c0 : 128;
        if ((rc = ssr::gemv_update<T, T>(s->n, k, W + (int64_t)c0 * s->n, s->n, s->hW.p + c0, c0 == 0 ? alpha : 1.0,
                                         beta, u)))
            return rc;
    }
    return SSR_OK;
}

template <class T> const T *aw_basis(const ssr_solver *s);
template <> const double *aw_basis<double>(const ssr_solver *s) { return s->kAW == s->kW ? s->AW.p : nullptr; }
template <> const float *aw_basis<float>(const ssr_solver *s)
{
    return s->kAW == s->kW && s->kAW32 == s->kAW ? s->AW32.p : nullptr;
}

// x <- x - W W' A x   (scratch t); with a symmetric operator and A W at hand,
// W'A x = (A W)' x needs no SpMV
template <class T>
int project(ssr_solver *s, T *x, T *t)
{
    if (s->kW == 0) return SSR_OK;
    const T *AWb = s->symmetric ? aw_basis<T>(s) : nullptr;
    if (AWb) return defl_apply<T>(s, x, 1.0, -1.0, x, AWb);
    int rc;
    if ((rc = spmv<T>(s, x, t))) return rc;
    return defl_apply<T>(s, t, 1.0, -1.0, x);
}

template <class T>
int precond(ssr_solver *s, const T *x, T *y)
{
    s->n_prec++;
    int rc = s->M->apply(x, y);
    if (rc) return rc;
    return ssr::mul<T>(s->n, s->mask, y);
}

int ensure_w32(ssr_solver *s)
{
    if (s->kW32 == s->kW) return SSR_OK;
    int rc;
    if ((rc = s->W32.alloc(s->n * (int64_t)(s->capW > 0 ? s->capW : 1)))) return rc;
    if (s->kW > 0 && (rc = ssr::convert<double, float>(s->n * (int64_t)s->kW, 1.0, s->W.p, s->W32.p))) return rc;
    s->kW32 = s->kW;
    return SSR_OK;
}

int ensure_aw32(ssr_solver *s)
{
    if (!s->symmetric || s->kAW != s->kW || s->kAW32 == s->kAW) return SSR_OK;
    int rc;
    if ((rc = s->AW32.alloc(s->n * (int64_t)(s->capW > 0 ? s->capW : 1)))) return rc;
    if (s->kAW > 0 && (rc = ssr::convert<double, float>(s->n * (int64_t)s->kAW, 1.0, s->AW.p, s->AW32.p))) return rc;
    s->kAW32 = s->kAW;
    return SSR_OK;
}

inline int64_t col_off(int j) { return (int64_t)j * j + 2 * (int64_t)j; }  // slot j holds 2(j+1)+1 values

// Flexible GMRES on A d = v0 * rnorm0 where V(:,0) = v0 (unit, precision T)
// is already in place.  Runs at most m iterations and stops when the residual
// estimate |g| drops below `target`; then x += Z y (x in double).
template <class T>
int fgmres_core(ssr_solver *s, T *V, T *Z, T *t, double rnorm0, double target, int m, double *x, int *iters,
                double *res_out)
{
    const int64_t n = s->n;
    int rc;
    std::vector<double> H((size_t)(m + 1) * (m > 0 ? m : 1), 0.0), cs(m + 1), sn(m + 1), g(m + 2, 0.0);
    g[0] = rnorm0;
    double res = rnorm0;
    int done_cols = -1;  // number of columns of the final solution once known
    int enq = 0;         // iterations enqueued
    int proc = 0;        // columns processed on the host
    const int lag = 1;
    auto process = [&](int j) -> int {
        SSR_HIP_CHECK(hipEventSynchronize(s->ev[j]));
        const double *h1 = s->hhost.p + col_off(j), *h2 = h1 + (j + 1), *nr2 = h2 + (j + 1);
        double *Hc = H.data() + (size_t)j * (m + 1);
        for (int i = 0; i <= j; ++i) Hc[i] = h1[i] + h2[i];
        const double hn = sqrt(nr2[0] > 0.0 ? nr2[0] : 0.0);
        Hc[j + 1] = hn;
        for (int i = 0; i < j; ++i) {
            const double a = cs[i] * Hc[i] + sn[i] * Hc[i + 1];
            Hc[i + 1] = -sn[i] * Hc[i] + cs[i] * Hc[i + 1];
            Hc[i] = a;
        }
        if (hn < 1e-14) {
            // exact solution within this Krylov space: close the column with
            // the previous rotations only
            cs[j] = 1.0;
            sn[j] = 0.0;
            res = fabs(g[j + 1]);
            done_cols = j + 1;
            return SSR_OK;
        }
        const double den = hypot(Hc[j], Hc[j + 1]);
        cs[j] = Hc[j] / den;
        sn[j] = Hc[j + 1] / den;
        Hc[j] = den;
        Hc[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        res = fabs(g[j + 1]);
        if (res < target || j + 1 == m) done_cols = j + 1;
        return SSR_OK;
    };
    while (done_cols < 0) {
        if (enq < m) {
            const int j = enq;
            T *zj = Z + (int64_t)j * n;
            T *u = V + (int64_t)(j + 1) * n;
            double *slot = s->hcol.p + col_off(j);
            if ((rc = precond<T>(s, V + (int64_t)j * n, zj))) return rc;
            if ((rc = project<T>(s, zj, t))) return rc;
            if ((rc = spmv<T>(s, zj, u))) return rc;
            const int k = j + 1;
            for (int pass = 0; pass < 2; ++pass) {
                double *hp = slot + pass * k;
                for (int c0 = 0; c0 < k; c0 += 128) {
                    const int kk = k - c0 < 128 ? k - c0 : 128;
                    if ((rc = ssr::multidot<T>(n, kk, V + (int64_t)c0 * n, n, u, hp + c0))) return rc;
                }
                for (int c0 = 0; c0 < k; c0 += 128) {
                    const int kk = k - c0 < 128 ? k - c0 : 128;
                    if ((rc = ssr::gemv_update<T, T>(n, kk, V + (int64_t)c0 * n, n, hp + c0, 1.0, -1.0, u))) return rc;
                }
            }
            if ((rc = ssr::dot<T>(n, u, u, nullptr, slot + 2 * k))) return rc;
            if ((rc = ssr::scale_inv_sqrt<T>(n, slot + 2 * k, 1e-14, u))) return rc;
            SSR_HIP_CHECK(hipMemcpyAsync(s->hhost.p + col_off(j), slot, sizeof(double) * (2 * k + 1),
                                         hipMemcpyDeviceToHost, ssr_stream()));
            SSR_HIP_CHECK(hipEventRecord(s->ev[j], ssr_stream()));
            enq++;
        }
        if (proc < enq && (enq - proc > lag || enq == m)) {
            if ((rc = process(proc))) return rc;
            proc++;
        }
    }
    const int it = done_cols;
    // back substitution R y = g
    std::vector<double> y(it, 0.0);
    for (int i = it - 1; i >= 0; --i) {
        double tt = g[i];
        for (int c = i + 1; c < it; ++c) tt -= H[(size_t)c * (m + 1) + i] * y[c];
        const double d = H[(size_t)i * (m + 1) + i];
        y[i] = d != 0.0 ? tt / d : 0.0;
    }
    // iterations enqueued beyond `it` only wrote V/Z columns past it; drain
    // them before reusing the host buffers
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    memcpy(s->yhost.p, y.data(), sizeof(double) * it);
    SSR_HIP_CHECK(hipMemcpyAsync(s->ydev.p, s->yhost.p, sizeof(double) * it, hipMemcpyHostToDevice, ssr_stream()));
    for (int c0 = 0; c0 < it; c0 += 128) {
        const int k = it - c0 < 128 ? it - c0 : 128;
        if ((rc = ssr::gemv_update<T, double>(n, k, Z + (int64_t)c0 * n, n, s->ydev.p + c0, 1.0, 1.0, x))) return rc;
    }
    *iters = it;
    *res_out = res;
    return SSR_OK;
}

int ensure_arnoldi(ssr_solver *s, int m)
{
    int rc;
    const int64_t nh = col_off(m) + 1;
    if ((rc = s->hcol.alloc(nh)) || (rc = s->hhost.alloc(nh)) || (rc = s->ydev.alloc(m + 1)) ||
        (rc = s->yhost.alloc(m + 1)))
        return rc;
    while ((int)s->ev.size() < m) {
        hipEvent_t e;
        SSR_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        s->ev.push_back(e);
    }
    return SSR_OK;
}

}  // namespace

extern "C" {

int ssr_solver_create(const ssr_pattern *P, int maxit, double tol, double defl_tol, int precond_type,
                      ssr_solver **out)
{
    SSR_REQUIRE(P && out, "solver_create: null");
    SSR_REQUIRE(maxit >= 0 && maxit <= 4096, "solver_create: maxit out of range");
    ssr_solver *s = new ssr_solver();
    int64_t n_p, n_e, nnzf;
    ssr_pattern_sizes(P, &s->dim, &s->n_n, &n_p, &n_e, &s->nnzb, &nnzf);
    s->P = P;
    s->n = s->dim * s->n_n;
    s->row_ptr = ssr_pattern_d_row_ptr(P);
    s->col = ssr_pattern_d_col(P);
    s->mask = ssr_pattern_d_mask(P);
    s->maxit = maxit;
    s->tol = tol;
    s->defl_tol = defl_tol;
    if (precond_type == 0) {
        s->Msp.reset(new BlockJacobi(P));
        s->M = s->Msp.get();
    } else {
        delete s;
        ssr_set_error("solver_create: unknown preconditioner %d", precond_type);
        return SSR_ERR_ARG;
    }
    int rc;
    if ((rc = s->t1.alloc(s->n)) || (rc = s->t2.alloc(s->n)) || (rc = s->r.alloc(s->n)) ||
        (rc = s->hW.alloc(8192)) || (rc = s->scal.alloc(8))) {
        delete s;
        return rc;
    }
    *out = s;
    return SSR_OK;
}

int ssr_solver_destroy(ssr_solver *s)
{
    if (!s) return SSR_OK;
    delete s;
    return SSR_OK;
}

// Replace the preconditioner (takes ownership).
int ssr_solver_set_precond(ssr_solver *s, ssr_precond *M)
{
    SSR_REQUIRE(s && M, "solver_set_precond: null");
    s->Msp.reset(M);
    s->M = M;
    return SSR_OK;
}

// The operator is symmetric (K_r is symmetrised by construction): the
// deflation projector then uses (A W)' x for W'A x.
int ssr_solver_set_symmetric(ssr_solver *s, int symmetric)
{
    SSR_REQUIRE(s, "set_symmetric: null");
    s->symmetric = symmetric != 0;
    s->kAW = -1;
    s->kAW32 = -1;
    return SSR_OK;
}

// The next setup of the solver's preconditioner rebuilds it in full.
int ssr_solver_invalidate_preconditioner(ssr_solver *s)
{
    SSR_REQUIRE(s && s->M, "invalidate_preconditioner: null");
    s->M->invalidate();
    return SSR_OK;
}

// dst applies (and sets up) src's preconditioner object from now on; the
// object lives until the last solver using it is destroyed.  Calls on the
// two solvers must not overlap.
int ssr_solver_share_preconditioner(ssr_solver *dst, const ssr_solver *src)
{
    SSR_REQUIRE(dst && src && dst->P == src->P, "share_preconditioner: solvers on different patterns");
    dst->Msp = src->Msp;
    dst->M = dst->Msp.get();
    return SSR_OK;
}

// 64: plain double; 32: float Krylov + preconditioner with double refinement.
int ssr_solver_set_precision(ssr_solver *s, int bits)
{
    SSR_REQUIRE(s && (bits == 64 || bits == 32), "solver_set_precision: bits must be 32 or 64");
    s->precision = bits;
    s->K32_valid = false;
    s->kW32 = -1;
    s->kAW32 = -1;
    return SSR_OK;
}

int ssr_solver_get_precision(const ssr_solver *s, int *bits)
{
    SSR_REQUIRE(s && bits, "solver_get_precision: null");
    *bits = s->precision;
    return SSR_OK;
}

// Operator values (device, node-block CSR of the pattern; borrowed) and
// preconditioner setup on them (setup_preconditioner / update_values).
int ssr_solver_set_matrix(ssr_solver *s, const double *Kvals, int setup_precond)
{
    SSR_REQUIRE(s && Kvals, "solver_set_matrix: null");
    s->K = Kvals;
    s->kAW = -1;  // values may change behind the same pointer
    int rc;
    s->K32_valid = false;
    if (s->precision == 32) {
        const int64_t nv = s->nnzb * s->dim * s->dim;
        if ((rc = s->K32.alloc(nv))) return rc;
        if ((rc = ssr::convert<double, float>(nv, 1.0, Kvals, s->K32.p))) return rc;
        s->K32_valid = true;
    }
    if (setup_precond) {
        s->M->set_symmetric(s->symmetric);
        return s->M->setup(Kvals, s->precision == 32 ? s->K32.p : nullptr);
    }
    return SSR_OK;
}

int ssr_solver_deflation_size(const ssr_solver *s, int *k)
{
    *k = s->kW;
    return SSR_OK;
}

int ssr_solver_clear_deflation(ssr_solver *s)
{
    s->kW = 0;
    s->kW32 = -1;
    s->kAW = -1;
    return SSR_OK;
}

// Append one column (device vector of length dim*n_n; Dirichlet entries are
// zeroed by the mask).
int ssr_solver_expand_deflation(ssr_solver *s, const double *v)
{
    SSR_REQUIRE(s && v, "expand_deflation: null");
    SSR_REQUIRE(s->kW < 8000, "expand_deflation: basis too large");
    if (s->kW == s->capW) {
        const int cap = s->capW ? 2 * s->capW : 16;
        double *p = nullptr;
        int rc = ssr_pool_malloc((void **)&p, sizeof(double) * s->n * cap);
        if (rc) return rc;
        if (s->kW) SSR_HIP_CHECK(hipMemcpyAsync(p, s->W.p, sizeof(double) * s->n * s->kW, hipMemcpyDeviceToDevice, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        ssr_pool_free(s->W.p);
        s->W.p = p;
        s->W.n = s->n * cap;
        s->capW = cap;
    }
    double *dst = s->W.p + (int64_t)s->kW * s->n;
    SSR_HIP_CHECK(hipMemcpyAsync(dst, v, sizeof(double) * s->n, hipMemcpyDeviceToDevice, ssr_stream()));
    int rc = ssr::mul<double>(s->n, s->mask, dst);
    if (rc) return rc;
    s->kW++;
    s->kW32 = -1;
    s->kAW = -1;
    return SSR_OK;
}

// DFGMRES.copy (DFGMRES.m:204-221): a new solver with src's settings and a
// copy of its deflation basis; with share_precond the copy applies (and sets
// up) src's preconditioner object, else it gets a node-block Jacobi of its
// own, as a fresh solver does.
int ssr_solver_copy_deflation(ssr_solver *dst, const ssr_solver *src);

int ssr_solver_clone(const ssr_solver *src, int share_precond, ssr_solver **out)
{
    SSR_REQUIRE(src && out, "solver_clone: null");
    *out = nullptr;
    ssr_solver *d = nullptr;
    int rc = ssr_solver_create(src->P, src->maxit, src->tol, src->defl_tol, 0, &d);
    if (rc) return rc;
    d->precision = src->precision;
    d->max_refine = src->max_refine;
    d->symmetric = src->symmetric;
    if (share_precond) {
        d->Msp = src->Msp;
        d->M = d->Msp.get();
    }
    if ((rc = ssr_solver_copy_deflation(d, src))) {
        delete d;
        return rc;
    }
    *out = d;
    return SSR_OK;
}

// dst's basis := src's basis (DFGMRES.copy keeps the deflation basis).
int ssr_solver_copy_deflation(ssr_solver *dst, const ssr_solver *src)
{
    SSR_REQUIRE(dst && src && dst->n == src->n, "copy_deflation: mismatched solvers");
    dst->kW = 0;
    dst->kW32 = -1;
    dst->kAW = -1;
    if (src->kW == 0) return SSR_OK;
    // one allocation at the source's capacity and one copy (the source's
    // columns are already zero at the Dirichlet DOFs, as expand leaves them)
    if (dst->capW < src->kW) {
        const int cap = src->capW > 16 ? src->capW : 16;
        int rc = dst->W.alloc(dst->n * (int64_t)cap);
        if (rc) return rc;
        dst->capW = cap;
    }
    SSR_HIP_CHECK(hipMemcpyAsync(dst->W.p, src->W.p, sizeof(double) * src->n * src->kW, hipMemcpyDeviceToDevice,
                                 ssr_stream()));
    dst->kW = src->kW;
    return SSR_OK;
}

// Copy column c of the basis to a device vector.
int ssr_solver_get_deflation(const ssr_solver *s, int c, double *out)
{
    SSR_REQUIRE(c >= 0 && c < s->kW, "get_deflation: column out of range");
    SSR_HIP_CHECK(hipMemcpyAsync(out, s->W.p + (int64_t)c * s->n, sizeof(double) * s->n, hipMemcpyDeviceToDevice,
                                 ssr_stream()));
    return SSR_OK;
}

// A-orthonormalise the basis against the current operator (A_orthogonalize.m),
// in double.  The reference runs Gram-Schmidt in the A-inner product over the
// columns in reverse order, one SpMV, one set of inner products and one
// host decision (drop when |w'Aw| <= tol, keep the sign) per column.  Here
// the same process runs on coefficient vectors: AW = A W (multi-vector SpMV,
// K read once per 8 columns), the Gram matrix G = W'AW (one tiled pass, one
// download), the reverse-order decisions on the host (an accepted column is
// W t_c, and (W t_c)'A(W e_j) = t_c'G e_j; t_c is supported on indices >= the
// column it started from), and the new basis W T (one tiled pass).  Same
// mathematics, one host round trip instead of k.
int ssr_solver_A_orthogonalize(ssr_solver *s)
{
    SSR_REQUIRE(s && s->K, "A_orthogonalize: operator not set");
    const int k = s->kW;
    if (k == 0) return SSR_OK;
    const int64_t n = s->n;
    int rc;
    const int64_t wl = ssr::gram_work_len(n, k);
    if ((rc = s->AW.alloc(n * (int64_t)s->capW)) || (rc = s->Gd.alloc((int64_t)k * k)) ||
        (rc = s->gwork.alloc(wl)))
        return rc;
    for (int c0 = 0; c0 < k; c0 += 8) {
        const int nv = k - c0 < 8 ? k - c0 : 8;
        if ((rc = ssr::spmm_bsr<double>(s->dim, s->n_n, s->row_ptr, s->col, s->K, s->W.p + (int64_t)c0 * n, n,
                                        s->AW.p + (int64_t)c0 * n, n, nv)))
            return rc;
        s->n_spmv += nv;
    }
    if ((rc = ssr::gram(n, k, s->W.p, s->AW.p, s->Gd.p, s->gwork.p, wl, s->symmetric ? 1 : 0))) return rc;
    std::vector<double> G((size_t)k * k);  // column j = W'A w_j
    SSR_HIP_CHECK(hipMemcpyAsync(G.data(), s->Gd.p, sizeof(double) * G.size(), hipMemcpyDeviceToHost, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    // accepted coefficient vectors, in acceptance order, as rows of Tm (k
    // wide) with their first nonzero index lo[c]
    std::vector<double> Tm, sgn, u(k), tg;
    std::vector<int> lo;
    Tm.reserve((size_t)k * 16);
    for (int i = 0; i < k; ++i) {
        const int j = k - 1 - i;
        const double *Gj = G.data() + (size_t)j * k;
        const size_t acc = sgn.size();
        // coefficients against every accepted column first (classical form,
        // as the reference subtracts W_orth * (A w)'W_orth in one product)
        tg.resize(acc);
        for (size_t c = 0; c < acc; ++c) {
            const double *t = Tm.data() + c * k;
            double d = 0.0;
            for (int q = lo[c]; q < k; ++q) d += t[q] * Gj[q];
            tg[c] = sgn[c] * d;
        }
        std::fill(u.begin() + j, u.end(), 0.0);
        u[j] = 1.0;
        for (size_t c = 0; c < acc; ++c) {
            const double *t = Tm.data() + c * k;
            const double g = tg[c];
            for (int q = lo[c]; q < k; ++q) u[q] -= g * t[q];
        }
        double nt = 0.0;  // (w - W_orth h)' A w
        for (int q = j; q < k; ++q) nt += u[q] * Gj[q];
        if (fabs(nt) > s->defl_tol) {
            const double sc = 1.0 / sqrt(fabs(nt));
            const size_t off = Tm.size();
            Tm.resize(off + k, 0.0);
            for (int q = j; q < k; ++q) Tm[off + q] = u[q] * sc;
            sgn.push_back(nt > 0 ? 1.0 : -1.0);
            lo.push_back(j);
        }
    }
    const int acc = (int)sgn.size();
    if (acc == 0) {
        s->kW = 0;
        s->kW32 = -1;
        s->kAW = -1;
        return SSR_OK;
    }
    // W_new(:, c) = W * T[acc-1-c]  (accepted columns back in forward order);
    // T is zero above each column's starting row, so each 64-column tile of
    // the product starts at the smallest starting row of its columns
    std::vector<double> Th((size_t)k * acc);
    const int nct = (acc + 63) / 64;
    std::vector<int> qst(nct, k);
    for (int c = 0; c < acc; ++c) {
        std::copy(Tm.begin() + (size_t)(acc - 1 - c) * k, Tm.begin() + (size_t)(acc - c) * k,
                  Th.begin() + (size_t)c * k);
        qst[c / 64] = std::min(qst[c / 64], lo[acc - 1 - c]);
    }
    if ((rc = s->Td.alloc((int64_t)k * acc)) || (rc = s->W2.alloc(n * (int64_t)s->capW)) ||
        (rc = s->Tq.alloc(nct)))
        return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(s->Td.p, Th.data(), sizeof(double) * Th.size(), hipMemcpyHostToDevice, ssr_stream()));
    SSR_HIP_CHECK(hipMemcpyAsync(s->Tq.p, qst.data(), sizeof(int) * qst.size(), hipMemcpyHostToDevice, ssr_stream()));
    // in float mode the same passes also write the float copies the float
    // solves read (the double results rounded: what converting them gives)
    const bool f32 = s->precision == 32;
    const int64_t capn = n * (int64_t)s->capW;
    if (f32 && (rc = s->W32.alloc(capn))) return rc;
    if ((rc = ssr::gemm_tall(n, k, acc, s->W.p, s->Td.p, s->W2.p, s->Tq.p, f32 ? s->W32.p : nullptr))) return rc;
    std::swap(s->W.p, s->W2.p);
    std::swap(s->W.n, s->W2.n);
    if (s->symmetric) {
        // A (W T) = (A W) T: the projector of the solves that follow reads it
        if ((rc = s->AW2.alloc(capn))) return rc;
        if (f32 && (rc = s->AW32.alloc(capn))) return rc;
        if ((rc = ssr::gemm_tall(n, k, acc, s->AW.p, s->Td.p, s->AW2.p, s->Tq.p, f32 ? s->AW32.p : nullptr)))
            return rc;
        std::swap(s->AW.p, s->AW2.p);
        std::swap(s->AW.n, s->AW2.n);
    }
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));  // Th leaves scope
    s->kW = acc;
    s->kW32 = f32 ? acc : -1;
    s->kAW = s->symmetric ? acc : -1;
    s->kAW32 = f32 && s->symmetric ? acc : -1;
    return SSR_OK;
}

// Deflated flexible GMRES (dfgmres_solver.m).  b, x device vectors (length
// dim*n_n).  x is overwritten.
int ssr_solver_dfgmres(ssr_solver *s, const double *b, double *x, int *iters, double *relres)
{
    SSR_REQUIRE(s && s->K && b && x, "dfgmres: operator or vectors missing");
    SSR_REQUIRE(s->precision == 64 || s->K32_valid, "dfgmres: float operator missing (set_matrix after set_precision)");
    const int64_t n = s->n;
    const int m = s->maxit;
    int rc;
    if ((rc = ensure_arnoldi(s, m))) return rc;
    // x0 = W W' b (double)
    if (s->kW > 0) {
        if ((rc = defl_apply<double>(s, b, 0.0, 1.0, x))) return rc;
    } else {
        if ((rc = ssr::fill<double>(n, 0.0, x))) return rc;
    }
    // r0 = b - A x
    if ((rc = spmv<double>(s, x, s->r.p, -1.0, 1.0, b))) return rc;
    double bnorm = hnorm(s, b, &rc);
    if (rc) return rc;
    if (bnorm == 0.0) bnorm = 1.0;
    double res = hnorm(s, s->r.p, &rc);
    if (rc) return rc;
    s->n_solves++;
    int it_total = 0, refine = 0;
    if (res / bnorm < s->tol || m == 0) {
        // converged at x0
    } else if (s->precision == 64) {
        if ((rc = s->V.alloc(n * (int64_t)(m + 1))) || (rc = s->Z.alloc(n * (int64_t)m))) return rc;
        if ((rc = ssr::axpby<double>(n, 1.0 / res, s->r.p, 0.0, s->V.p))) return rc;
        double est = res;
        if ((rc = fgmres_core<double>(s, s->V.p, s->Z.p, s->t2.p, res, s->tol * bnorm, m, x, &it_total, &est))) return rc;
        res = est;
        refine = 1;
    } else {
        if ((rc = s->V32.alloc(n * (int64_t)(m + 1))) || (rc = s->Z32.alloc(n * (int64_t)m)) || (rc = s->t2f.alloc(n)))
            return rc;
        if ((rc = ensure_w32(s)) || (rc = ensure_aw32(s))) return rc;
        while (it_total < m && refine < s->max_refine) {
            if (refine > 0 && s->kW > 0) {
                // The projected Krylov space has no component in span(W), so a
                // restart first takes the coarse correction of its residual,
                // as x0 = W W' b does for the first pass.
                if ((rc = defl_apply<double>(s, s->r.p, 1.0, 1.0, x))) return rc;
                if ((rc = spmv<double>(s, x, s->r.p, -1.0, 1.0, b))) return rc;
                res = hnorm(s, s->r.p, &rc);
                if (rc) return rc;
                if (res / bnorm < s->tol) break;
            }
            if ((rc = ssr::convert<double, float>(n, 1.0 / res, s->r.p, s->V32.p))) return rc;
            // a float Krylov space loses orthogonality well before it could
            // resolve 1e-6 of its start with a weak preconditioner; each pass
            // asks for 1e-4 and deeper reductions come from restarts on the
            // double residual
            const double target = fmax(s->tol * bnorm, 1e-4 * res);
            int it = 0;
            double est = 0.0;
            if ((rc = fgmres_core<float>(s, s->V32.p, s->Z32.p, s->t2f.p, res, target, m - it_total, x, &it, &est)))
                return rc;
            it_total += it;
            refine++;
            // true residual in double
            if ((rc = spmv<double>(s, x, s->r.p, -1.0, 1.0, b))) return rc;
            res = hnorm(s, s->r.p, &rc);
            if (rc) return rc;
            if (res / bnorm < s->tol || it == 0) break;
        }
    }
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    *iters = it_total;
    *relres = res / bnorm;
    s->last_iters = it_total;
    s->last_refine = refine;
    s->last_relres = *relres;
    s->total_iters += it_total;
    return SSR_OK;
}

// Deflated preconditioned CG (dcg_solver.m), double.
int ssr_solver_dcg(ssr_solver *s, const double *b, double *x, int *iters, double *relres)
{
    SSR_REQUIRE(s && s->K && b && x, "dcg: operator or vectors missing");
    const int64_t n = s->n;
    int rc;
    if ((rc = s->V.alloc(4 * n))) return rc;
    double *r = s->V.p, *z = r + n, *p = z + n, *q = p + n;
    if (s->kW > 0) {
        if ((rc = defl_apply<double>(s, b, 0.0, 1.0, x))) return rc;
    } else {
        if ((rc = ssr::fill<double>(n, 0.0, x))) return rc;
    }
    if ((rc = spmv<double>(s, x, r, -1.0, 1.0, b))) return rc;
    const double bnorm = hnorm(s, b, &rc);
    if (rc) return rc;
    double res = hnorm(s, r, &rc) / bnorm;
    if (rc) return rc;
    s->n_solves++;
    if (res < s->tol || s->maxit == 0 || bnorm == 0.0) {
        *iters = 0;
        *relres = res;
        return SSR_OK;
    }
    if ((rc = precond<double>(s, r, z))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, ssr_stream()));
    if ((rc = project<double>(s, p, s->t2.p))) return rc;
    double gamma_old = 0.0;
    if ((rc = ssr_dot(n, r, z, nullptr, &gamma_old))) return rc;
    int j = 0;
    for (j = 1; j <= s->maxit; ++j) {
        if ((rc = spmv<double>(s, p, q))) return rc;
        double sp = 0.0;
        if ((rc = ssr_dot(n, q, p, nullptr, &sp))) return rc;
        const double alpha = gamma_old / sp;
        if ((rc = ssr::axpby<double>(n, alpha, p, 1.0, x))) return rc;
        if ((rc = ssr::axpby<double>(n, -alpha, q, 1.0, r))) return rc;
        res = hnorm(s, r, &rc) / bnorm;
        if (rc) return rc;
        if (res < s->tol) break;
        if ((rc = precond<double>(s, r, z))) return rc;
        double gamma_new = 0.0;
        if ((rc = ssr_dot(n, r, z, nullptr, &gamma_new))) return rc;
        const double beta = gamma_new / gamma_old;
        // p = P(z) + beta p   (q as scratch for P(z))
        SSR_HIP_CHECK(hipMemcpyAsync(q, z, sizeof(double) * n, hipMemcpyDeviceToDevice, ssr_stream()));
        if ((rc = project<double>(s, q, s->t2.p))) return rc;
        if ((rc = ssr::axpby<double>(n, 1.0, q, beta, p))) return rc;
        gamma_old = gamma_new;
    }
    if (j > s->maxit) j = s->maxit;
    *iters = j;
    *relres = res;
    s->last_iters = j;
    s->last_relres = res;
    s->total_iters += j;
    return SSR_OK;
}

int ssr_solver_precond_apply(ssr_solver *s, const double *x, double *y) { return precond<double>(s, x, y); }

const ssr_pattern *ssr_solver_pattern(const ssr_solver *s) { return s->P; }

int ssr_solver_stats(const ssr_solver *s, int64_t *out /* 4 */)
{
    out[0] = s->total_iters;
    out[1] = s->n_solves;
    out[2] = s->n_prec;
    out[3] = s->n_spmv;
    return SSR_OK;
}

}  // extern "C"
