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
//   * A happy breakdown (H(j+1,j) < 1e-14) returns the exact Krylov solution
//     with j columns (dfgmres_solver.m:136-138 would reuse y of step j-1).
// Vectors span all dim*n_n DOFs; Dirichlet entries are kept at zero by the
// mask after each preconditioner application, so every inner product equals
// the reference's inner product over the free DOFs Q.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "ssr_common.h"

struct ssr_pattern;
const int32_t *ssr_pattern_d_row_ptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_col(const ssr_pattern *p);
const int32_t *ssr_pattern_d_diag(const ssr_pattern *p);
const double *ssr_pattern_d_mask(const ssr_pattern *p);
void ssr_pattern_sizes(const ssr_pattern *p, int *dim, int64_t *n_n, int64_t *n_p, int64_t *n_e,
                       int64_t *nnzb, int64_t *nnz_free);

extern "C" {
int ssr_dev_spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col, const double *val,
                     const double *x, double *y, double alpha, double beta);
int ssr_dot(int64_t n, const double *a, const double *b, const double *mask, double *host_result);
int ssr_dev_multidot(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h);
int ssr_dev_gemv_update(int64_t n, int k, const double *V, int64_t ld, const double *h, double alpha,
                        double beta, double *u);
int ssr_dev_axpby(int64_t n, double a, const double *x, double b, double *y);
int ssr_dev_lincomb(int64_t n, double a, const double *x, double b, const double *y, double *z);
int ssr_dev_mul(int64_t n, const double *m, double *x);
int ssr_dev_block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, double *inv);
int ssr_dev_block_jacobi_apply(int dim, int64_t n_n, const double *inv, const double *x, const double *mask,
                               double *y);
}


struct BlockJacobi : ssr_precond {
    int dim;
    int64_t n_n;
    const int32_t *diag;
    const double *mask;
    double *inv = nullptr;
    BlockJacobi(const ssr_pattern *P)
    {
        int64_t n_p, n_e, nnzb, nnzf;
        ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
        diag = ssr_pattern_d_diag(P);
        mask = ssr_pattern_d_mask(P);
    }
    ~BlockJacobi() { (void)hipFree(inv); }
    int setup(const double *K) override
    {
        if (!inv) SSR_HIP_CHECK(hipMalloc((void **)&inv, sizeof(double) * n_n * dim * dim));
        return ssr_dev_block_jacobi_setup(dim, n_n, diag, K, inv);
    }
    int apply(const double *x, double *y) override { return ssr_dev_block_jacobi_apply(dim, n_n, inv, x, mask, y); }
};


struct DevVec {
    double *p = nullptr;
    int64_t n = 0;
    ~DevVec() { (void)hipFree(p); }
    int alloc(int64_t len)
    {
        if (p && n >= len) return SSR_OK;
        (void)hipFree(p);
        p = nullptr;
        n = len;
        SSR_HIP_CHECK(hipMalloc((void **)&p, sizeof(double) * (len > 0 ? len : 1)));
        return SSR_OK;
    }
};

struct ssr_solver {
    const ssr_pattern *P = nullptr;
    int dim = 0;
    int64_t n_n = 0, n = 0;
    const int32_t *row_ptr = nullptr, *col = nullptr;
    const double *mask = nullptr;
    const double *K = nullptr;  // operator values (borrowed)
    int maxit = 100;
    double tol = 1e-1, defl_tol = 1e-3;
    ssr_precond *M = nullptr;
    bool own_M = true;
    // deflation basis (columns of length n)
    DevVec W;
    int kW = 0, capW = 0;
    // Krylov storage
    DevVec V, Z, t1, t2, hdev;
    // stats
    int last_iters = 0;
    double last_relres = 0.0;
    int64_t total_iters = 0, n_solves = 0, n_prec = 0, n_spmv = 0;
};

namespace {

int spmv(ssr_solver *s, const double *x, double *y)
{
    s->n_spmv++;
    return ssr_dev_spmv_bsr(s->dim, s->n_n, s->row_ptr, s->col, s->K, x, y, 1.0, 0.0);
}

double hnorm(ssr_solver *s, const double *x, int *rc)
{
    double v = 0.0;
    *rc = ssr_dot(s->n, x, x, nullptr, &v);
    return sqrt(v);
}

// h (device) = W' x, chunked by 128 columns; result on device at s->hdev
int defl_coef(ssr_solver *s, const double *x)
{
    int rc;
    for (int c0 = 0; c0 < s->kW; c0 += 128) {
        const int k = s->kW - c0 < 128 ? s->kW - c0 : 128;
        if ((rc = ssr_dev_multidot(s->n, k, s->W.p + (int64_t)c0 * s->n, s->n, x, s->hdev.p + c0))) return rc;
    }
    return SSR_OK;
}

// u = alpha*u + beta*W h(dev)
int defl_update(ssr_solver *s, double alpha, double beta, double *u)
{
    int rc;
    for (int c0 = 0; c0 < s->kW; c0 += 128) {
        const int k = s->kW - c0 < 128 ? s->kW - c0 : 128;
        if ((rc = ssr_dev_gemv_update(s->n, k, s->W.p + (int64_t)c0 * s->n, s->n, s->hdev.p + c0,
                                      c0 == 0 ? alpha : 1.0, beta, u)))
            return rc;
    }
    return SSR_OK;
}

// x <- x - W W' A x   (t2 used as scratch)
int project(ssr_solver *s, double *x)
{
    if (s->kW == 0) return SSR_OK;
    int rc;
    if ((rc = spmv(s, x, s->t2.p))) return rc;
    if ((rc = defl_coef(s, s->t2.p))) return rc;
    return defl_update(s, 1.0, -1.0, x);
}

int precond(ssr_solver *s, const double *x, double *y)
{
    s->n_prec++;
    int rc = s->M->apply(x, y);
    if (rc) return rc;
    return ssr_dev_mul(s->n, s->mask, y);
}

}  // namespace

extern "C" {

int ssr_solver_create(const ssr_pattern *P, int maxit, double tol, double defl_tol, int precond_type,
                      ssr_solver **out)
{
    SSR_REQUIRE(P && out, "solver_create: null");
    SSR_REQUIRE(maxit >= 0 && maxit <= 4096, "solver_create: maxit out of range");
    ssr_solver *s = new ssr_solver();
    int64_t n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &s->dim, &s->n_n, &n_p, &n_e, &nnzb, &nnzf);
    s->P = P;
    s->n = s->dim * s->n_n;
    s->row_ptr = ssr_pattern_d_row_ptr(P);
    s->col = ssr_pattern_d_col(P);
    s->mask = ssr_pattern_d_mask(P);
    s->maxit = maxit;
    s->tol = tol;
    s->defl_tol = defl_tol;
    if (precond_type == 0) {
        s->M = new BlockJacobi(P);
    } else {
        delete s;
        ssr_set_error("solver_create: unknown preconditioner %d", precond_type);
        return SSR_ERR_ARG;
    }
    int rc;
    if ((rc = s->t1.alloc(s->n)) || (rc = s->t2.alloc(s->n)) || (rc = s->hdev.alloc(8192))) {
        delete s;
        return rc;
    }
    *out = s;
    return SSR_OK;
}

int ssr_solver_destroy(ssr_solver *s)
{
    if (!s) return SSR_OK;
    if (s->own_M) delete s->M;
    delete s;
    return SSR_OK;
}

// Replace the preconditioner (takes ownership).
int ssr_solver_set_precond(ssr_solver *s, ssr_precond *M)
{
    SSR_REQUIRE(s && M, "solver_set_precond: null");
    if (s->own_M) delete s->M;
    s->M = M;
    s->own_M = true;
    return SSR_OK;
}

// Operator values (device, node-block CSR of the pattern; borrowed) and
// preconditioner setup on them (setup_preconditioner / update_values).
int ssr_solver_set_matrix(ssr_solver *s, const double *Kvals, int setup_precond)
{
    SSR_REQUIRE(s && Kvals, "solver_set_matrix: null");
    s->K = Kvals;
    if (setup_precond) return s->M->setup(Kvals);
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
        SSR_HIP_CHECK(hipMalloc((void **)&p, sizeof(double) * s->n * cap));
        if (s->kW) SSR_HIP_CHECK(hipMemcpyAsync(p, s->W.p, sizeof(double) * s->n * s->kW, hipMemcpyDeviceToDevice, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        (void)hipFree(s->W.p);
        s->W.p = p;
        s->W.n = s->n * cap;
        s->capW = cap;
    }
    double *dst = s->W.p + (int64_t)s->kW * s->n;
    SSR_HIP_CHECK(hipMemcpyAsync(dst, v, sizeof(double) * s->n, hipMemcpyDeviceToDevice, ssr_stream()));
    int rc = ssr_dev_mul(s->n, s->mask, dst);
    if (rc) return rc;
    s->kW++;
    return SSR_OK;
}

// dst's basis := src's basis (DFGMRES.copy keeps the deflation basis).
int ssr_solver_copy_deflation(ssr_solver *dst, const ssr_solver *src)
{
    SSR_REQUIRE(dst && src && dst->n == src->n, "copy_deflation: mismatched solvers");
    dst->kW = 0;
    for (int c = 0; c < src->kW; ++c) {
        int rc = ssr_solver_expand_deflation(dst, src->W.p + (int64_t)c * src->n);
        if (rc) return rc;
    }
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

// A-orthonormalise the basis against the current operator (A_orthogonalize.m).
int ssr_solver_A_orthogonalize(ssr_solver *s)
{
    SSR_REQUIRE(s && s->K, "A_orthogonalize: operator not set");
    const int k = s->kW;
    if (k == 0) return SSR_OK;
    int rc;
    DevVec Wn;
    if ((rc = Wn.alloc(s->n * (int64_t)k))) return rc;
    std::vector<double> sgn;  // signs of accepted columns, in acceptance order
    int acc = 0;
    std::vector<double> hc;
    for (int i = 0; i < k; ++i) {
        const double *wsrc = s->W.p + (int64_t)(k - 1 - i) * s->n;
        double *temp = Wn.p + (int64_t)acc * s->n;  // candidate slot
        SSR_HIP_CHECK(hipMemcpyAsync(temp, wsrc, sizeof(double) * s->n, hipMemcpyDeviceToDevice, ssr_stream()));
        if ((rc = spmv(s, temp, s->t1.p))) return rc;  // A_temp
        if (acc > 0) {
            // temp -= (W_orth .* norms') * (A_temp' W_orth)'
            for (int c0 = 0; c0 < acc; c0 += 128) {
                const int kk = acc - c0 < 128 ? acc - c0 : 128;
                if ((rc = ssr_dev_multidot(s->n, kk, Wn.p + (int64_t)c0 * s->n, s->n, s->t1.p, s->hdev.p + c0))) return rc;
            }
            hc.resize(acc);
            SSR_HIP_CHECK(hipMemcpyAsync(hc.data(), s->hdev.p, sizeof(double) * acc, hipMemcpyDeviceToHost, ssr_stream()));
            SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
            for (int c = 0; c < acc; ++c) hc[c] *= sgn[c];
            SSR_HIP_CHECK(hipMemcpyAsync(s->hdev.p, hc.data(), sizeof(double) * acc, hipMemcpyHostToDevice, ssr_stream()));
            for (int c0 = 0; c0 < acc; c0 += 128) {
                const int kk = acc - c0 < 128 ? acc - c0 : 128;
                if ((rc = ssr_dev_gemv_update(s->n, kk, Wn.p + (int64_t)c0 * s->n, s->n, s->hdev.p + c0, 1.0, -1.0, temp)))
                    return rc;
            }
        }
        double nt = 0.0;
        if ((rc = ssr_dot(s->n, temp, s->t1.p, nullptr, &nt))) return rc;
        if (fabs(nt) > s->defl_tol) {
            double sg = 1.0;
            if (nt > 0) nt = sqrt(nt);
            else { nt = sqrt(fabs(nt)); sg = -1.0; }
            if ((rc = ssr_dev_axpby(s->n, 1.0 / nt, temp, 0.0, temp))) return rc;
            sgn.push_back(sg);
            acc++;
        }
    }
    // reverse the accepted columns back into W
    for (int c = 0; c < acc; ++c)
        SSR_HIP_CHECK(hipMemcpyAsync(s->W.p + (int64_t)c * s->n, Wn.p + (int64_t)(acc - 1 - c) * s->n,
                                     sizeof(double) * s->n, hipMemcpyDeviceToDevice, ssr_stream()));
    SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    s->kW = acc;
    return SSR_OK;
}

// Deflated flexible GMRES (dfgmres_solver.m).  b, x device vectors (length
// dim*n_n).  x is overwritten.
int ssr_solver_dfgmres(ssr_solver *s, const double *b, double *x, int *iters, double *relres)
{
    SSR_REQUIRE(s && s->K && b && x, "dfgmres: operator or vectors missing");
    const int64_t n = s->n;
    const int m = s->maxit;
    int rc;
    if ((rc = s->V.alloc(n * (int64_t)(m + 1))) || (rc = s->Z.alloc(n * (int64_t)(m > 0 ? m : 1)))) return rc;
    // x0
    if (s->kW > 0) {
        if ((rc = defl_coef(s, b))) return rc;
        if ((rc = defl_update(s, 0.0, 1.0, x))) return rc;
    } else {
        SSR_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n, ssr_stream()));
    }
    // r0 = b - A x  (into V(:,0))
    double *V = s->V.p, *Zp = s->Z.p;
    if ((rc = spmv(s, x, V))) return rc;
    if ((rc = ssr_dev_axpby(n, 1.0, b, -1.0, V))) return rc;
    double bnorm = hnorm(s, b, &rc);
    if (rc) return rc;
    if (bnorm == 0.0) bnorm = 1.0;
    double res = hnorm(s, V, &rc);
    if (rc) return rc;
    s->n_solves++;
    if (res / bnorm < s->tol || m == 0) {
        *iters = 0;
        *relres = res / bnorm;
        s->last_iters = 0;
        s->last_relres = *relres;
        return SSR_OK;
    }
    if ((rc = ssr_dev_axpby(n, 1.0 / res, V, 0.0, V))) return rc;
    std::vector<double> H((size_t)(m + 1) * m, 0.0), cs(m), sn(m), g(m + 1, 0.0), h(m + 1), h2(m + 1);
    g[0] = res;
    int it = 0;
    bool done = false;
    for (int j = 0; j < m && !done; ++j) {
        double *zj = Zp + (int64_t)j * n;
        double *u = V + (int64_t)(j + 1) * n;
        if ((rc = precond(s, V + (int64_t)j * n, zj))) return rc;
        if ((rc = project(s, zj))) return rc;
        if ((rc = spmv(s, zj, u))) return rc;
        // CGS2
        const int k = j + 1;
        for (int pass = 0; pass < 2; ++pass) {
            // all coefficients first (classical Gram-Schmidt), in chunks of 128 columns
            for (int c0 = 0; c0 < k; c0 += 128) {
                const int kk = k - c0 < 128 ? k - c0 : 128;
                if ((rc = ssr_dev_multidot(n, kk, V + (int64_t)c0 * n, n, u, s->hdev.p + c0))) return rc;
            }
            std::vector<double> &hv = pass == 0 ? h : h2;
            SSR_HIP_CHECK(hipMemcpyAsync(hv.data(), s->hdev.p, sizeof(double) * k, hipMemcpyDeviceToHost, ssr_stream()));
            for (int c0 = 0; c0 < k; c0 += 128) {
                const int kk = k - c0 < 128 ? k - c0 : 128;
                if ((rc = ssr_dev_gemv_update(n, kk, V + (int64_t)c0 * n, n, s->hdev.p + c0, 1.0, -1.0, u))) return rc;
            }
        }
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        for (int i = 0; i < k; ++i) h[i] += h2[i];
        const double hn = hnorm(s, u, &rc);
        if (rc) return rc;
        double *Hc = H.data() + (size_t)j * (m + 1);  // column j
        for (int i = 0; i < k; ++i) Hc[i] = h[i];
        Hc[k] = hn;
        it = j + 1;
        if (hn < 1e-14) {
            // exact solution within this Krylov space: close the column with
            // the previous rotations only
            for (int i = 0; i < j; ++i) {
                const double a = cs[i] * Hc[i] + sn[i] * Hc[i + 1];
                Hc[i + 1] = -sn[i] * Hc[i] + cs[i] * Hc[i + 1];
                Hc[i] = a;
            }
            cs[j] = 1.0;
            sn[j] = 0.0;
            res = fabs(g[j + 1]);
            done = true;
            break;
        }
        if ((rc = ssr_dev_axpby(n, 1.0 / hn, u, 0.0, u))) return rc;
        for (int i = 0; i < j; ++i) {
            const double a = cs[i] * Hc[i] + sn[i] * Hc[i + 1];
            Hc[i + 1] = -sn[i] * Hc[i] + cs[i] * Hc[i + 1];
            Hc[i] = a;
        }
        const double den = hypot(Hc[j], Hc[j + 1]);
        cs[j] = Hc[j] / den;
        sn[j] = Hc[j + 1] / den;
        Hc[j] = den;
        Hc[j + 1] = 0.0;
        g[j + 1] = -sn[j] * g[j];
        g[j] = cs[j] * g[j];
        res = fabs(g[j + 1]);
        if (res / bnorm < s->tol) done = true;
    }
    // back substitution R y = g
    std::vector<double> y(it, 0.0);
    for (int i = it - 1; i >= 0; --i) {
        double t = g[i];
        for (int c = i + 1; c < it; ++c) t -= H[(size_t)c * (m + 1) + i] * y[c];
        const double d = H[(size_t)i * (m + 1) + i];
        y[i] = d != 0.0 ? t / d : 0.0;
    }
    for (int c0 = 0; c0 < it; c0 += 128) {
        const int k = it - c0 < 128 ? it - c0 : 128;
        SSR_HIP_CHECK(hipMemcpyAsync(s->hdev.p, y.data() + c0, sizeof(double) * k, hipMemcpyHostToDevice, ssr_stream()));
        if ((rc = ssr_dev_gemv_update(n, k, Zp + (int64_t)c0 * n, n, s->hdev.p, 1.0, 1.0, x))) return rc;
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    }
    *iters = it;
    *relres = res / bnorm;
    s->last_iters = it;
    s->last_relres = *relres;
    s->total_iters += it;
    return SSR_OK;
}

// Deflated preconditioned CG (dcg_solver.m).
int ssr_solver_dcg(ssr_solver *s, const double *b, double *x, int *iters, double *relres)
{
    SSR_REQUIRE(s && s->K && b && x, "dcg: operator or vectors missing");
    const int64_t n = s->n;
    int rc;
    if ((rc = s->V.alloc(4 * n))) return rc;
    double *r = s->V.p, *z = r + n, *p = z + n, *q = p + n;
    if (s->kW > 0) {
        if ((rc = defl_coef(s, b))) return rc;
        if ((rc = defl_update(s, 0.0, 1.0, x))) return rc;
    } else {
        SSR_HIP_CHECK(hipMemsetAsync(x, 0, sizeof(double) * n, ssr_stream()));
    }
    if ((rc = spmv(s, x, r))) return rc;
    if ((rc = ssr_dev_axpby(n, 1.0, b, -1.0, r))) return rc;
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
    if ((rc = precond(s, r, z))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, ssr_stream()));
    if ((rc = project(s, p))) return rc;
    double gamma_old = 0.0;
    if ((rc = ssr_dot(n, r, z, nullptr, &gamma_old))) return rc;
    int j = 0;
    for (j = 1; j <= s->maxit; ++j) {
        if ((rc = spmv(s, p, q))) return rc;
        double sp = 0.0;
        if ((rc = ssr_dot(n, q, p, nullptr, &sp))) return rc;
        const double alpha = gamma_old / sp;
        if ((rc = ssr_dev_axpby(n, alpha, p, 1.0, x))) return rc;
        if ((rc = ssr_dev_axpby(n, -alpha, q, 1.0, r))) return rc;
        res = hnorm(s, r, &rc) / bnorm;
        if (rc) return rc;
        if (res < s->tol) break;
        if ((rc = precond(s, r, z))) return rc;
        double gamma_new = 0.0;
        if ((rc = ssr_dot(n, r, z, nullptr, &gamma_new))) return rc;
        const double beta = gamma_new / gamma_old;
        // p = P(z) + beta p   (q as scratch for P(z))
        SSR_HIP_CHECK(hipMemcpyAsync(q, z, sizeof(double) * n, hipMemcpyDeviceToDevice, ssr_stream()));
        if ((rc = project(s, q))) return rc;
        if ((rc = ssr_dev_axpby(n, 1.0, q, beta, p))) return rc;
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

int ssr_solver_precond_apply(ssr_solver *s, const double *x, double *y) { return precond(s, x, y); }

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
