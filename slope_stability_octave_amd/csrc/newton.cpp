// Bordered semismooth Newton of the indirect SSR continuation, driven from
// the library: newton_ind_SSR.m:1-206 with damping_ALG5.m:1-60.
//
// Every field (U, F, f, V, W, d_U and the constitutive work arrays) stays in
// device memory; the host keeps only the scalars the control flow needs (the
// residual norms, f'V, f'W, f'U and the damping factor), each read back with
// one synchronising reduction, as the Octave loop reads them.  The calls are
// the library's own C entry points, so the kernels and their order are those
// of the Python driver (newton.py) and the two give the same numbers.
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <vector>

#include "ssr_common.h"
#include "../../include/ssr_hip.h"

namespace {

#define NCHECK(expr)               \
    do {                           \
        int _rc = (expr);          \
        if (_rc != SSR_OK) return _rc; \
    } while (0)

// The element phases of a Newton solve replayed as captured HIP graphs
// (tuning key 16): the strain, stress (and tangent) and internal-force
// kernels of one build_F for a given (U, F, tangent or not), and the K_r
// assembly for a given r.  A phase runs eagerly the first time (so that any
// scratch a kernel allocates on first use exists), is captured the second
// time and replayed from then on.  The graphs belong to one Newton solve:
// its buffers do not change while it runs, and the graphs are destroyed, the
// stream drained first, when it returns.  The Davis reduction (lambda as a
// kernel argument) stays an ordinary launch before each replay.
//
// Capture happens only while this is the one Newton solve running in the
// process and no host thread runs the library on a stream of its own
// (concurrent instances replay nothing): another thread's synchronising
// calls (a hipFree, a blocking copy) can invalidate a capture in progress.  A capture that fails anyway
// is dropped and the phase runs eagerly from then on (nothing captured has
// executed, so the eager run gives the same results).
std::atomic<int> g_active_solves{0};

struct PhaseGraph {
    int kind;        // 0: build_F body, 1: K_r assembly
    const void *a;   // U (kind 0)
    const void *b;   // F (kind 0)
    int flag;        // with the tangent (kind 0)
    double r;        // regularisation (kind 1)
    int uses;
    bool nocap;      // capture failed: eager from now on
    hipGraphExec_t exec;
};

struct GraphCache {
    std::vector<PhaseGraph> g;
    hipStream_t cap = nullptr;  // capture stream (default priority)
    GraphCache() { g_active_solves.fetch_add(1); }
    GraphCache(const GraphCache &) = delete;
    GraphCache &operator=(const GraphCache &) = delete;
    ~GraphCache()
    {
        bool any = false;
        for (auto &e : g) any = any || e.exec;
        if (any) (void)hipStreamSynchronize(ssr_stream());
        for (auto &e : g)
            if (e.exec) (void)hipGraphExecDestroy(e.exec);
        if (cap) (void)hipStreamDestroy(cap);
        g_active_solves.fetch_sub(1);
    }
};

struct NewtonCtx {
    const ssr_constitutive *C;
    ssr_solver *solver;
    int method;
    int64_t n_dof;
    double *F, *F_eps, *G, *rhs, *V, *W, *dU, *Ua, *Fa, *tmp;
    GraphCache *graphs;
};

// Run body() (kernel launches on ssr_stream() only), or replay it from the
// graph captured for this key.  The kernel-only element forms are the
// default ones (the FP64-atomic forms clear their outputs with a memset,
// which a captured graph must not hold, so they always run eagerly).
template <class Body>
int replay(const NewtonCtx &x, int kind, const void *a, const void *b, int flag, double r, Body body)
{
    GraphCache *gc = x.graphs;
    if (!gc || !ssr_tuning(SSR_TUNE_NEWTON_GRAPH) || ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) != 0) return body();
    PhaseGraph *e = nullptr;
    for (auto &q : gc->g)
        if (q.kind == kind && q.a == a && q.b == b && q.flag == flag && q.r == r) e = &q;
    if (!e) {
        gc->g.push_back(PhaseGraph{kind, a, b, flag, r, 0, false, nullptr});
        e = &gc->g.back();
    }
    if (!e->exec) {
        if (e->nocap || e->uses++ == 0 || g_active_solves.load() > 1 || ssr_thread_streams_used()) return body();
        if (!gc->cap) SSR_HIP_CHECK(hipStreamCreateWithFlags(&gc->cap, hipStreamNonBlocking));
        SSR_HIP_CHECK(hipStreamBeginCapture(gc->cap, hipStreamCaptureModeThreadLocal));
        ssr_capture_stream_push(gc->cap);
        const int rc = body();
        ssr_capture_stream_pop();
        hipGraph_t g = nullptr;
        const hipError_t ec = hipStreamEndCapture(gc->cap, &g);
        hipGraphExec_t ge = nullptr;
        hipError_t ei = hipErrorUnknown;
        if (!rc && ec == hipSuccess && g) ei = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        if (g) (void)hipGraphDestroy(g);
        if (ei != hipSuccess) {
            // nothing of the capture ran: clear its error and run the phase
            (void)hipGetLastError();
            e->nocap = true;
            return body();
        }
        e->exec = ge;
    }
    SSR_HIP_CHECK(hipGraphLaunch(e->exec, ssr_stream()));
    return SSR_OK;
}

// host seconds spent in a phase (the calls that end in a synchronisation
// absorb the device work queued before them)
struct Lap {
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double take()
    {
        const auto t1 = std::chrono::steady_clock::now();
        const double s = std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
        return s;
    }
};

// CONSTITUTIVE.m:178-200 / reduction.m:25-58 at a scalar lambda
int reduction(const NewtonCtx &x, double lam)
{
    const ssr_constitutive *C = x.C;
    int c_all = 1;
    if (C->davis == 'C') {
        // Octave's vector `if` of reduction.m:47 holds only when every point has
        // atan(tan(phi)/lambda) > psi; the distinct materials decide it
        const int64_t n_host = C->n_friction > 0 ? C->n_friction : C->n_int;
        for (int64_t i = 0; i < n_host; ++i)
            if (!(atan(tan(C->phi_host[i]) / lam) > C->psi_host[i])) {
                c_all = 0;
                break;
            }
    }
    // lambda reaches the device by a one-element fill on the stream: no host
    // synchronisation (a host copy would drain the queue)
    NCHECK(ssr_dev_fill(1, lam, C->lambda));
    if (C->mat_id)  // one evaluation per material (the same values as per point)
        return ssr_dev_reduction(C->n_mat, C->davis, C->c0_m, C->phi_m, C->psi_m, C->lambda, 0, c_all, C->c_bar_m,
                                 C->sin_phi_m);
    return ssr_dev_reduction(C->n_int, C->davis, C->c0, C->phi, C->psi, C->lambda, 0, c_all, C->c_bar,
                             C->sin_phi);
}

// build_F_all / build_F_and_DS_all (CONSTITUTIVE.m:202-285): reduction, E = B U,
// S (and DS), F = B'(w S); the *_reduced forms (reduce = false) keep the
// material the caller reduced
int build_F_body(const NewtonCtx &x, const double *U, double *F, int with_ds)
{
    const ssr_constitutive *C = x.C;
    if (C->coord)
        NCHECK(ssr_dev_strain_geo(C->pattern, C->n_q, C->coord, U, C->E));
    else
        NCHECK(ssr_dev_strain(C->pattern, C->n_q, C->DPhi1, C->DPhi2, C->DPhi3, U, C->E));
    if (C->mat_id)
        NCHECK(ssr_dev_constitutive_mat(C->dim, C->n_int, C->n_q, C->E, C->mat_id, C->n_mat, C->c_bar_m,
                                        C->sin_phi_m, C->shear_m, C->bulk_m, C->lame_m, C->S,
                                        with_ds ? C->DS : nullptr, nullptr));
    else if (C->dim == 3)
        NCHECK(ssr_dev_constitutive_3D(C->n_int, C->E, C->c_bar, C->sin_phi, C->shear, C->bulk, C->lame, C->S,
                                       with_ds ? C->DS : nullptr, nullptr));
    else
        NCHECK(ssr_dev_constitutive_2D(C->n_int, C->E, C->c_bar, C->sin_phi, C->shear, C->bulk, C->lame, C->S,
                                       with_ds ? C->DS : nullptr, nullptr));
    if (C->coord) return ssr_dev_force_geo(C->pattern, C->n_q, C->coord, C->S, C->s_rows, C->WEIGHT, F);
    return ssr_dev_force(C->pattern, C->n_q, C->DPhi1, C->DPhi2, C->DPhi3, C->S, C->s_rows, C->WEIGHT, F);
}

int build_F(const NewtonCtx &x, double lam, const double *U, double *F, int with_ds, bool reduce = true)
{
    if (reduce) NCHECK(reduction(x, lam));
    return replay(x, 0, U, F, with_ds, 0.0, [&] { return build_F_body(x, U, F, with_ds); });
}

// K_r (CONSTITUTIVE.m:487-532) from the tangent of the last build_F with DS
int build_K_r(const NewtonCtx &x, double r)
{
    const ssr_constitutive *C = x.C;
    return replay(x, 1, nullptr, nullptr, 0, r, [&] {
        return ssr_dev_assemble_K_r(C->pattern, C->n_q, C->DPhi1, C->DPhi2, C->DPhi3, C->DS, 0, C->WEIGHT, r,
                                    C->K_elast, C->K_r);
    });
}

// ||a(Q) - b(Q)||
int norm_Q(const NewtonCtx &x, const double *a, const double *b, double *out)
{
    NCHECK(ssr_dev_lincomb(x.n_dof, 1.0, a, -1.0, b, x.tmp));
    double s = 0.0;
    NCHECK(ssr_dot(x.n_dof, x.tmp, x.tmp, x.C->mask, &s));
    *out = sqrt(s > 0.0 ? s : 0.0);
    return SSR_OK;
}

// damping_ALG5.m:26-58: halve alpha until the residual norm decreases.  A NaN
// trial residual counts as no decrease (Octave's `>=` would accept it).
int damping_ALG5(const NewtonCtx &x, int it_damp_max, const double *U_it, double lambda_it, double d_l,
                 double criterion, double *alpha_out)
{
    if (isnan(d_l)) {
        *alpha_out = 0.0;
        return SSR_OK;
    }
    double alpha = 1.0;
    int it_damp = 0;
    while (true) {
        ++it_damp;
        NCHECK(ssr_dev_lincomb(x.n_dof, 1.0, U_it, alpha, x.dU, x.Ua));
        NCHECK(build_F(x, lambda_it + alpha * d_l, x.Ua, x.Fa, 0));
        double crit_alpha;
        NCHECK(norm_Q(x, x.Fa, x.C->f, &crit_alpha));
        if (crit_alpha >= criterion || isnan(crit_alpha))
            alpha = alpha / 2;
        else
            break;
        if (it_damp >= it_damp_max) {
            alpha = 0.0;
            break;
        }
    }
    *alpha_out = alpha;
    return SSR_OK;
}

// damping.m:25-54: bisection on the directional derivative (F(U + a dU) - f)'dU
// over all DOFs, at the material the caller reduced (f: the load vector, or
// a zero field for the limit-load Newton)
int damping(const NewtonCtx &x, int it_damp_max, const double *U_it, const double *F, const double *f,
            double *alpha_out)
{
    const int64_t n = x.n_dof;
    double decrease = 0.0, dd = 0.0;
    NCHECK(ssr_dev_lincomb(n, 1.0, F, -1.0, f, x.tmp));
    NCHECK(ssr_dot(n, x.tmp, x.dU, nullptr, &decrease));
    NCHECK(ssr_dot(n, x.dU, x.dU, nullptr, &dd));
    if (isnan(decrease) || isinf(sqrt(dd)) || decrease >= 0) {
        *alpha_out = 0.0;
        return SSR_OK;
    }
    double alpha = 1.0, a_min = 0.0, a_max = 1.0;
    for (int it_damp = 0; it_damp < it_damp_max;) {
        ++it_damp;
        NCHECK(ssr_dev_lincomb(n, 1.0, U_it, alpha, x.dU, x.Ua));
        NCHECK(build_F(x, 0.0, x.Ua, x.Fa, 0, false));
        NCHECK(ssr_dev_lincomb(n, 1.0, x.Fa, -1.0, f, x.tmp));
        NCHECK(ssr_dot(n, x.tmp, x.dU, nullptr, &decrease));
        if (decrease < 0) {
            if (alpha == 1) break;
            a_min = alpha;
        } else {
            a_max = alpha;
        }
        alpha = (a_min + a_max) / 2;
    }
    *alpha_out = alpha;
    return SSR_OK;
}

int solve(const NewtonCtx &x, const double *b, double *sol, int *iters)
{
    double rr = 0.0;
    if (x.method == 1) return ssr_solver_dcg(x.solver, b, sol, iters, &rr);
    return ssr_solver_dfgmres(x.solver, b, sol, iters, &rr);
}

// argument checks shared by the entry points; fills the context
int prepare(const ssr_constitutive *C, ssr_solver *solver, int method, int it_newt_max, int it_damp_max,
            double *work, int64_t work_len, NewtonCtx *x)
{
    SSR_REQUIRE(C->pattern && C->DPhi1 && C->DPhi2 && (C->dim == 2 || C->DPhi3) && C->WEIGHT && C->c0 &&
                    C->phi && C->psi && C->shear && C->bulk && C->lame && C->lambda && C->c_bar && C->sin_phi &&
                    C->E && C->S && C->DS && C->K_elast && C->K_tang && C->K_r && C->mask && C->f,
                "newton: incomplete constitutive description");
    SSR_REQUIRE(C->dim == 2 || C->dim == 3, "newton: dim must be 2 or 3");
    SSR_REQUIRE(!C->mat_id || (C->n_mat > 0 && C->c0_m && C->phi_m && C->psi_m && C->shear_m && C->bulk_m &&
                               C->lame_m && C->c_bar_m && C->sin_phi_m),
                "newton: incomplete per-material fields");
    SSR_REQUIRE(C->s_rows == (C->dim == 3 ? 6 : 4), "newton: S has 6 rows in 3D, 4 in 2D");
    SSR_REQUIRE(C->davis == 'A' || C->davis == 'B' || C->davis == 'C', "Incorrect choice of the Davis approach.");
    SSR_REQUIRE(C->davis != 'C' || (C->phi_host && C->psi_host), "newton: Davis C needs phi, psi on the host");
    SSR_REQUIRE(method == 0 || method == 1, "newton: method 0 (DFGMRES) or 1 (DCG)");
    SSR_REQUIRE(it_newt_max >= 1 && it_damp_max >= 1, "newton: iteration limits must be positive");
    int64_t info[8];
    NCHECK(ssr_pattern_info(C->pattern, info));
    SSR_REQUIRE(info[0] == C->dim, "newton: pattern dimension %lld != %d", (long long)info[0], C->dim);
    SSR_REQUIRE(info[3] > 0 && C->n_q > 0 && C->n_int == info[3] * C->n_q,
                "newton: n_int %lld != n_e %lld x n_q %d", (long long)C->n_int, (long long)info[3], C->n_q);
    const int64_t n_dof = info[0] * info[1];
    SSR_REQUIRE(work_len >= 10 * n_dof, "newton: work holds %lld doubles, %lld needed", (long long)work_len,
                (long long)(10 * n_dof));
    x->C = C;
    x->solver = solver;
    x->method = method;
    x->n_dof = n_dof;
    x->graphs = nullptr;
    double **slots[10] = {&x->F, &x->F_eps, &x->G, &x->rhs, &x->V, &x->W, &x->dU, &x->Ua, &x->Fa, &x->tmp};
    for (int k = 0; k < 10; ++k) *slots[k] = work + k * n_dof;
    return SSR_OK;
}

}  // namespace

extern "C" int ssr_newton_work_len(const ssr_constitutive *C, int64_t *len)
{
    SSR_REQUIRE(C && C->pattern && len, "newton_work_len: null argument");
    int64_t info[8];
    NCHECK(ssr_pattern_info(C->pattern, info));
    *len = 10 * info[0] * info[1];
    return SSR_OK;
}

extern "C" int ssr_newton_ind_SSR(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini,
                                  double omega, double lambda_ini, int it_newt_max, int it_damp_max, double tol,
                                  double r_min, double *work, int64_t work_len, double *U_it, double *lambda_out,
                                  int *flag_out, int *it_out, int *status_out, double *hist, double *times)
{
    SSR_REQUIRE(C && solver && U_ini && work && U_it && lambda_out && flag_out && it_out && status_out,
                "newton_ind_SSR: null argument");
    NewtonCtx x;
    NCHECK(prepare(C, solver, method, it_newt_max, it_damp_max, work, work_len, &x));
    GraphCache graphs;
    x.graphs = &graphs;
    const int64_t n_dof = x.n_dof;

    // newton_ind_SSR.m:22-33
    NCHECK(ssr_dev_lincomb(n_dof, 1.0, U_ini, 0.0, U_ini, U_it));
    double lambda_it = lambda_ini;
    const double eps = tol / 1000;
    double ff = 0.0;
    NCHECK(ssr_dot(n_dof, C->f, C->f, C->mask, &ff));
    const double norm_f = sqrt(ff);
    int flag_N = 0;
    bool compute_diffs = true;
    double r = r_min;
    int it = 0;
    double rel_resid = NAN, criterion = NAN;
    int status = SSR_NEWTON_UNKNOWN;
    if (hist)
        for (int64_t k = 0; k < 6 * (int64_t)it_newt_max; ++k) hist[k] = NAN;
    double tm[4] = {0.0, 0.0, 0.0, 0.0};  // solves, A-orthogonalisation, preconditioner, the rest
    Lap lap;

    // newton_ind_SSR.m:68-175
    while (true) {
        ++it;
        double *h = hist ? hist + 6 * (int64_t)(it - 1) : nullptr;
        if (compute_diffs) {
            NCHECK(build_F(x, lambda_it, U_it, x.F, 1));
            NCHECK(norm_Q(x, x.F, C->f, &criterion));
            rel_resid = criterion / norm_f;
            if (h) h[0] = rel_resid;
            if (rel_resid < tol && it > 1) {
                status = SSR_NEWTON_CONVERGED;
                break;
            }
        } else if (h) {
            h[0] = rel_resid;
        }
        if (h) h[1] = r;
        // K_r (CONSTITUTIVE.m:487-532)
        NCHECK(build_K_r(x, r));
        // G = (F(lambda + eps) - F) / eps   (:94-96)
        NCHECK(build_F(x, lambda_it + eps, U_it, x.F_eps, 0));
        NCHECK(ssr_dev_lincomb(n_dof, 1.0 / eps, x.F_eps, -1.0 / eps, x.F, x.G));
        // preconditioner and deflation basis for K_r (:99-110)
        tm[3] += lap.take();
        NCHECK(ssr_solver_set_matrix(solver, C->K_r, 1));
        tm[2] += lap.take();
        NCHECK(ssr_solver_A_orthogonalize(solver));
        tm[1] += lap.take();
        if (h) {
            int k = 0;
            NCHECK(ssr_solver_deflation_size(solver, &k));
            h[5] = k;
        }
        // W = K \ (-G(Q)),  V = K \ (f(Q) - F(Q))   (:112-115)
        int lin_w = 0, lin_v = 0;
        NCHECK(ssr_dev_lincomb(n_dof, -1.0, x.G, 0.0, x.G, x.rhs));
        NCHECK(ssr_dev_mul(n_dof, C->mask, x.rhs));
        NCHECK(solve(x, x.rhs, x.W, &lin_w));
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, C->f, -1.0, x.F, x.rhs));
        NCHECK(ssr_dev_mul(n_dof, C->mask, x.rhs));
        NCHECK(solve(x, x.rhs, x.V, &lin_v));
        tm[0] += lap.take();
        if (h) {
            h[3] = lin_w;
            h[4] = lin_v;
        }
        // d_l = -f'V / f'W,  d_U = V + d_l W   (:121-122)
        double fV = 0.0, fW = 0.0;
        NCHECK(ssr_dot(n_dof, C->f, x.V, C->mask, &fV));
        NCHECK(ssr_dot(n_dof, C->f, x.W, C->mask, &fW));
        const double d_l = fW != 0.0 ? -fV / fW : NAN;
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, x.V, d_l, x.W, x.dU));
        double alpha = 0.0;
        NCHECK(damping_ALG5(x, it_damp_max, U_it, lambda_it, d_l, criterion, &alpha));
        if (h) h[2] = alpha;
        // regularisation and deflation-basis update (:129-143)
        compute_diffs = true;
        if (alpha < 1e-1) {
            if (alpha == 0) {
                compute_diffs = false;
                r = r * 2;
            } else {
                r = r * pow(2.0, 1.0 / 4);
            }
        } else {
            NCHECK(ssr_solver_expand_deflation(solver, x.W));
            NCHECK(ssr_solver_expand_deflation(solver, x.V));
            if (alpha > 0.5) r = fmax(r / sqrt(2.0), r_min);
        }
        if (alpha == 0 && r > 1) {
            status = SSR_NEWTON_STALLED;
            if (rel_resid > 10 * tol) flag_N = 1;
            break;
        }
        // U += alpha d_U; U = omega U / (f(Q)'U(Q)); lambda += alpha d_l   (:154-156)
        NCHECK(ssr_dev_axpby(n_dof, alpha, x.dU, 1.0, U_it));
        double fU = 0.0;
        NCHECK(ssr_dot(n_dof, C->f, U_it, C->mask, &fU));
        NCHECK(ssr_dev_axpby(n_dof, 0.0, U_it, omega / fU, U_it));
        lambda_it = lambda_it + alpha * d_l;
        // (:164-174; a NaN residual fails the solve)
        if (isnan(rel_resid) || it == it_newt_max) {
            if (rel_resid > 10 * tol || isnan(rel_resid)) flag_N = 1;
            status = isnan(rel_resid) ? SSR_NEWTON_NAN : SSR_NEWTON_MAXIT;
            break;
        }
    }
    NCHECK(ssr_synchronize());
    tm[3] += lap.take();
    if (times)
        for (int q = 0; q < 4; ++q) times[q] += tm[q];
    *lambda_out = lambda_it;
    *flag_out = flag_N;
    *it_out = it;
    *status_out = status;
    return SSR_OK;
}

extern "C" int ssr_newton(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini,
                          int it_newt_max, int it_damp_max, double tol, double r_min, double *work, int64_t work_len,
                          double *U_it, int *flag_out, int *it_out, int *status_out, double *hist, double *times)
{
    SSR_REQUIRE(C && solver && U_ini && work && U_it && flag_out && it_out && status_out, "newton: null argument");
    NewtonCtx x;
    NCHECK(prepare(C, solver, method, it_newt_max, it_damp_max, work, work_len, &x));
    GraphCache graphs;
    x.graphs = &graphs;
    const int64_t n_dof = x.n_dof;

    // newton.m:36-44
    NCHECK(ssr_dev_lincomb(n_dof, 1.0, U_ini, 0.0, U_ini, U_it));
    double ff = 0.0;
    NCHECK(ssr_dot(n_dof, C->f, C->f, C->mask, &ff));
    const double norm_f = sqrt(ff);
    int flag_N = 0;
    bool compute_diffs = true;
    double r = r_min;
    int it = 0;
    double rel_resid = NAN, criterion = NAN;
    int status = SSR_NEWTON_UNKNOWN;
    if (hist)
        for (int64_t k = 0; k < 5 * (int64_t)it_newt_max; ++k) hist[k] = NAN;
    double tm[4] = {0.0, 0.0, 0.0, 0.0};
    Lap lap;

    // newton.m:74-160
    while (true) {
        ++it;
        double *h = hist ? hist + 5 * (int64_t)(it - 1) : nullptr;
        if (compute_diffs) {
            NCHECK(build_F(x, 0.0, U_it, x.F, 1, false));
            NCHECK(norm_Q(x, x.F, C->f, &criterion));
            rel_resid = criterion / norm_f;
            if (h) h[0] = rel_resid;
            if (rel_resid < tol) {
                status = SSR_NEWTON_CONVERGED;
                break;
            }
        } else if (h) {
            h[0] = rel_resid;
        }
        if (h) h[1] = r;
        NCHECK(build_K_r(x, r));
        tm[3] += lap.take();
        NCHECK(ssr_solver_set_matrix(solver, C->K_r, 1));
        tm[2] += lap.take();
        NCHECK(ssr_solver_A_orthogonalize(solver));
        tm[1] += lap.take();
        if (h) {
            int k = 0;
            NCHECK(ssr_solver_deflation_size(solver, &k));
            h[4] = k;
        }
        // dU = K \ (f(Q) - F(Q))   (:112-113)
        int lin_it = 0;
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, C->f, -1.0, x.F, x.rhs));
        NCHECK(ssr_dev_mul(n_dof, C->mask, x.rhs));
        NCHECK(solve(x, x.rhs, x.dU, &lin_it));
        tm[0] += lap.take();
        if (h) h[3] = lin_it;
        double alpha = 0.0;
        NCHECK(damping(x, it_damp_max, U_it, x.F, C->f, &alpha));
        if (h) h[2] = alpha;
        // (:121-143)
        compute_diffs = true;
        if (alpha < 1e-1) {
            if (alpha == 0) {
                compute_diffs = false;
                r = r * 2;
            } else {
                r = r * pow(2.0, 1.0 / 4);
            }
        } else {
            NCHECK(ssr_solver_expand_deflation(solver, x.dU));
            if (alpha > 0.5) r = fmax(r / sqrt(2.0), r_min);
        }
        if (alpha == 0 && r > 1) {
            status = SSR_NEWTON_STALLED;
            if (rel_resid > 10 * tol) flag_N = 1;
            break;
        }
        NCHECK(ssr_dev_axpby(n_dof, alpha, x.dU, 1.0, U_it));
        // (:153-160; a NaN residual fails the solve)
        if (isnan(rel_resid) || it == it_newt_max) {
            if (rel_resid > 10 * tol || isnan(rel_resid)) flag_N = 1;
            status = isnan(rel_resid) ? SSR_NEWTON_NAN : SSR_NEWTON_MAXIT;
            break;
        }
    }
    NCHECK(ssr_synchronize());
    tm[3] += lap.take();
    if (times)
        for (int q = 0; q < 4; ++q) times[q] += tm[q];
    *flag_out = flag_N;
    *it_out = it;
    *status_out = status;
    return SSR_OK;
}

extern "C" int ssr_newton_ind_LL(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini,
                                 double t_ini, double omega, int it_newt_max, int it_damp_max, double tol,
                                 double r_min, double *work, int64_t work_len, double *U_it, double *t_out,
                                 int *flag_out, int *it_out, int *status_out, double *hist, double *times)
{
    SSR_REQUIRE(C && solver && U_ini && work && U_it && t_out && flag_out && it_out && status_out,
                "newton_ind_LL: null argument");
    NewtonCtx x;
    NCHECK(prepare(C, solver, method, it_newt_max, it_damp_max, work, work_len, &x));
    GraphCache graphs;
    x.graphs = &graphs;
    const int64_t n_dof = x.n_dof;
    double *zero = x.G;  // the damping's target: F(U) against f = 0
    NCHECK(ssr_dev_fill(n_dof, 0.0, zero));

    // newton_ind_LL.m setup (the lines before the loop at :80)
    NCHECK(ssr_dev_lincomb(n_dof, 1.0, U_ini, 0.0, U_ini, U_it));
    double t_it = t_ini;
    double ff = 0.0;
    NCHECK(ssr_dot(n_dof, C->f, C->f, C->mask, &ff));
    const double norm_f = sqrt(ff);
    int flag_N = 0;
    bool compute_diffs = true;
    double r = r_min;
    int it = 0;
    double rel_resid = NAN;
    int status = SSR_NEWTON_UNKNOWN;
    if (hist)
        for (int64_t k = 0; k < 6 * (int64_t)it_newt_max; ++k) hist[k] = NAN;
    double tm[4] = {0.0, 0.0, 0.0, 0.0};
    Lap lap;

    // newton_ind_LL.m:80-177: F(U) = t f, f'U = omega at the unreduced material
    while (true) {
        ++it;
        double *h = hist ? hist + 6 * (int64_t)(it - 1) : nullptr;
        if (compute_diffs) NCHECK(build_F(x, 0.0, U_it, x.F, 1, false));
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, x.F, -t_it, C->f, x.tmp));
        double s2 = 0.0;
        NCHECK(ssr_dot(n_dof, x.tmp, x.tmp, C->mask, &s2));
        const double criterion = sqrt(s2 > 0.0 ? s2 : 0.0);
        rel_resid = criterion / norm_f;
        if (h) h[0] = rel_resid;
        if (criterion < tol * norm_f && it > 1) {
            status = SSR_NEWTON_CONVERGED;
            break;
        }
        if (h) h[1] = r;
        NCHECK(build_K_r(x, r));
        tm[3] += lap.take();
        NCHECK(ssr_solver_set_matrix(solver, C->K_r, 1));
        tm[2] += lap.take();
        NCHECK(ssr_solver_A_orthogonalize(solver));
        tm[1] += lap.take();
        if (h) {
            int k = 0;
            NCHECK(ssr_solver_deflation_size(solver, &k));
            h[5] = k;
        }
        // W = K \ f(Q),  V = K \ (t f(Q) - F(Q))
        int lin_w = 0, lin_v = 0;
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, C->f, 0.0, C->f, x.rhs));
        NCHECK(ssr_dev_mul(n_dof, C->mask, x.rhs));
        NCHECK(solve(x, x.rhs, x.W, &lin_w));
        NCHECK(ssr_dev_lincomb(n_dof, t_it, C->f, -1.0, x.F, x.rhs));
        NCHECK(ssr_dev_mul(n_dof, C->mask, x.rhs));
        NCHECK(solve(x, x.rhs, x.V, &lin_v));
        tm[0] += lap.take();
        if (h) {
            h[3] = lin_w;
            h[4] = lin_v;
        }
        // d_t = -f'V / f'W,  d_U = V + d_t W
        double fV = 0.0, fW = 0.0;
        NCHECK(ssr_dot(n_dof, C->f, x.V, C->mask, &fV));
        NCHECK(ssr_dot(n_dof, C->f, x.W, C->mask, &fW));
        const double d_t = fW != 0.0 ? -fV / fW : NAN;
        NCHECK(ssr_dev_lincomb(n_dof, 1.0, x.V, d_t, x.W, x.dU));
        double alpha = 0.0;
        NCHECK(damping(x, it_damp_max, U_it, x.F, zero, &alpha));  // :130, against F_int*0
        if (h) h[2] = alpha;
        compute_diffs = true;
        if (alpha < 1e-1) {
            if (alpha == 0) {
                compute_diffs = false;
                r = r * 2;
            } else {
                r = r * pow(2.0, 1.0 / 4);
            }
        } else {
            NCHECK(ssr_solver_expand_deflation(solver, x.W));
            NCHECK(ssr_solver_expand_deflation(solver, x.V));
            if (alpha > 0.5) r = fmax(r / sqrt(2.0), r_min);
        }
        if (alpha == 0 && r > 1) {
            status = SSR_NEWTON_STALLED;
            flag_N = 1;
            break;
        }
        // U += alpha d_U; U = omega U / (f(Q)'U(Q)); t takes the full d_t
        NCHECK(ssr_dev_axpby(n_dof, alpha, x.dU, 1.0, U_it));
        double fU = 0.0;
        NCHECK(ssr_dot(n_dof, C->f, U_it, C->mask, &fU));
        NCHECK(ssr_dev_axpby(n_dof, 0.0, U_it, omega / fU, U_it));
        t_it = t_it + d_t;
        // reaching it_newt_max is a failure here
        if (isnan(rel_resid) || it == it_newt_max) {
            flag_N = 1;
            status = isnan(rel_resid) ? SSR_NEWTON_NAN : SSR_NEWTON_MAXIT;
            break;
        }
    }
    NCHECK(ssr_synchronize());
    tm[3] += lap.take();
    if (times)
        for (int q = 0; q < 4; ++q) times[q] += tm[q];
    *t_out = t_it;
    *flag_out = flag_N;
    *it_out = it;
    *status_out = status;
    return SSR_OK;
}

// ---------------------------------------------------------------------------
// The indirect SSR continuation in the library: SSR_indirect_continuation.m
// (omega stepping, secant predictor :119, retries with half the omega step
// :164-167, omega step doubling :214-216, stop rules :203-211 and :220-227)
// with its initial phase init_phase_SSR_indirect_continuation.m, over the
// Newton loops above.  Every field stays on the device; the host holds the
// histories and the step control.  The order of operations (reductions,
// Newton solves on solver copies, basis expansions, dot products) is that of
// the host driver continuation.py, so the two give the same numbers.
// ---------------------------------------------------------------------------
namespace {

struct DevVec {
    double *p = nullptr;
    ~DevVec() { ssr_pool_free(p); }
    int alloc(int64_t n) { return ssr_pool_malloc((void **)&p, sizeof(double) * (n > 0 ? n : 1)); }
};

struct SolverCopy {
    ssr_solver *s = nullptr;
    ~SolverCopy() { (void)ssr_solver_destroy(s); }
};

// Newton iterations and the Krylov iterations of their solves, from the
// per-iteration records (columns `kcols` hold the Krylov counts)
void add_krylov(const std::vector<double> &hist, int width, int its, int kc0, int kc1, int64_t *krylov)
{
    for (int i = 0; i < its; ++i)
        for (int c = kc0; c <= kc1; ++c) {
            const double v = hist[(size_t)width * i + c];
            if (v == v) *krylov += (int64_t)v;
        }
}

}  // namespace

extern "C" int ssr_SSR_indirect_continuation(const ssr_constitutive *C, ssr_solver *solver, int method,
                                             int share_precond, const double *params, double *U_out,
                                             double *lambda_hist, double *omega_hist, double *Umax_hist,
                                             int hist_cap, int *n_hist, int *attempt_its, int attempt_cap,
                                             int *n_attempts, int *init_its, int init_cap, int *n_init,
                                             int64_t *krylov_its, int *stop)
{
    SSR_REQUIRE(C && solver && params && U_out && lambda_hist && omega_hist && Umax_hist && n_hist && attempt_its &&
                    n_attempts && init_its && n_init && krylov_its && stop && hist_cap >= 2 && attempt_cap >= 1 &&
                    init_cap >= 2,
                "SSR_indirect_continuation: bad arguments");
    const double lambda_init = params[0], d_lambda_init = params[1], d_lambda_min = params[2],
                 d_lambda_diff_scaled_min = params[3], omega_max_stop = params[5], tol = params[8], r_min = params[9];
    const int step_max = (int)params[4], it_newt_max = (int)params[6], it_damp_max = (int)params[7];
    SSR_REQUIRE(it_newt_max > 0 && it_damp_max > 0 && step_max > 0, "SSR_indirect_continuation: bad parameters");
    int64_t info[8];
    NCHECK(ssr_pattern_info(C->pattern, info));
    const int64_t n_dof = info[0] * info[1], n_n = info[1];
    const int dim = (int)info[0];
    int64_t wl = 0;
    NCHECK(ssr_newton_work_len(C, &wl));
    DevVec work, Ua, Ub, Uc, Uini;
    NCHECK(work.alloc(wl));
    NCHECK(Ua.alloc(n_dof));
    NCHECK(Ub.alloc(n_dof));
    NCHECK(Uc.alloc(n_dof));
    NCHECK(Uini.alloc(n_dof));
    std::vector<double> hist((size_t)6 * it_newt_max);
    double tm[4] = {0, 0, 0, 0};
    NewtonCtx x{};
    x.C = C;
    *n_hist = *n_attempts = *n_init = 0;
    *krylov_its = 0;
    *stop = 0;

    // init_phase_SSR_indirect_continuation.m, on one copy of the solver
    double *U1 = Ua.p, *U2 = Ub.p, *Ut = Uc.p;
    double omega1 = 0.0, omega2 = 0.0, lambda1 = lambda_init, lambda2 = 0.0;
    {
        SolverCopy init;
        NCHECK(ssr_solver_clone(solver, share_precond, &init.s));
        NCHECK(ssr_dev_fill(n_dof, 0.0, Uini.p));
        NCHECK(reduction(x, lambda_init));
        int flag = 0, it = 0, st = 0;
        std::fill(hist.begin(), hist.end(), NAN);
        NCHECK(ssr_newton(C, init.s, method, Uini.p, it_newt_max, it_damp_max, tol, r_min, work.p, wl, U1, &flag, &it,
                          &st, hist.data(), tm));
        add_krylov(hist, 5, it, 3, 3, krylov_its);
        init_its[(*n_init)++] = it;
        if (flag == 1) {
            *stop = -1;
            ssr_set_error("Initial choice of lambda seems to be too large.");
            return SSR_ERR_STATE;
        }
        NCHECK(ssr_dot(n_dof, C->f, U1, C->mask, &omega1));
        double d_lambda = d_lambda_init;
        NCHECK(ssr_solver_expand_deflation(init.s, U1));
        while (true) {
            const double lambda_it = lambda1 + d_lambda;
            NCHECK(reduction(x, lambda_it));
            std::fill(hist.begin(), hist.end(), NAN);
            NCHECK(ssr_newton(C, init.s, method, U1, it_newt_max, it_damp_max, tol, r_min, work.p, wl, Ut, &flag, &it,
                              &st, hist.data(), tm));
            add_krylov(hist, 5, it, 3, 3, krylov_its);
            if (*n_init < init_cap) init_its[(*n_init)++] = it;
            if (flag == 1) {
                d_lambda = d_lambda / 2;
            } else {
                std::swap(U2, Ut);
                NCHECK(ssr_dot(n_dof, C->f, U2, C->mask, &omega2));
                lambda2 = lambda_it;
                if ((omega2 - omega1) / fmax(1.0, omega1) < 1e-5) {
                    std::swap(U1, U2);
                    lambda1 = lambda2;
                    omega1 = omega2;
                } else {
                    break;
                }
            }
            if (d_lambda < d_lambda_min) {
                *stop = -2;
                ssr_set_error("It seems that the FoS is equal to lambda_init.");
                return SSR_ERR_STATE;
            }
            if (lambda1 > 10) {
                *stop = -3;
                ssr_set_error("It seems that the FoS is greater than 10.");
                return SSR_ERR_STATE;
            }
        }
    }
    // SSR_indirect_continuation.m:86-110
    double *U_old = U1, *U = U2;  // Ut is free
    NCHECK(ssr_solver_expand_deflation(solver, U_old));
    int nh = 0;
    auto push_hist = [&](double lam, double om, const double *Uf) -> int {
        SSR_REQUIRE(nh < hist_cap, "SSR_indirect_continuation: history capacity %d exceeded", hist_cap);
        double um = 0.0;
        NCHECK(ssr_node_norm_max(dim, n_n, Uf, &um));
        lambda_hist[nh] = lam;
        omega_hist[nh] = om;
        Umax_hist[nh] = um;
        ++nh;
        *n_hist = nh;
        return SSR_OK;
    };
    NCHECK(push_hist(lambda1, omega1, U_old));
    NCHECK(push_hist(lambda2, omega2, U));
    double omega_old = omega1, omega = omega2, lam = lambda2;
    double d_omega = omega - omega_old;
    if (omega_max_stop < omega + d_omega) {
        *stop = -4;
        ssr_set_error("Too small value of omega_max_stop. It is necessary to increase this input quantinty.");
        return SSR_ERR_STATE;
    }
    int step = 2, n_omega = 0;
    const int n_omega_max = 5;
    while (true) {
        const double omega_it = fmin(omega + d_omega, omega_max_stop);
        d_omega = omega_it - omega;
        // secant predictor (:119): U_ini = a (U - U_old) + U
        const double a = d_omega / (omega - omega_old);
        NCHECK(ssr_dev_lincomb(n_dof, 1.0 + a, U, -a, U_old, Uini.p));
        SolverCopy cp;
        NCHECK(ssr_solver_clone(solver, share_precond, &cp.s));
        double lambda_it = 0.0;
        int flag = 0, it = 0, st = 0;
        std::fill(hist.begin(), hist.end(), NAN);
        NCHECK(ssr_newton_ind_SSR(C, cp.s, method, Uini.p, omega_it, lam, it_newt_max, it_damp_max, tol, r_min, work.p,
                                  wl, Ut, &lambda_it, &flag, &it, &st, hist.data(), tm));
        add_krylov(hist, 6, it, 3, 4, krylov_its);
        if (*n_attempts < attempt_cap) attempt_its[(*n_attempts)++] = it;
        if (flag == 1) {
            // a failed solve asks for a fresh preconditioner before the retry
            NCHECK(ssr_solver_invalidate_preconditioner(solver));
            d_omega = d_omega / 2;
            n_omega += 1;
        } else {
            step += 1;
            double *freed = U_old;
            U_old = U;
            U = Ut;
            Ut = freed;
            NCHECK(ssr_solver_expand_deflation(solver, U));
            omega_old = omega;
            omega = omega_it;
            const double d_lambda = lambda_it - lam;
            lam = lambda_it;
            n_omega = 0;
            NCHECK(push_hist(lam, omega, U));
            const double scaled = d_lambda / d_omega * (omega_hist[nh - 1] - omega_hist[0]);
            if (scaled < d_lambda_diff_scaled_min) {
                *stop = 1;  // minimal increase of lambda
                break;
            }
            if (omega >= omega_max_stop) {
                *stop = 2;  // maximal omega
                break;
            }
            if ((lambda_hist[nh - 1] - lambda_hist[nh - 2]) < 0.9 * (lambda_hist[nh - 2] - lambda_hist[nh - 3]))
                d_omega = 2 * d_omega;
        }
        if (n_omega >= n_omega_max) {
            *stop = 3;  // impossible to increase omega
            break;
        }
        if (step >= step_max) {
            *stop = 4;  // maximal number of steps
            break;
        }
    }
    NCHECK(ssr_dev_lincomb(n_dof, 1.0, U, 0.0, U, U_out));
    return ssr_synchronize();
}
