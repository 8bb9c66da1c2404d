// Smoothed-aggregation AMG preconditioner for the tangent operator.
//
// Reference counterpart: the BoomerAMG binding (hypre_boomeramg_mex.cpp:
// setup_instance_ijv :855-1010, command_update_values :1151-1211, command_apply
// :751-796) with the rigid-body near-null space of
// near_null_space_elasticity_3D.m:46-103 (centred coordinates).  BoomerAMG is a
// third-party, unpinned dependency; this is an algorithmic substitute, not a
// port: classical smoothed aggregation (Vanek-Mandel-Brezina).
//
//   setup  (host, C++)   node-block strength of connection, greedy aggregation,
//                        tentative prolongator by local QR of the near-null
//                        space, one damped-Jacobi smoothing step of P,
//                        Galerkin coarse operators, dense coarsest level.
//   refresh (host)       same aggregates and P, new Galerkin products for
//                        new operator values (update_values).
//   apply  (GPU)         V(1,1)-cycle with Chebyshev(2) smoothing on D^-1 A,
//                        scalar-CSR SpMV on coarse levels, the node-block
//                        operator on the finest level.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <omp.h>

#include <algorithm>
#include <array>
#include <chrono>
#include <memory>
#include <vector>

#include "amg_device.h"
#include "linalg.h"
#include "ssr_common.h"

struct ssr_pattern;
struct ssr_solver;
const int32_t *ssr_pattern_d_row_ptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_col(const ssr_pattern *p);
const double *ssr_pattern_d_mask(const ssr_pattern *p);
void ssr_pattern_sizes(const ssr_pattern *p, int *dim, int64_t *n_n, int64_t *n_p, int64_t *n_e, int64_t *nnzb,
                       int64_t *nnz_free);

const int32_t *ssr_pattern_d_diag(const ssr_pattern *p);
const int32_t *ssr_pattern_d_tblk(const ssr_pattern *p);
const int32_t *ssr_pattern_d_brow(const ssr_pattern *p);

extern "C" {
int ssr_pattern_get(const ssr_pattern *p, int which, void *dst);
int ssr_solver_set_precond(ssr_solver *s, ssr_precond *M);
const ssr_pattern *ssr_solver_pattern(const ssr_solver *s);
}


namespace {

// Wall-clock phase totals of the host setup (printed with SSR_AMG_VERBOSE).
struct PhaseClock {
    double t[12] = {0};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    void lap(int k)
    {
        const auto t1 = std::chrono::steady_clock::now();
        t[k] += std::chrono::duration<double>(t1 - t0).count();
        t0 = t1;
    }
};
const char *kPhaseNames[12] = {"download", "fine_csr", "rho", "aggregate", "tentative", "A*T",
                               "P", "R=P'", "RAP", "upload", "coarse_inv", "other"};

struct HCsr {
    int64_t nr = 0, nc = 0;
    std::vector<int64_t> rp;
    std::vector<int32_t> ci;
    std::vector<double> v;
};

// A' with the rows of each column in ascending order: per-thread column
// counts over fixed row ranges give each thread its place in every column,
// so the result does not depend on the thread count.
HCsr transpose(const HCsr &A)
{
    HCsr T;
    T.nr = A.nc;
    T.nc = A.nr;
    T.rp.assign(T.nr + 1, 0);
    T.ci.resize(A.ci.size());
    T.v.resize(A.v.size());
    const int nt = std::max(1, std::min<int>(omp_get_max_threads(), (int)(A.nr / 4096) + 1));
    std::vector<int64_t> cnt((size_t)nt * T.nr, 0);
    auto range = [&](int t, int64_t &r0, int64_t &r1) {
        r0 = A.nr * t / nt;
        r1 = A.nr * (t + 1) / nt;
    };
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        int64_t r0, r1;
        range(t, r0, r1);
        int64_t *c = cnt.data() + (size_t)t * T.nr;
        for (int64_t k = A.rp[r0]; k < A.rp[r1]; ++k) c[A.ci[k]]++;
    }
    // column c of thread t starts after column c of threads < t
    for (int64_t c = 0; c < T.nr; ++c) {
        int64_t o = 0;
        for (int t = 0; t < nt; ++t) {
            const int64_t x = cnt[(size_t)t * T.nr + c];
            cnt[(size_t)t * T.nr + c] = o;
            o += x;
        }
        T.rp[c + 1] = o;
    }
    for (int64_t c = 0; c < T.nr; ++c) T.rp[c + 1] += T.rp[c];
#pragma omp parallel for num_threads(nt) schedule(static, 1)
    for (int t = 0; t < nt; ++t) {
        int64_t r0, r1;
        range(t, r0, r1);
        int64_t *f = cnt.data() + (size_t)t * T.nr;
        for (int64_t r = r0; r < r1; ++r)
            for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
                const int32_t c = A.ci[k];
                const int64_t d = T.rp[c] + f[c]++;
                T.ci[d] = (int32_t)r;
                T.v[d] = A.v[k];
            }
    }
    return T;
}

// C = A*B, Gustavson with a dense accumulator per thread; columns sorted per
// row.  Rows are independent, so the result does not depend on the thread
// count.
HCsr spgemm(const HCsr &A, const HCsr &B)
{
    HCsr C;
    C.nr = A.nr;
    C.nc = B.nc;
    C.rp.assign(C.nr + 1, 0);
#pragma omp parallel
    {
        std::vector<int64_t> mark(B.nc, -1);
#pragma omp for schedule(dynamic, 512)
        for (int64_t r = 0; r < A.nr; ++r) {
            int64_t c = 0;
            for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
                const int32_t j = A.ci[k];
                for (int64_t t = B.rp[j]; t < B.rp[j + 1]; ++t)
                    if (mark[B.ci[t]] != r) {
                        mark[B.ci[t]] = r;
                        c++;
                    }
            }
            C.rp[r + 1] = c;
        }
    }
    for (int64_t r = 0; r < C.nr; ++r) C.rp[r + 1] += C.rp[r];
    C.ci.resize(C.rp[C.nr]);
    C.v.resize(C.rp[C.nr]);
#pragma omp parallel
    {
        std::vector<int64_t> mark(B.nc, -1);
        std::vector<double> acc(B.nc, 0.0);
        std::vector<int32_t> cols;
#pragma omp for schedule(dynamic, 512)
        for (int64_t r = 0; r < A.nr; ++r) {
            cols.clear();
            for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
                const int32_t j = A.ci[k];
                const double a = A.v[k];
                for (int64_t t = B.rp[j]; t < B.rp[j + 1]; ++t) {
                    const int32_t c = B.ci[t];
                    if (mark[c] != r) {
                        mark[c] = r;
                        acc[c] = 0.0;
                        cols.push_back(c);
                    }
                    acc[c] += a * B.v[t];
                }
            }
            std::sort(cols.begin(), cols.end());
            int64_t o = C.rp[r];
            for (int32_t c : cols) {
                C.ci[o] = c;
                C.v[o++] = acc[c];
            }
        }
    }
    return C;
}

// Same sparsity as `C` (computed before), new values.
void spgemm_numeric(const HCsr &A, const HCsr &B, HCsr &C)
{
    std::vector<double> acc(B.nc, 0.0);
    for (int64_t r = 0; r < A.nr; ++r) {
        for (int64_t k = C.rp[r]; k < C.rp[r + 1]; ++k) acc[C.ci[k]] = 0.0;
        for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) {
            const int32_t j = A.ci[k];
            const double a = A.v[k];
            for (int64_t t = B.rp[j]; t < B.rp[j + 1]; ++t) acc[B.ci[t]] += a * B.v[t];
        }
        for (int64_t k = C.rp[r]; k < C.rp[r + 1]; ++k) C.v[k] = acc[C.ci[k]];
    }
}

void spmv_host(const HCsr &A, const double *x, double *y)
{
#pragma omp parallel for schedule(static)
    for (int64_t r = 0; r < A.nr; ++r) {
        double s = 0.0;
        for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) s += A.v[k] * x[A.ci[k]];
        y[r] = s;
    }
}

template <class T>
int dupload(T **d, const std::vector<T> &h)
{
    (void)hipFree(*d);
    *d = nullptr;
    SSR_HIP_CHECK(hipMalloc((void **)d, sizeof(T) * (h.empty() ? 1 : h.size())));
    if (!h.empty()) {
        SSR_HIP_CHECK(hipMemcpyAsync(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    }
    return SSR_OK;
}

// Device copy of a level matrix: scalar CSR, or block CSR with dense rb x cb
// blocks when its rows come in whole nodes of rb unknowns and its columns in
// whole nodes of cb (the coarse unknowns of SA-AMG are aggregate x mode, and
// every aggregate of a level usually keeps all its modes).  The blocks hold
// the CSR entries at their places and zeros elsewhere; each row's sum runs
// over the same columns in the same order.
template <class T>
struct DCsr {
    int64_t nr = 0, nnz = 0;
    int64_t *rp = nullptr;
    int32_t *ci = nullptr;
    T *v = nullptr;
    int rb = 0, cb = 0;  // block shape (0: scalar CSR)
    int64_t nbr = 0, nblk = 0;
    int32_t *brp = nullptr, *bcol = nullptr;
    T *bval = nullptr;
    ~DCsr() { release(); }
    void release()
    {
        (void)hipFree(rp);
        (void)hipFree(ci);
        (void)hipFree(v);
        (void)hipFree(brp);
        (void)hipFree(bcol);
        (void)hipFree(bval);
        rp = nullptr;
        ci = nullptr;
        v = nullptr;
        brp = bcol = nullptr;
        bval = nullptr;
        rb = cb = 0;
    }
    int upload(const HCsr &A, int row_block = 0, int col_block = 0)
    {
        release();
        nr = A.nr;
        nnz = (int64_t)A.ci.size();
        if (row_block > 0 && col_block > 0 && ssr::brc_supported(row_block, col_block) && A.nr % row_block == 0 &&
            A.nc % col_block == 0 && A.nr / row_block < ((int64_t)1 << 31))
            return upload_blocks(A, row_block, col_block);
        int rc;
        if ((rc = dupload(&rp, A.rp)) || (rc = dupload(&ci, A.ci))) return rc;
        std::vector<T> vt(A.v.begin(), A.v.end());
        return dupload(&v, vt);
    }
    int upload_blocks(const HCsr &A, int r_b, int c_b)
    {
        const int64_t n_br = A.nr / r_b;
        std::vector<int32_t> hrp(n_br + 1, 0), hcol;
        std::vector<std::vector<int32_t>> cols(n_br);
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t I = 0; I < n_br; ++I) {
            auto &c = cols[I];
            for (int64_t r = I * r_b; r < (I + 1) * r_b; ++r)
                for (int64_t t = A.rp[r]; t < A.rp[r + 1]; ++t) c.push_back(A.ci[t] / c_b);
            std::sort(c.begin(), c.end());
            c.erase(std::unique(c.begin(), c.end()), c.end());
        }
        for (int64_t I = 0; I < n_br; ++I) hrp[I + 1] = hrp[I] + (int32_t)cols[I].size();
        const int64_t nb = hrp[n_br];
        hcol.resize(nb);
        std::vector<T> hv((size_t)nb * r_b * c_b, T(0));
#pragma omp parallel for schedule(dynamic, 256)
        for (int64_t I = 0; I < n_br; ++I) {
            const auto &c = cols[I];
            std::copy(c.begin(), c.end(), hcol.begin() + hrp[I]);
            for (int64_t r = I * r_b; r < (I + 1) * r_b; ++r)
                for (int64_t t = A.rp[r]; t < A.rp[r + 1]; ++t) {
                    const int32_t J = A.ci[t] / c_b;
                    const int64_t k = std::lower_bound(c.begin(), c.end(), J) - c.begin() + hrp[I];
                    hv[(size_t)k * r_b * c_b + (r - I * r_b) * c_b + (A.ci[t] - (int64_t)J * c_b)] = (T)A.v[t];
                }
        }
        int rc;
        if ((rc = dupload(&brp, hrp)) || (rc = dupload(&bcol, hcol)) || (rc = dupload(&bval, hv))) return rc;
        rb = r_b;
        cb = c_b;
        nbr = n_br;
        nblk = nb;
        return SSR_OK;
    }
    // take ownership of device block arrays built elsewhere
    void adopt(int r_b, int c_b, int64_t n_br, int64_t n_blk, int32_t *d_brp, int32_t *d_bcol, T *d_bval)
    {
        release();
        rb = r_b;
        cb = c_b;
        nbr = n_br;
        nblk = n_blk;
        nr = n_br * r_b;
        nnz = n_blk * r_b * c_b;
        brp = d_brp;
        bcol = d_bcol;
        bval = d_bval;
    }
    // a copy in precision T of a block-form matrix
    template <class S>
    int mirror(const DCsr<S> &src)
    {
        release();
        if (!src.rb) {
            ssr_set_error("AMG: float mirror of a scalar-CSR level");
            return SSR_ERR_STATE;
        }
        SSR_HIP_CHECK(hipMalloc((void **)&brp, sizeof(int32_t) * (src.nbr + 1)));
        SSR_HIP_CHECK(hipMalloc((void **)&bcol, sizeof(int32_t) * (src.nblk > 0 ? src.nblk : 1)));
        SSR_HIP_CHECK(hipMalloc((void **)&bval, sizeof(T) * (src.nblk > 0 ? src.nblk : 1) * src.rb * src.cb));
        SSR_HIP_CHECK(hipMemcpyAsync(brp, src.brp, sizeof(int32_t) * (src.nbr + 1), hipMemcpyDeviceToDevice,
                                     ssr_stream()));
        if (src.nblk)
            SSR_HIP_CHECK(hipMemcpyAsync(bcol, src.bcol, sizeof(int32_t) * src.nblk, hipMemcpyDeviceToDevice,
                                         ssr_stream()));
        rb = src.rb;
        cb = src.cb;
        nbr = src.nbr;
        nblk = src.nblk;
        nr = src.nr;
        nnz = src.nnz;
        return src.nblk ? ssr::convert<S, T>(src.nblk * src.rb * src.cb, 1.0, src.bval, bval) : SSR_OK;
    }
    int spmv(const T *x, T *y, double a, double b, const T *z = nullptr) const
    {
        if (rb) return ssr::spmv_brc<T>(rb, cb, nbr, brp, bcol, bval, x, y, a, b, z, nblk);
        return ssr::spmv_csr<T>(nr, rp, ci, v, x, y, a, b, z, nnz);
    }
};

// unknowns per node when every node of the level has the same number, else 0
inline int uniform_node_size(const std::vector<int64_t> &node_ptr)
{
    if (node_ptr.size() < 2) return 0;
    const int64_t d = node_ptr[1] - node_ptr[0];
    for (size_t i = 1; i + 1 < node_ptr.size(); ++i)
        if (node_ptr[i + 1] - node_ptr[i] != d) return 0;
    return d > 0 && d <= 8 ? (int)d : 0;
}

// Device copy of one level in precision T.
template <class T>
struct DevLevel {
    DCsr<T> A, P, R;
    T *dinv = nullptr, *x = nullptr, *b = nullptr, *r = nullptr, *d = nullptr;
    ~DevLevel()
    {
        (void)hipFree(dinv);
        (void)hipFree(x);
        (void)hipFree(b);
        (void)hipFree(r);
        (void)hipFree(d);
    }
};

struct Level {
    HCsr A;                        // operator (host)
    HCsr P, R;                     // prolongation to this level from the next coarser / restriction
    std::vector<int64_t> node_ptr; // dofs of node i: node_ptr[i]..node_ptr[i+1]
    std::vector<double> dinv;
    double rho = 1.0;    // power-iteration estimate of rho(D^-1 A) (prolongator smoothing)
    double rho_up = 1.0; // Gershgorin bound of rho(D^-1 A) (Chebyshev interval)
    DevLevel<double> d64;
    DevLevel<float> d32;
    template <class T> DevLevel<T> &dev();
};
template <> DevLevel<double> &Level::dev<double>() { return d64; }
template <> DevLevel<float> &Level::dev<float>() { return d32; }

// max_i sum_j |a_ij| |dinv_i|: an upper bound of the spectral radius of D^-1 A
double gershgorin_rho(const HCsr &A, const std::vector<double> &dinv)
{
    double m = 0.0;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (int64_t r = 0; r < A.nr; ++r) {
        double s = 0.0;
        for (int64_t k = A.rp[r]; k < A.rp[r + 1]; ++k) s += fabs(A.v[k]);
        m = std::max(m, s * fabs(dinv[r]));
    }
    return m > 0 ? m : 1.0;
}

// sum_i f(i) over fixed chunks of rows, the chunk sums added in order: the
// same value for any thread count
template <class F>
double chunked_sum(int64_t n, F f)
{
    constexpr int64_t CH = 16384;
    const int64_t nc = (n + CH - 1) / CH;
    std::vector<double> part(nc > 0 ? nc : 1, 0.0);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) {
        double a = 0.0;
        const int64_t e = std::min(n, (c + 1) * CH);
        for (int64_t i = c * CH; i < e; ++i) a += f(i);
        part[c] = a;
    }
    double t = 0.0;
    for (int64_t c = 0; c < nc; ++c) t += part[c];
    return t;
}

double power_rho(const HCsr &A, const std::vector<double> &dinv, int its)
{
    const int64_t n = A.nr;
    std::vector<double> x(n), y(n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < n; ++i) x[i] = 1.0 + 0.1 * (double)((i * 7919) % 13);
    double lam = 0.0;
    for (int k = 0; k < its; ++k) {
        const double nx = sqrt(chunked_sum(n, [&](int64_t i) { return x[i] * x[i]; }));
        if (nx == 0.0) return 1.0;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) x[i] /= nx;
        spmv_host(A, x.data(), y.data());
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < n; ++i) y[i] *= dinv[i];
        lam = chunked_sum(n, [&](int64_t i) { return x[i] * y[i]; });
        x.swap(y);
    }
    return lam > 0 ? lam : 1.0;
}

struct Options {
    double theta = 0.0;     // strength threshold on node blocks
    int max_levels = 12;
    int64_t max_coarse = 600;
    int cheb_degree = 2;
    double cheb_lower = 1.0 / 30.0;
    double cheb_upper = 1.1;
    int refresh_every = 1;  // Galerkin refresh every n-th setup call (aggregates/P kept)
    int rebuild_every = 0;  // full rebuild every n-th setup call (0 = never after the first)
};

}  // namespace

struct Amg : ssr_precond {
    const ssr_pattern *Pat;
    int dim;
    int64_t n_n, n, nnzb;
    std::vector<int32_t> row_ptr, col;
    std::vector<uint8_t> mask;
    std::vector<double> B;  // near-null space, n x k column-major
    int k = 0;
    Options opt;
    std::vector<std::unique_ptr<Level>> L;
    std::vector<double> coarse_inv;
    double *d_coarse_inv = nullptr;
    float *d_coarse_inv32 = nullptr;
    int64_t n_coarse = 0;
    const double *Kdev = nullptr;
    const float *Kdev32 = nullptr;
    bool want32 = false;  // keep a float copy of the device hierarchy
    // finest-level smoother scaled by the node-block diagonal inverses of the
    // current operator (refreshed at every setup) instead of the scalar diagonal
    bool block0 = false;
    double rho_up_block = 1.0;  // Gershgorin bound of rho(D_B^-1 A) on the finest level
    double *binv64 = nullptr;
    float *binv32 = nullptr;
    // page-locked host copy of the fine operator's block values (downloaded
    // at each setup; kept for the next one)
    struct PinnedVals {
        double *p = nullptr;
        size_t n = 0;
        int reserve(size_t m)
        {
            if (m <= n) return SSR_OK;
            if (p) (void)hipHostFree(p);
            p = nullptr;
            n = 0;
            SSR_HIP_CHECK(hipHostMalloc((void **)&p, sizeof(double) * m, hipHostMallocDefault));
            n = m;
            return SSR_OK;
        }
        double &operator[](size_t i) { return p[i]; }
        const double &operator[](size_t i) const { return p[i]; }
        double *data() { return p; }
        const double *data() const { return p; }
        size_t size() const { return n; }
        ~PinnedVals()
        {
            if (p) (void)hipHostFree(p);
        }
    } Kh;
    int64_t n_setup = 0;
    int64_t generation = 0;  // bumped whenever device level data is (re)uploaded
    bool fine_sym = false;   // finest-level values symmetric bit for bit (mirror SpMV allowed)
    bool built = false;
    bool on_device = false;  // the hierarchy was built on the device (amg_device.hip)
    double *d_B = nullptr;   // device copy of the near-null space (device build)
    double dev_phase[8] = {0};
    bool force_rebuild = false;  // set by invalidate(): rebuild at the next setup

    Amg(const ssr_pattern *P, const double *near_null, int n_null, const Options &o) : Pat(P), opt(o)
    {
        int64_t n_p, n_e, nnzf;
        ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
        n = dim * n_n;
        row_ptr.resize(n_n + 1);
        col.resize(nnzb);
        mask.resize(n);
        ssr_pattern_get(P, 0, row_ptr.data());
        ssr_pattern_get(P, 1, col.data());
        ssr_pattern_get(P, 6, mask.data());
        k = n_null;
        block0 = ssr_tuning(SSR_TUNE_AMG_BLOCK_SMOOTH) != 0;
        B.assign(near_null, near_null + (size_t)n * k);
        for (int c = 0; c < k; ++c)
            for (int64_t i = 0; i < n; ++i)
                if (!mask[i]) B[(size_t)c * n + i] = 0.0;
    }
    ~Amg()
    {
        (void)hipFree(d_coarse_inv);
        (void)hipFree(d_coarse_inv32);
        (void)hipFree(binv64);
        (void)hipFree(binv32);
        (void)hipFree(d_B);
        vg64.reset();
        vg32.reset();
        if (cap_stream) (void)hipStreamDestroy(cap_stream);
    }

    // scalar CSR of the downloaded node-block operator (zeros dropped, the
    // diagonal kept): count per row, prefix sum, fill, rows in parallel
    void fine_from_bsr(HCsr &A)
    {
        const int bs = dim * dim;
        A.nr = A.nc = n;
        A.rp.assign(n + 1, 0);
        auto keep = [&](int64_t r, int i, int32_t b, int j, double v) { return v != 0.0 || (col[b] == r && i == j); };
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < n_n; ++r)
            for (int i = 0; i < dim; ++i) {
                int64_t c = 0;
                for (int32_t b = row_ptr[r]; b < row_ptr[r + 1]; ++b)
                    for (int j = 0; j < dim; ++j) c += keep(r, i, b, j, Kh[(size_t)b * bs + i * dim + j]) ? 1 : 0;
                A.rp[r * dim + i + 1] = c;
            }
        for (int64_t q = 0; q < n; ++q) A.rp[q + 1] += A.rp[q];
        A.ci.resize(A.rp[n]);
        A.v.resize(A.rp[n]);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t r = 0; r < n_n; ++r)
            for (int i = 0; i < dim; ++i) {
                int64_t o = A.rp[r * dim + i];
                for (int32_t b = row_ptr[r]; b < row_ptr[r + 1]; ++b)
                    for (int j = 0; j < dim; ++j) {
                        const double v = Kh[(size_t)b * bs + i * dim + j];
                        if (keep(r, i, b, j, v)) {
                            A.ci[o] = col[b] * dim + j;
                            A.v[o++] = v;
                        }
                    }
            }
    }

    // max over rows of sum_j |(D_B^-1 A)_ij| with D_B the node diagonal blocks
    // of the downloaded operator Kh (an upper bound of rho(D_B^-1 A))
    double block_gershgorin() const
    {
        const int bs = dim * dim;
        double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
        for (int64_t r = 0; r < n_n; ++r) {
            std::array<double, 9> Bi{};
            std::array<double, 3> acc{};
            const double *D = nullptr;
            for (int32_t b = row_ptr[r]; b < row_ptr[r + 1]; ++b)
                if (col[b] == r) D = Kh.data() + (size_t)b * bs;
            if (!D) continue;
            if (dim == 2) {
                const double det = D[0] * D[3] - D[1] * D[2];
                Bi = {D[3] / det, -D[1] / det, -D[2] / det, D[0] / det};
            } else {
                const double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8],
                             c02 = D[3] * D[7] - D[4] * D[6];
                const double id = 1.0 / (D[0] * c00 + D[1] * c01 + D[2] * c02);
                Bi = {c00 * id, (D[2] * D[7] - D[1] * D[8]) * id, (D[1] * D[5] - D[2] * D[4]) * id,
                      c01 * id, (D[0] * D[8] - D[2] * D[6]) * id, (D[2] * D[3] - D[0] * D[5]) * id,
                      c02 * id, (D[1] * D[6] - D[0] * D[7]) * id, (D[0] * D[4] - D[1] * D[3]) * id};
            }
            std::fill(acc.begin(), acc.end(), 0.0);
            for (int32_t b = row_ptr[r]; b < row_ptr[r + 1]; ++b) {
                const double *A = Kh.data() + (size_t)b * bs;
                for (int i = 0; i < dim; ++i)
                    for (int j = 0; j < dim; ++j) {
                        double v = 0.0;
                        for (int q = 0; q < dim; ++q) v += Bi[i * dim + q] * A[q * dim + j];
                        acc[i] += fabs(v);
                    }
            }
            for (int i = 0; i < dim; ++i) mx = std::max(mx, acc[i]);
        }
        return mx > 0 ? mx : 1.0;
    }

    static void diag_inv(const HCsr &A, std::vector<double> &d)
    {
        d.assign(A.nr, 1.0);
#pragma omp parallel for schedule(static)
        for (int64_t r = 0; r < A.nr; ++r)
            for (int64_t t = A.rp[r]; t < A.rp[r + 1]; ++t)
                if (A.ci[t] == r) d[r] = A.v[t] != 0.0 ? 1.0 / A.v[t] : 1.0;
    }

    // Aggregate the nodes of level l; returns aggregate id per node (-1 = none).
    std::vector<int32_t> aggregate(const Level &lv, int32_t &n_agg)
    {
        const HCsr &A = lv.A;
        const int64_t nn = (int64_t)lv.node_ptr.size() - 1;
        std::vector<int32_t> node_of(A.nr);
        for (int64_t i = 0; i < nn; ++i)
            for (int64_t d = lv.node_ptr[i]; d < lv.node_ptr[i + 1]; ++d) node_of[d] = (int32_t)i;
        // node diagonal block norms (each node sums its own rows, in order)
        std::vector<double> dn(nn, 0.0);
#pragma omp parallel for schedule(dynamic, 1024)
        for (int64_t i = 0; i < nn; ++i) {
            double s = 0.0;
            for (int64_t r = lv.node_ptr[i]; r < lv.node_ptr[i + 1]; ++r)
                for (int64_t t = A.rp[r]; t < A.rp[r + 1]; ++t)
                    if (node_of[A.ci[t]] == i) s += A.v[t] * A.v[t];
            dn[i] = s;
        }
        // strong node graph: nodes in parallel, counted first, then filled
        std::vector<int64_t> sp(nn + 1, 0);
        std::vector<int32_t> sc;
        auto strong = [&](int64_t i, std::vector<double> &acc, std::vector<int64_t> &mk, std::vector<int32_t> &nb,
                          std::vector<int32_t> &out) {
            nb.clear();
            out.clear();
            for (int64_t d = lv.node_ptr[i]; d < lv.node_ptr[i + 1]; ++d)
                for (int64_t t = A.rp[d]; t < A.rp[d + 1]; ++t) {
                    const int32_t j = node_of[A.ci[t]];
                    if (j == i) continue;
                    if (mk[j] != i) {
                        mk[j] = i;
                        acc[j] = 0.0;
                        nb.push_back(j);
                    }
                    acc[j] += A.v[t] * A.v[t];
                }
            for (int32_t j : nb)
                if (acc[j] > 0.0 && acc[j] >= opt.theta * opt.theta * sqrt(dn[i] * dn[j])) out.push_back(j);
        };
        for (int pass = 0; pass < 2; ++pass) {
            if (pass == 1) {
                for (int64_t i = 0; i < nn; ++i) sp[i + 1] += sp[i];
                sc.resize(sp[nn]);
            }
#pragma omp parallel
            {
                std::vector<double> acc(nn, 0.0);
                std::vector<int64_t> mk(nn, -1);
                std::vector<int32_t> nb, out;
#pragma omp for schedule(dynamic, 1024)
                for (int64_t i = 0; i < nn; ++i) {
                    strong(i, acc, mk, nb, out);
                    if (pass == 0) sp[i + 1] = (int64_t)out.size();
                    else std::copy(out.begin(), out.end(), sc.begin() + sp[i]);
                }
            }
        }
        return ssr_amg::aggregate_graph(nn, sp, sc, n_agg);
    }

    // Tentative prolongator of level lv given aggregates; returns coarse
    // near-null space and coarse node pointers.
    void tentative(const Level &lv, const std::vector<double> &Bl, int kl, const std::vector<int32_t> &agg,
                   int32_t n_agg, HCsr &T, std::vector<double> &Bc, std::vector<int64_t> &cnode_ptr)
    {
        const int64_t nn = (int64_t)lv.node_ptr.size() - 1;
        const int64_t nf = lv.A.nr;
        // the fine unknowns of each aggregate, in node order
        std::vector<int64_t> dptr(n_agg + 1, 0), dofs;
        for (int64_t i = 0; i < nn; ++i)
            if (agg[i] >= 0) dptr[agg[i] + 1] += lv.node_ptr[i + 1] - lv.node_ptr[i];
        for (int32_t a = 0; a < n_agg; ++a) dptr[a + 1] += dptr[a];
        dofs.resize(dptr[n_agg]);
        {
            std::vector<int64_t> f(dptr.begin(), dptr.end() - 1);
            for (int64_t i = 0; i < nn; ++i)
                if (agg[i] >= 0)
                    for (int64_t d = lv.node_ptr[i]; d < lv.node_ptr[i + 1]; ++d) dofs[f[agg[i]]++] = d;
        }
        // per aggregate (independent, in parallel): the near-null space
        // restricted to it, orthonormalised column by column (CGS2) with the
        // dependent columns dropped; Q (m x kept) and the kept rows of R
        std::vector<std::vector<double>> Qs(n_agg), Rs(n_agg);
        std::vector<int> kept_of(n_agg, 0);
#pragma omp parallel for schedule(dynamic, 64)
        for (int32_t a = 0; a < n_agg; ++a) {
            const int64_t *D = dofs.data() + dptr[a];
            const int64_t m = dptr[a + 1] - dptr[a];
            std::vector<double> Qm((size_t)m * kl, 0.0), v(m);
            std::vector<std::vector<double>> R;  // rows of R for kept columns
            int kept = 0;
            for (int c = 0; c < kl; ++c) {
                double n0 = 0.0;
                for (int64_t i = 0; i < m; ++i) {
                    v[i] = Bl[(size_t)c * nf + D[i]];
                    n0 += v[i] * v[i];
                }
                n0 = sqrt(n0);
                std::vector<double> rcol(kept, 0.0);
                for (int pass = 0; pass < 2; ++pass)
                    for (int q = 0; q < kept; ++q) {
                        double s = 0.0;
                        for (int64_t i = 0; i < m; ++i) s += Qm[(size_t)q * m + i] * v[i];
                        for (int64_t i = 0; i < m; ++i) v[i] -= s * Qm[(size_t)q * m + i];
                        rcol[q] += s;
                    }
                double nv = 0.0;
                for (double t : v) nv += t * t;
                nv = sqrt(nv);
                // R entries for this column: rcol (kept rows), nv for the new row
                for (int q = 0; q < kept; ++q) R[q][c] = rcol[q];
                if (nv > 1e-10 * n0 && nv > 1e-300) {
                    for (int64_t i = 0; i < m; ++i) Qm[(size_t)kept * m + i] = v[i] / nv;
                    R.emplace_back(kl, 0.0);
                    R.back()[c] = nv;
                    kept++;
                }
            }
            Qm.resize((size_t)kept * m);
            Qs[a].swap(Qm);
            std::vector<double> rr;
            rr.reserve((size_t)kept * kl);
            for (int q = 0; q < kept; ++q) rr.insert(rr.end(), R[q].begin(), R[q].end());
            Rs[a].swap(rr);
            kept_of[a] = kept;
        }
        // coarse numbering in aggregate order
        std::vector<int64_t> cofs(n_agg + 1, 0);
        for (int32_t a = 0; a < n_agg; ++a) cofs[a + 1] = cofs[a] + kept_of[a];
        const int64_t nc = cofs[n_agg];
        cnode_ptr.assign(cofs.begin(), cofs.end());
        std::vector<double> bcv((size_t)nc * kl);
        for (int32_t a = 0; a < n_agg; ++a) std::copy(Rs[a].begin(), Rs[a].end(), bcv.begin() + cofs[a] * kl);
        // T: fine unknown D[i] of aggregate a has entries Q(i, q) at columns cofs[a] + q
        std::vector<int32_t> agg_of(nf, -1), pos_of(nf, 0);
        for (int32_t a = 0; a < n_agg; ++a)
            for (int64_t i = dptr[a]; i < dptr[a + 1]; ++i) {
                agg_of[dofs[i]] = a;
                pos_of[dofs[i]] = (int32_t)(i - dptr[a]);
            }
        T.nr = nf;
        T.nc = nc;
        T.rp.assign(nf + 1, 0);
        for (int64_t d = 0; d < nf; ++d) T.rp[d + 1] = T.rp[d] + (agg_of[d] >= 0 ? kept_of[agg_of[d]] : 0);
        T.ci.resize(T.rp[nf]);
        T.v.resize(T.rp[nf]);
#pragma omp parallel for schedule(static)
        for (int64_t d = 0; d < nf; ++d) {
            const int32_t a = agg_of[d];
            if (a < 0) continue;
            const int64_t m = dptr[a + 1] - dptr[a];
            for (int q = 0; q < kept_of[a]; ++q) {
                T.ci[T.rp[d] + q] = (int32_t)(cofs[a] + q);
                T.v[T.rp[d] + q] = Qs[a][(size_t)q * m + pos_of[d]];
            }
        }
        // Bc column-major nc x kl from row-major bcv
        Bc.assign((size_t)nc * kl, 0.0);
        for (int64_t r = 0; r < nc; ++r)
            for (int c = 0; c < kl; ++c) Bc[(size_t)c * nc + r] = bcv[(size_t)r * kl + c];
    }

    template <class T>
    int alloc_level_vectors(Level &lv)
    {
        const int64_t m = lv.A.nr > 0 ? lv.A.nr : 1;
        DevLevel<T> &dl = lv.dev<T>();
        T **ps[4] = {&dl.x, &dl.b, &dl.r, &dl.d};
        for (auto p : ps) {
            (void)hipFree(*p);
            *p = nullptr;
            SSR_HIP_CHECK(hipMalloc((void **)p, sizeof(T) * m));
        }
        return SSR_OK;
    }

    // Device copy of level l (operator except on the finest level, P, R,
    // D^-1, work vectors) in precision T.
    template <class T>
    int upload_level(size_t l)
    {
        generation++;
        Level &lv = *L[l];
        DevLevel<T> &dl = lv.dev<T>();
        int rc;
        const int d_l = ssr_tuning(SSR_TUNE_AMG_COARSE_BLOCKS) ? uniform_node_size(lv.node_ptr) : 0;
        if (l > 0 && (rc = dl.A.upload(lv.A, d_l, d_l))) return rc;
        if (l + 1 < L.size()) {
            const int d_c = d_l ? uniform_node_size(L[l + 1]->node_ptr) : 0;
            if ((rc = dl.P.upload(lv.P, d_l, d_c)) || (rc = dl.R.upload(lv.R, d_c, d_l))) return rc;
        }
        std::vector<T> di(lv.dinv.begin(), lv.dinv.end());
        if ((rc = dupload(&dl.dinv, di))) return rc;
        return alloc_level_vectors<T>(lv);
    }

    PhaseClock clk;

    // float copy of a hierarchy built on the device
    int mirror_float(size_t l)
    {
        generation++;
        Level &lv = *L[l];
        DevLevel<double> &d = lv.d64;
        DevLevel<float> &f = lv.d32;
        int rc;
        if (l > 0 && (rc = f.A.mirror(d.A))) return rc;
        if (l + 1 < L.size() && ((rc = f.P.mirror(d.P)) || (rc = f.R.mirror(d.R)))) return rc;
        (void)hipFree(f.dinv);
        f.dinv = nullptr;
        SSR_HIP_CHECK(hipMalloc((void **)&f.dinv, sizeof(float) * (lv.A.nr > 0 ? lv.A.nr : 1)));
        if ((rc = ssr::convert<double, float>(lv.A.nr, 1.0, d.dinv, f.dinv))) return rc;
        return alloc_level_vectors<float>(lv);
    }

    // The hierarchy on the device (amg_device.hip): the same algorithm and
    // aggregates as build(), every level in block CSR.  *fallback is set
    // when an aggregate loses a near-null vector (non-uniform coarse
    // unknowns) or the problem is too small for a second level; build() then
    // runs instead.
    int build_device(bool *fallback)
    {
        *fallback = false;
        const auto t_start = std::chrono::steady_clock::now();
        int rc;
        if (!d_B) {
            SSR_HIP_CHECK(hipMalloc((void **)&d_B, sizeof(double) * B.size()));
            SSR_HIP_CHECK(hipMemcpyAsync(d_B, B.data(), sizeof(double) * B.size(), hipMemcpyHostToDevice,
                                         ssr_stream()));
        }
        std::vector<ssr_amg::DevLevelOut> out;
        ssr_amg::DevOptions o;
        o.theta = opt.theta;
        o.max_levels = opt.max_levels;
        o.max_coarse = opt.max_coarse;
        rc = ssr_amg::amg_device_build(dim, n_n, ssr_pattern_d_row_ptr(Pat), ssr_pattern_d_col(Pat),
                                       ssr_pattern_d_brow(Pat), ssr_pattern_d_diag(Pat), Kdev, row_ptr, col, d_B, k,
                                       o, out, dev_phase);
        if (rc == ssr_amg::kNotUniform || (rc == SSR_OK && out.size() < 2)) {
            for (auto &l : out) l.release();
            *fallback = true;
            return SSR_OK;
        }
        if (rc) return rc;
        L.clear();
        for (size_t l = 0; l < out.size(); ++l) {
            ssr_amg::DevLevelOut &o_l = out[l];
            L.emplace_back(new Level());
            Level &lv = *L.back();
            const int64_t nl = (int64_t)o_l.rb * o_l.n_nodes;
            lv.A.nr = lv.A.nc = nl;
            lv.rho = o_l.rho;
            lv.rho_up = o_l.rho_up;
            lv.node_ptr.resize(o_l.n_nodes + 1);
            for (int64_t i = 0; i <= o_l.n_nodes; ++i) lv.node_ptr[i] = i * o_l.rb;
            DevLevel<double> &dl = lv.d64;
            if (l > 0) {
                dl.A.adopt(o_l.A.rb, o_l.A.cb, o_l.A.nbr, o_l.A.nblk, o_l.A.brp, o_l.A.bcol, o_l.A.bval);
                (void)hipFree(o_l.A.brow);
                o_l.A = ssr_amg::DevBlocks();
            }
            if (l + 1 < out.size()) {
                dl.P.adopt(o_l.P.rb, o_l.P.cb, o_l.P.nbr, o_l.P.nblk, o_l.P.brp, o_l.P.bcol, o_l.P.bval);
                dl.R.adopt(o_l.R.rb, o_l.R.cb, o_l.R.nbr, o_l.R.nblk, o_l.R.brp, o_l.R.bcol, o_l.R.bval);
                (void)hipFree(o_l.P.brow);
                (void)hipFree(o_l.R.brow);
                o_l.P = ssr_amg::DevBlocks();
                o_l.R = ssr_amg::DevBlocks();
            }
            (void)hipFree(dl.dinv);
            dl.dinv = o_l.dinv;
            o_l.dinv = nullptr;
            (void)hipFree(o_l.diag);
            o_l.diag = nullptr;
            if ((rc = alloc_level_vectors<double>(lv))) return rc;
        }
        generation++;
        // the coarsest operator on the host for the dense inverse
        {
            Level &c = *L.back();
            const DCsr<double> &Ad = c.d64.A;
            const int rbk = Ad.rb, cbk = Ad.cb;
            std::vector<int32_t> brp(Ad.nbr + 1), bcol(Ad.nblk);
            std::vector<double> bval((size_t)Ad.nblk * rbk * cbk);
            SSR_HIP_CHECK(hipMemcpyAsync(brp.data(), Ad.brp, sizeof(int32_t) * brp.size(), hipMemcpyDeviceToHost,
                                         ssr_stream()));
            SSR_HIP_CHECK(hipMemcpyAsync(bcol.data(), Ad.bcol, sizeof(int32_t) * bcol.size(), hipMemcpyDeviceToHost,
                                         ssr_stream()));
            SSR_HIP_CHECK(hipMemcpyAsync(bval.data(), Ad.bval, sizeof(double) * bval.size(), hipMemcpyDeviceToHost,
                                         ssr_stream()));
            SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
            HCsr &A = c.A;
            A.nr = A.nc = Ad.nbr * rbk;
            A.rp.assign(A.nr + 1, 0);
            A.ci.clear();
            A.v.clear();
            for (int64_t I = 0; I < Ad.nbr; ++I)
                for (int r = 0; r < rbk; ++r) {
                    for (int32_t t = brp[I]; t < brp[I + 1]; ++t)
                        for (int q = 0; q < cbk; ++q) {
                            A.ci.push_back(bcol[t] * cbk + q);
                            A.v.push_back(bval[((size_t)t * rbk + r) * cbk + q]);
                        }
                    A.rp[I * rbk + r + 1] = (int64_t)A.ci.size();
                }
        }
        if (block0 &&
            (rc = ssr_amg::fine_block_gershgorin(dim, n_n, ssr_pattern_d_row_ptr(Pat), ssr_pattern_d_diag(Pat), Kdev,
                                                 &rho_up_block)))
            return rc;
        if (want32)
            for (size_t l = 0; l < L.size(); ++l)
                if ((rc = mirror_float(l))) return rc;
        if ((rc = build_coarse())) return rc;
        built = true;
        on_device = true;
        if (getenv("SSR_AMG_VERBOSE")) {
            for (size_t l = 0; l < L.size(); ++l)
                fprintf(stderr, "[amg] device level %zu: n=%lld blocks=%lld rho=%.6g rho_up=%.6g\n", l,
                        (long long)L[l]->A.nr, (long long)(l ? L[l]->d64.A.nblk : (int64_t)col.size()), L[l]->rho,
                        L[l]->rho_up);
            const double tot =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count();
            fprintf(stderr,
                    "[amg] device setup %.3f s (cumulative phases: dinv+rho %.3f, strength+aggregation %.3f, "
                    "tentative %.3f, P %.3f, AP %.3f, R %.3f, RAP %.3f)\n",
                    tot, dev_phase[0], dev_phase[1], dev_phase[2], dev_phase[3], dev_phase[4], dev_phase[5],
                    dev_phase[6]);
        }
        return SSR_OK;
    }

    // build on the device when enabled, else (or when the device build
    // declines) on the host from a downloaded copy of the operator
    int build_any()
    {
        if (ssr_tuning(SSR_TUNE_AMG_DEVICE_SETUP) && ssr_tuning(SSR_TUNE_AMG_COARSE_BLOCKS)) {
            bool fallback = false;
            int rc = build_device(&fallback);
            if (rc || !fallback) return rc;
        }
        on_device = false;
        int rc;
        if ((rc = Kh.reserve((size_t)nnzb * dim * dim))) return rc;
        SSR_HIP_CHECK(hipMemcpyAsync(Kh.data(), Kdev, sizeof(double) * (size_t)nnzb * dim * dim,
                                     hipMemcpyDeviceToHost, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        clk.lap(0);
        return build();
    }

    int build()
    {
        clk.lap(11);
        L.clear();
        L.emplace_back(new Level());
        Level &f = *L[0];
        fine_from_bsr(f.A);
        if (block0) rho_up_block = block_gershgorin();
        clk.lap(1);
        f.node_ptr.resize(n_n + 1);
        for (int64_t i = 0; i <= n_n; ++i) f.node_ptr[i] = i * dim;
        std::vector<double> Bl = B;
        int kl = k;
        int rc;
        for (int lev = 0;; ++lev) {
            Level &lv = *L[lev];
            diag_inv(lv.A, lv.dinv);
            lv.rho = power_rho(lv.A, lv.dinv, 30);
            lv.rho_up = gershgorin_rho(lv.A, lv.dinv);
            clk.lap(2);
            if (lv.A.nr <= opt.max_coarse || lev + 1 >= opt.max_levels) break;
            int32_t n_agg = 0;
            std::vector<int32_t> agg = aggregate(lv, n_agg);
            clk.lap(3);
            if (n_agg == 0) break;
            HCsr T;
            std::vector<double> Bc;
            std::vector<int64_t> cnode;
            tentative(lv, Bl, kl, agg, n_agg, T, Bc, cnode);
            clk.lap(4);
            if (T.nc == 0 || T.nc >= lv.A.nr * 0.9) break;
            // P = T - omega D^-1 A T, omega = 4/(3 rho)
            HCsr AT = spgemm(lv.A, T);
            clk.lap(5);
            const double om = 4.0 / (3.0 * lv.rho);
            HCsr P;
            {
                // P = T - om * Dinv * AT (union of the two patterns, columns
                // sorted per row); rows in parallel, counted first
                P.nr = T.nr;
                P.nc = T.nc;
                P.rp.assign(P.nr + 1, 0);
                auto row = [&](int64_t r, std::vector<double> &accv, std::vector<int64_t> &mk,
                               std::vector<int32_t> &cs) {
                    cs.clear();
                    for (int64_t t = T.rp[r]; t < T.rp[r + 1]; ++t) {
                        if (mk[T.ci[t]] != r) { mk[T.ci[t]] = r; accv[T.ci[t]] = 0.0; cs.push_back(T.ci[t]); }
                        accv[T.ci[t]] += T.v[t];
                    }
                    for (int64_t t = AT.rp[r]; t < AT.rp[r + 1]; ++t) {
                        if (mk[AT.ci[t]] != r) { mk[AT.ci[t]] = r; accv[AT.ci[t]] = 0.0; cs.push_back(AT.ci[t]); }
                        accv[AT.ci[t]] -= om * lv.dinv[r] * AT.v[t];
                    }
                    std::sort(cs.begin(), cs.end());
                };
#pragma omp parallel
                {
                    std::vector<double> accv(T.nc, 0.0);
                    std::vector<int64_t> mk(T.nc, -1);
                    std::vector<int32_t> cs;
#pragma omp for schedule(dynamic, 512)
                    for (int64_t r = 0; r < P.nr; ++r) {
                        row(r, accv, mk, cs);
                        P.rp[r + 1] = (int64_t)cs.size();
                    }
                }
                for (int64_t r = 0; r < P.nr; ++r) P.rp[r + 1] += P.rp[r];
                P.ci.resize(P.rp[P.nr]);
                P.v.resize(P.rp[P.nr]);
#pragma omp parallel
                {
                    std::vector<double> accv(T.nc, 0.0);
                    std::vector<int64_t> mk(T.nc, -1);
                    std::vector<int32_t> cs;
#pragma omp for schedule(dynamic, 512)
                    for (int64_t r = 0; r < P.nr; ++r) {
                        row(r, accv, mk, cs);
                        int64_t o = P.rp[r];
                        for (int32_t c : cs) {
                            P.ci[o] = c;
                            P.v[o++] = accv[c];
                        }
                    }
                }
            }
            clk.lap(6);
            HCsr R = transpose(P);
            clk.lap(7);
            HCsr AP = spgemm(lv.A, P);
            L.emplace_back(new Level());
            Level &c = *L[lev + 1];
            c.A = spgemm(R, AP);
            clk.lap(8);
            c.node_ptr = cnode;
            lv.P = std::move(P);
            lv.R = std::move(R);
            Bl.swap(Bc);
            kl = k;
        }
        // device upload
        for (size_t l = 0; l < L.size(); ++l) {
            if ((rc = upload_level<double>(l))) return rc;
            if (want32 && (rc = upload_level<float>(l))) return rc;
        }
        clk.lap(9);
        if ((rc = build_coarse())) return rc;
        clk.lap(10);
        built = true;
        if (getenv("SSR_AMG_VERBOSE")) {
            for (size_t l = 0; l < L.size(); ++l)
                fprintf(stderr, "[amg] level %zu: n=%lld nnz=%lld nodes=%lld rho=%.3g rho_up=%.3g\n", l,
                        (long long)L[l]->A.nr, (long long)L[l]->A.ci.size(), (long long)L[l]->node_ptr.size() - 1,
                        L[l]->rho, L[l]->rho_up);
            fprintf(stderr, "[amg] setup totals (s):");
            for (int q = 0; q < 12; ++q) fprintf(stderr, " %s=%.3f", kPhaseNames[q], clk.t[q]);
            fprintf(stderr, "\n");
        }
        return SSR_OK;
    }

    // numeric refresh: same P/R, new fine values -> new Galerkin products
    int refresh()
    {
        Level &f = *L[0];
        fine_from_bsr(f.A);
        if (block0) rho_up_block = block_gershgorin();
        int rc;
        for (size_t l = 0; l < L.size(); ++l) {
            Level &lv = *L[l];
            diag_inv(lv.A, lv.dinv);
            lv.rho = power_rho(lv.A, lv.dinv, 30);
            lv.rho_up = gershgorin_rho(lv.A, lv.dinv);
            if (l + 1 < L.size()) {
                Level &c = *L[l + 1];
                HCsr AP = spgemm(lv.A, lv.P);
                HCsr Ac = spgemm(lv.R, AP);
                c.A = std::move(Ac);
            }
        }
        for (size_t l = 0; l < L.size(); ++l) {
            if ((rc = upload_level<double>(l))) return rc;
            if (want32 && (rc = upload_level<float>(l))) return rc;
        }
        return build_coarse();
    }

    int build_coarse()
    {
        Level &c = *L.back();
        const int64_t m = c.A.nr;
        if (m > 8000) {
            ssr_set_error("AMG: coarsening stopped at %lld unknowns (too many for the dense coarse solve)",
                          (long long)m);
            return SSR_ERR_STATE;
        }
        n_coarse = m;
        std::vector<double> M((size_t)m * m, 0.0), I((size_t)m * m, 0.0);
        for (int64_t r = 0; r < m; ++r) {
            for (int64_t t = c.A.rp[r]; t < c.A.rp[r + 1]; ++t) M[(size_t)r * m + c.A.ci[t]] += c.A.v[t];
            I[(size_t)r * m + r] = 1.0;
        }
        double scale = 0.0;
        for (int64_t r = 0; r < m; ++r) scale = std::max(scale, fabs(M[(size_t)r * m + r]));
        // Gauss-Jordan with partial pivoting
        for (int64_t col = 0; col < m; ++col) {
            int64_t piv = col;
            for (int64_t r = col + 1; r < m; ++r)
                if (fabs(M[(size_t)r * m + col]) > fabs(M[(size_t)piv * m + col])) piv = r;
            if (fabs(M[(size_t)piv * m + col]) < 1e-14 * (scale > 0 ? scale : 1.0)) {
                M[(size_t)piv * m + col] = scale > 0 ? scale * 1e-14 : 1.0;  // singular direction
            }
            if (piv != col)
                for (int64_t j = 0; j < m; ++j) {
                    std::swap(M[(size_t)piv * m + j], M[(size_t)col * m + j]);
                    std::swap(I[(size_t)piv * m + j], I[(size_t)col * m + j]);
                }
            const double d = M[(size_t)col * m + col];
            for (int64_t j = 0; j < m; ++j) {
                M[(size_t)col * m + j] /= d;
                I[(size_t)col * m + j] /= d;
            }
            for (int64_t r = 0; r < m; ++r) {
                if (r == col) continue;
                const double f = M[(size_t)r * m + col];
                if (f == 0.0) continue;
                for (int64_t j = 0; j < m; ++j) {
                    M[(size_t)r * m + j] -= f * M[(size_t)col * m + j];
                    I[(size_t)r * m + j] -= f * I[(size_t)col * m + j];
                }
            }
        }
        coarse_inv.swap(I);
        generation++;
        int rc = dupload(&d_coarse_inv, coarse_inv);
        if (rc || !want32) return rc;
        std::vector<float> ci32(coarse_inv.begin(), coarse_inv.end());
        return dupload(&d_coarse_inv32, ci32);
    }

    void invalidate() override { force_rebuild = true; }
    void set_symmetric(bool v) override { fine_sym = v; }

    int setup(const double *Kvals, const float *K32) override
    {
        // The finest level always applies the current operator (Kvals / K32
        // on the device); the coarse hierarchy is rebuilt / refreshed on the
        // host only when the schedule asks for it.
        Kdev = Kvals;
        Kdev32 = K32;
        const bool need32 = K32 != nullptr && !want32;
        if (K32) want32 = true;
        const bool rebuild =
            !built || force_rebuild || (opt.rebuild_every > 0 && n_setup % opt.rebuild_every == 0);
        force_rebuild = false;
        const bool refresh_now = !rebuild && opt.refresh_every > 0 && n_setup % opt.refresh_every == 0;
        int rc = SSR_OK;
        if (rebuild || (refresh_now && on_device)) {
            clk.lap(11);
            rc = build_any();
        } else if (refresh_now) {
            clk.lap(11);
            if ((rc = Kh.reserve((size_t)nnzb * dim * dim))) return rc;
            SSR_HIP_CHECK(hipMemcpyAsync(Kh.data(), Kvals, sizeof(double) * (size_t)nnzb * dim * dim,
                                         hipMemcpyDeviceToHost, ssr_stream()));
            SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
            clk.lap(0);
            rc = refresh();
        }
        if (!rc && block0) {
            const int64_t nb = n_n * dim * dim;
            const int32_t *dg = ssr_pattern_d_diag(Pat);
            if (!binv64) SSR_HIP_CHECK(hipMalloc((void **)&binv64, sizeof(double) * nb));
            rc = ssr::block_jacobi_setup<double>(dim, n_n, dg, Kvals, binv64);
            if (!rc && K32) {
                if (!binv32) SSR_HIP_CHECK(hipMalloc((void **)&binv32, sizeof(float) * nb));
                rc = ssr::block_jacobi_setup<float>(dim, n_n, dg, Kvals, binv32);
            }
        }
        if (rc) {
            n_setup++;
            return rc;
        }
        if (!(rebuild || refresh_now) && need32) {
            // first float use of an existing hierarchy
            for (size_t l = 0; l < L.size() && !rc; ++l) rc = on_device ? mirror_float(l) : upload_level<float>(l);
            if (!rc) {
                std::vector<float> ci32(coarse_inv.begin(), coarse_inv.end());
                rc = dupload(&d_coarse_inv32, ci32);
            }
        }
        n_setup++;
        return rc;
    }

    template <class T> const T *fine_vals() const;
    template <class T> const T *coarse_inverse() const;

    template <class T>
    int level_spmv(size_t l, const T *x, T *y, double a, double b, const T *z)
    {
        if (l == 0)
            return ssr::spmv_bsr<T>(dim, n_n, ssr_pattern_d_row_ptr(Pat), ssr_pattern_d_col(Pat), fine_vals<T>(), x, y,
                                    a, b, z, fine_sym ? ssr_pattern_d_diag(Pat) : nullptr,
                                    fine_sym ? ssr_pattern_d_tblk(Pat) : nullptr);
        return L[l]->dev<T>().A.spmv(x, y, a, b, z);
    }

    // Chebyshev smoothing of A x = b on level l (x updated), D^-1 scaled.
    template <class T> const T *block_inv() const;

    template <class T>
    int smooth(size_t l, const T *b, T *x, bool x_zero)
    {
        if (l == 0 && block0) return smooth_block<T>(b, x, x_zero);
        Level &lv = *L[l];
        DevLevel<T> &dl = lv.dev<T>();
        const int64_t m = lv.A.nr;
        // interval [lower, upper] * rho_up with rho_up a guaranteed upper bound, so the
        // polynomial never amplifies the highest modes
        const double up = opt.cheb_upper * lv.rho_up, lo = opt.cheb_lower * lv.rho_up;
        const double theta = 0.5 * (up + lo), delta = 0.5 * (up - lo);
        const double sigma = theta / delta;
        double rho_k = 1.0 / sigma;
        int rc;
        // d = (1/theta) D^-1 r ; x += d   (r = b when x = 0)
        if (x_zero) {
            if ((rc = ssr::cheb_first<T>(m, 1.0 / theta, dl.dinv, b, dl.d, x))) return rc;
        } else {
            if ((rc = level_spmv<T>(l, x, dl.r, -1.0, 1.0, b))) return rc;
            if ((rc = ssr::cheb_step<T>(m, 0.0, 1.0 / theta, dl.dinv, dl.r, dl.d, x))) return rc;
        }
        for (int it = 1; it < opt.cheb_degree; ++it) {
            const double rho_new = 1.0 / (2.0 * sigma - rho_k);
            if ((rc = level_spmv<T>(l, x, dl.r, -1.0, 1.0, b))) return rc;  // r = b - A x
            if ((rc = ssr::cheb_step<T>(m, rho_new * rho_k, 2.0 * rho_new / delta, dl.dinv, dl.r, dl.d, x)))
                return rc;
            rho_k = rho_new;
        }
        return SSR_OK;
    }

    // Chebyshev on D_B^-1 A (finest level, node-block scaling)
    template <class T>
    int smooth_block(const T *b, T *x, bool x_zero)
    {
        DevLevel<T> &dl = L[0]->dev<T>();
        const double up = opt.cheb_upper * rho_up_block, lo = opt.cheb_lower * rho_up_block;
        const double theta = 0.5 * (up + lo), delta = 0.5 * (up - lo);
        const double sigma = theta / delta;
        double rho_k = 1.0 / sigma;
        const T *Bi = block_inv<T>();
        int rc;
        if (x_zero) {
            if ((rc = ssr::cheb_block_step<T>(dim, n_n, 0.0, 1.0 / theta, Bi, b, dl.d, x, 1))) return rc;
        } else {
            if ((rc = level_spmv<T>(0, x, dl.r, -1.0, 1.0, b))) return rc;
            if ((rc = ssr::cheb_block_step<T>(dim, n_n, 0.0, 1.0 / theta, Bi, dl.r, dl.d, x, 0))) return rc;
        }
        for (int it = 1; it < opt.cheb_degree; ++it) {
            const double rho_new = 1.0 / (2.0 * sigma - rho_k);
            if ((rc = level_spmv<T>(0, x, dl.r, -1.0, 1.0, b))) return rc;
            if ((rc = ssr::cheb_block_step<T>(dim, n_n, rho_new * rho_k, 2.0 * rho_new / delta, Bi, dl.r, dl.d, x, 0)))
                return rc;
            rho_k = rho_new;
        }
        return SSR_OK;
    }

    template <class T>
    int vcycle(size_t l, const T *b, T *x)
    {
        int rc;
        Level &lv = *L[l];
        DevLevel<T> &dl = lv.dev<T>();
        if (l + 1 == L.size()) return ssr::dense_gemv<T>((int)n_coarse, coarse_inverse<T>(), b, x);
        if ((rc = smooth<T>(l, b, x, true))) return rc;
        // r = b - A x ; bc = R r
        if ((rc = level_spmv<T>(l, x, dl.r, -1.0, 1.0, b))) return rc;
        DevLevel<T> &dc = L[l + 1]->dev<T>();
        if ((rc = dl.R.spmv(dl.r, dc.b, 1.0, 0.0))) return rc;
        if ((rc = vcycle<T>(l + 1, dc.b, dc.x))) return rc;
        // x += P xc
        if ((rc = dl.P.spmv(dc.x, x, 1.0, 1.0))) return rc;
        return smooth<T>(l, b, x, false);
    }

    // The V-cycle as one captured HIP graph over fixed input/output buffers
    // (about twenty kernels become one launch plus two copies).  The graph
    // bakes in every device pointer of the hierarchy and the finest-level
    // operator, so it is re-captured when the hierarchy is rebuilt or
    // refreshed (generation) or the operator buffer changes.  The V-cycle
    // contains kernels only (no memset or host synchronisation).
    template <class T>
    struct VGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        T *bin = nullptr, *xout = nullptr;
        const void *key = nullptr;
        int64_t gen = -1;
        void reset()
        {
            if (ge) (void)hipGraphExecDestroy(ge);
            if (g) (void)hipGraphDestroy(g);
            ge = nullptr;
            g = nullptr;
        }
        ~VGraph()
        {
            reset();
            (void)hipFree(bin);
            (void)hipFree(xout);
        }
    };
    VGraph<double> vg64;
    VGraph<float> vg32;
    hipStream_t cap_stream = nullptr;  // capture only (default priority, non-blocking)
    template <class T> VGraph<T> &vgraph();

    template <class T>
    int apply_graph(const T *x, T *y)
    {
        VGraph<T> &G = vgraph<T>();
        if (!G.ge || G.key != (const void *)fine_vals<T>() || G.gen != generation) {
            G.reset();
            if (!G.bin) {
                SSR_HIP_CHECK(hipMalloc((void **)&G.bin, sizeof(T) * n));
                SSR_HIP_CHECK(hipMalloc((void **)&G.xout, sizeof(T) * n));
            }
            if (!cap_stream) SSR_HIP_CHECK(hipStreamCreateWithFlags(&cap_stream, hipStreamNonBlocking));
            SSR_HIP_CHECK(hipStreamBeginCapture(cap_stream, hipStreamCaptureModeRelaxed));
            ssr_capture_stream_push(cap_stream);
            const int rc = vcycle<T>(0, G.bin, G.xout);
            ssr_capture_stream_pop();
            hipGraph_t g = nullptr;
            const hipError_t e = hipStreamEndCapture(cap_stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            SSR_HIP_CHECK(e);
            G.g = g;
            SSR_HIP_CHECK(hipGraphInstantiate(&G.ge, G.g, nullptr, nullptr, 0));
            G.key = (const void *)fine_vals<T>();
            G.gen = generation;
        }
        SSR_HIP_CHECK(hipMemcpyAsync(G.bin, x, sizeof(T) * n, hipMemcpyDeviceToDevice, ssr_stream()));
        SSR_HIP_CHECK(hipGraphLaunch(G.ge, ssr_stream()));
        SSR_HIP_CHECK(hipMemcpyAsync(y, G.xout, sizeof(T) * n, hipMemcpyDeviceToDevice, ssr_stream()));
        return SSR_OK;
    }

    int apply(const double *x, double *y) override
    {
        if (!built) {
            ssr_set_error("AMG: apply before setup");
            return SSR_ERR_STATE;
        }
        if (ssr_tuning(SSR_TUNE_AMG_GRAPH)) return apply_graph<double>(x, y);
        return vcycle<double>(0, x, y);
    }

    int apply(const float *x, float *y) override
    {
        if (!built || !want32 || !Kdev32) {
            ssr_set_error("AMG: float apply before a mixed-precision setup");
            return SSR_ERR_STATE;
        }
        if (ssr_tuning(SSR_TUNE_AMG_GRAPH)) return apply_graph<float>(x, y);
        return vcycle<float>(0, x, y);
    }
};

template <> Amg::VGraph<double> &Amg::vgraph<double>() { return vg64; }
template <> Amg::VGraph<float> &Amg::vgraph<float>() { return vg32; }

template <> const double *Amg::block_inv<double>() const { return binv64; }
template <> const float *Amg::block_inv<float>() const { return binv32; }
template <> const double *Amg::fine_vals<double>() const { return Kdev; }
template <> const float *Amg::fine_vals<float>() const { return Kdev32; }
template <> const double *Amg::coarse_inverse<double>() const { return d_coarse_inv; }
template <> const float *Amg::coarse_inverse<float>() const { return d_coarse_inv32; }

extern "C" {

// Replace the solver's preconditioner by SA-AMG with the given near-null
// space (n_null columns over all dim*n_n DOFs, host, column-major).
// opts (may be NULL): double[8] = {theta, max_levels, max_coarse, cheb_degree,
// cheb_lower, cheb_upper, refresh_every, rebuild_every}.
int ssr_solver_set_amg(ssr_solver *s, const double *near_null, int n_null, const double *opts)
{
    SSR_REQUIRE(s && near_null && n_null > 0 && n_null <= 16, "set_amg: bad arguments");
    Options o;
    if (opts) {
        o.theta = opts[0];
        o.max_levels = (int)opts[1];
        o.max_coarse = (int64_t)opts[2];
        o.cheb_degree = (int)opts[3];
        o.cheb_lower = opts[4];
        o.cheb_upper = opts[5];
        o.refresh_every = (int)opts[6];
        o.rebuild_every = (int)opts[7];
    }
    Amg *M = new Amg(ssr_solver_pattern(s), near_null, n_null, o);
    return ssr_solver_set_precond(s, M);
}

}  // extern "C"
