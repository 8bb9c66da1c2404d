// AMG preconditioner instances keyed by an id string, built from COO
// triplets: the command surface of the reference's BoomerAMG binding
// (+LINEAR_SOLVERS/mex/hypre_boomeramg_mex.cpp) over this library's
// smoothed-aggregation AMG.
//
//   setup_ijv(I, J, V, n, Z, opts, id)  hypre_boomeramg_mex.cpp:1013-1149
//   update_values(V, id)                :1151-1211
//   apply(r, id)                        :751-796 (one V-cycle from x = 0)
//   clear([id]), info                   :797-830, :1215-1270
//
// I, J are 1-based, double-typed as Octave passes them.  The matrix is taken
// in node blocks of `dofs_per_node` consecutive unknowns (the reference's
// num_functions with dof_func = mod(q-1, num_functions),
// set_linear_solver.m:186-214); each triplet's place among the block values
// is kept (the reference's coo_to_rowmajor, :471-518), so update_values only
// scatters the new values.  A setup_ijv on an existing id with the same
// number of triplets takes that same values-only path (the reference's fast
// path); force_rebuild = 1 rebuilds from the new triplets.  Duplicate (I, J)
// entries are summed.  Every values update rebuilds the hierarchy in full,
// as the reference re-runs HYPRE_BoomerAMGSetup (:1200): the result is a
// preconditioner of the new matrix.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include <algorithm>
#include <chrono>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "ssr_common.h"
#include "../../include/ssr_hip.h"

int ssr_pattern_create_blocks(int dim, int64_t n_n, std::vector<int32_t> row_ptr, std::vector<int32_t> col,
                              ssr_pattern **out);

namespace {

struct Instance {
    ssr_pattern *P = nullptr;
    ssr_solver *S = nullptr;
    int64_t n = 0, n_pad = 0, nnz = 0, nnzb = 0;
    int dpn = 0, n_null = 0;
    std::vector<int64_t> pos;  // triplet k -> index into the block values
    int64_t diag_last = 0;     // block index of the last block row's diagonal
    std::vector<double> hK;
    double *dK = nullptr, *dx = nullptr, *dy = nullptr;
    ~Instance()
    {
        (void)ssr_solver_destroy(S);
        (void)ssr_pattern_destroy(P);
        (void)hipFree(dK);
        (void)hipFree(dx);
        (void)hipFree(dy);
    }
};

std::mutex g_mu;
std::unordered_map<std::string, std::unique_ptr<Instance>> g_inst;

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// block values from V; the unknowns that pad n to whole node blocks are
// decoupled with a unit diagonal
int scatter_values(Instance &I, const double *V)
{
    std::fill(I.hK.begin(), I.hK.end(), 0.0);
    for (int64_t k = 0; k < I.nnz; ++k) I.hK[I.pos[k]] += V[k];
    const int d = I.dpn;
    for (int64_t q = I.n; q < I.n_pad; ++q) {
        const int64_t blk = I.diag_last;  // the last block row's diagonal block
        I.hK[blk * d * d + (q % d) * d + (q % d)] = 1.0;
    }
    return ssr_htod(I.dK, I.hK.data(), sizeof(double) * I.hK.size());
}

int build(Instance &I, const double *Iv, const double *Jv, const double *V, int64_t nnz, int64_t n, const double *Z,
          int n_null, int dpn, const double *opts)
{
    SSR_REQUIRE(dpn == 2 || dpn == 3, "amg setup_ijv: dofs_per_node must be 2 or 3 (got %d)", dpn);
    SSR_REQUIRE(n > 0, "amg setup_ijv: n must be positive");
    SSR_REQUIRE(nnz > 0 && Iv && Jv && V, "amg setup_ijv: I, J, V must be non-empty");
    SSR_REQUIRE(n_null >= 0 && n_null <= 16 && (n_null == 0 || Z), "amg setup_ijv: bad null space");
    const int64_t nb = (n + dpn - 1) / dpn;  // node blocks; the last may be padded
    SSR_REQUIRE(nb < ((int64_t)1 << 31), "amg setup_ijv: n too large");
    std::vector<int64_t> bi(nnz), bj(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
        const double a = Iv[k], b = Jv[k];
        SSR_REQUIRE(a >= 1 && a <= (double)n && b >= 1 && b <= (double)n && a == floor(a) && b == floor(b),
                    "amg setup_ijv: triplet %lld has index (%g, %g) outside 1..%lld", (long long)k, a, b,
                    (long long)n);
        bi[k] = (int64_t)a - 1;
        bj[k] = (int64_t)b - 1;
    }
    // block pattern: the blocks the triplets touch plus every diagonal block
    std::vector<std::vector<int32_t>> rows(nb);
    for (int64_t r = 0; r < nb; ++r) rows[r].push_back((int32_t)r);
    for (int64_t k = 0; k < nnz; ++k) rows[bi[k] / dpn].push_back((int32_t)(bj[k] / dpn));
    std::vector<int32_t> rp(nb + 1, 0), col;
    for (int64_t r = 0; r < nb; ++r) {
        auto &v = rows[r];
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        rp[r + 1] = rp[r] + (int32_t)v.size();
    }
    col.reserve(rp[nb]);
    for (auto &v : rows) col.insert(col.end(), v.begin(), v.end());
    const int bs = dpn * dpn;
    I.pos.resize(nnz);
    for (int64_t k = 0; k < nnz; ++k) {
        const int64_t r = bi[k] / dpn;
        const int32_t *b0 = col.data() + rp[r], *b1 = col.data() + rp[r + 1];
        const int64_t blk = std::lower_bound(b0, b1, (int32_t)(bj[k] / dpn)) - col.data();
        I.pos[k] = blk * bs + (bi[k] % dpn) * dpn + (bj[k] % dpn);
    }
    I.n = n;
    I.n_pad = nb * dpn;
    I.nnz = nnz;
    I.nnzb = rp[nb];
    I.dpn = dpn;
    I.diag_last = std::lower_bound(col.data() + rp[nb - 1], col.data() + rp[nb], (int32_t)(nb - 1)) - col.data();
    int rc;
    if ((rc = ssr_pattern_create_blocks(dpn, nb, std::move(rp), std::move(col), &I.P))) return rc;
    I.hK.assign((size_t)I.nnzb * bs, 0.0);
    if ((rc = ssr_dmalloc((void **)&I.dK, sizeof(double) * I.hK.size())) ||
        (rc = ssr_dmalloc((void **)&I.dx, sizeof(double) * I.n_pad)) ||
        (rc = ssr_dmalloc((void **)&I.dy, sizeof(double) * I.n_pad)) || (rc = ssr_dzero(I.dx, sizeof(double) * I.n_pad)))
        return rc;
    if ((rc = scatter_values(I, V))) return rc;
    if ((rc = ssr_solver_create(I.P, 0, 0.0, 0.0, 0, &I.S))) return rc;
    // near-null space: the given columns, else one constant per function
    const int64_t np_ = I.n_pad;
    std::vector<double> B;
    int k = n_null;
    if (k > 0) {
        B.assign((size_t)np_ * k, 0.0);
        for (int c = 0; c < k; ++c) std::copy(Z + (size_t)c * n, Z + (size_t)(c + 1) * n, B.begin() + (size_t)c * np_);
    } else {
        k = dpn;
        B.assign((size_t)np_ * k, 0.0);
        for (int64_t q = 0; q < n; ++q) B[(size_t)(q % dpn) * np_ + q] = 1.0;
    }
    I.n_null = k;
    if ((rc = ssr_solver_set_amg(I.S, B.data(), k, opts))) return rc;
    return ssr_solver_set_matrix(I.S, I.dK, 1);
}

Instance *find(const char *id)
{
    auto it = g_inst.find(id);
    return it == g_inst.end() ? nullptr : it->second.get();
}

}  // namespace

extern "C" {

int ssr_amg_setup_ijv(const char *id, const double *I, const double *J, const double *V, int64_t nnz, int64_t n,
                      const double *Z, int n_null, int dofs_per_node, const double *opts, int force_rebuild,
                      double *info)
{
    SSR_REQUIRE(id, "amg setup_ijv: instance id missing");
    std::lock_guard<std::mutex> lk(g_mu);
    const auto t0 = std::chrono::steady_clock::now();
    Instance *old = find(id);
    int rc;
    if (old && !force_rebuild && old->nnz == nnz && old->n == n) {
        // values-only update of the stored pattern (hypre_boomeramg_mex.cpp:1054-1104)
        SSR_REQUIRE(V, "amg setup_ijv: V missing");
        if ((rc = scatter_values(*old, V)) || (rc = ssr_solver_invalidate_preconditioner(old->S)) ||
            (rc = ssr_solver_set_matrix(old->S, old->dK, 1)) || (rc = ssr_synchronize()))
            return rc;
    } else {
        if (old) g_inst.erase(id);
        std::unique_ptr<Instance> inst(new Instance());
        if ((rc = build(*inst, I, J, V, nnz, n, Z, n_null, dofs_per_node, opts)) || (rc = ssr_synchronize()))
            return rc;
        g_inst[id] = std::move(inst);
    }
    if (info) {
        const Instance &x = *g_inst[id];
        info[0] = (double)x.n;
        info[1] = (double)x.nnz;
        info[2] = (double)x.dpn;
        info[3] = (double)x.n_null;
        info[4] = seconds_since(t0);
    }
    return SSR_OK;
}

int ssr_amg_update_values(const char *id, const double *V, int64_t nnz, double *setup_seconds)
{
    SSR_REQUIRE(id && V, "amg update_values: bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    Instance *x = find(id);
    SSR_REQUIRE(x, "amg update_values: unknown instance id '%s' (call setup_ijv first)", id);
    SSR_REQUIRE(nnz == x->nnz, "amg update_values: V has %lld values, the instance's pattern %lld", (long long)nnz,
                (long long)x->nnz);
    const auto t0 = std::chrono::steady_clock::now();
    int rc;
    if ((rc = scatter_values(*x, V)) || (rc = ssr_solver_invalidate_preconditioner(x->S)) ||
        (rc = ssr_solver_set_matrix(x->S, x->dK, 1)) || (rc = ssr_synchronize()))
        return rc;
    if (setup_seconds) *setup_seconds = seconds_since(t0);
    return SSR_OK;
}

int ssr_amg_apply(const char *id, const double *r, int64_t n, double *y)
{
    SSR_REQUIRE(id && r && y, "amg apply: bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    Instance *x = find(id);
    SSR_REQUIRE(x, "amg apply: unknown instance id '%s' (call setup_ijv first)", id);
    SSR_REQUIRE(n == x->n, "amg apply: rhs length %lld does not match the preconditioner size %lld", (long long)n,
                (long long)x->n);
    int rc;
    if ((rc = ssr_htod(x->dx, r, sizeof(double) * n)) || (rc = ssr_solver_precond_apply(x->S, x->dx, x->dy)))
        return rc;
    return ssr_dtoh(y, x->dy, sizeof(double) * n);
}

int ssr_amg_apply_dev(const char *id, const double *r, double *y)
{
    SSR_REQUIRE(id && r && y, "amg apply_dev: bad arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    Instance *x = find(id);
    SSR_REQUIRE(x, "amg apply_dev: unknown instance id '%s' (call setup_ijv first)", id);
    // through the padded work vectors: the caller's vectors hold n entries
    int rc;
    if ((rc = ssr_dtod(x->dx, r, sizeof(double) * x->n)) || (rc = ssr_solver_precond_apply(x->S, x->dx, x->dy)))
        return rc;
    return ssr_dtod(y, x->dy, sizeof(double) * x->n);
}

int ssr_amg_clear(const char *id)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int rc = ssr_synchronize();
    if (rc) return rc;
    if (!id) {
        g_inst.clear();
        return SSR_OK;
    }
    g_inst.erase(id);
    return SSR_OK;
}

int ssr_amg_info(int64_t *count, int64_t *n_total)
{
    SSR_REQUIRE(count, "amg info: null");
    std::lock_guard<std::mutex> lk(g_mu);
    *count = (int64_t)g_inst.size();
    if (n_total) {
        int64_t s = 0;
        for (auto &kv : g_inst) s += kv.second->n;
        *n_total = s;
    }
    return SSR_OK;
}

}  // extern "C"
