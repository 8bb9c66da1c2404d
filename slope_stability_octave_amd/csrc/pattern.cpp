// Fixed sparsity pattern of the tangent operator (host-side setup).
//
// Reference behaviour: CONSTITUTIVE.init_K_r_pattern (CONSTITUTIVE.m:393-464)
// builds the symbolic pattern of B_Q' * D * B_Q for any block-diagonal D --
// free DOFs coupled whenever their nodes share an element -- and
// build_scatter_map (CONSTITUTIVE.m:582-635) maps each element-matrix entry
// to its position in that pattern.
//
// This build stores the operator as node-block CSR (BSR) over ALL nodes with
// dim x dim blocks: the block pattern is the node adjacency through elements,
// so every element contributes whole blocks and one block index per node
// pair replaces the reference's n_ld^2 int64 map (100 int32 per P2 tetrahedron
// instead of 900 int64).  Dirichlet DOFs are masked (zero rows/columns, unit
// diagonal), which leaves the restriction to the free DOFs equal to K(Q,Q).
// `ref_pos` lists, in the reference's order (column-major `find` of the
// free-DOF pattern, CONSTITUTIVE.m:445), where each K(Q,Q) entry lives in the
// block value array, so results can be handed back in the reference layout.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "pattern.h"
#include "ssr_common.h"

namespace {
template <class T>
int upload(T **dst, const std::vector<T> &src)
{
    size_t bytes = sizeof(T) * (src.empty() ? 1 : src.size());
    hipError_t e = hipMalloc((void **)dst, bytes);
    if (e != hipSuccess) {
        ssr_set_error("hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return SSR_ERR_ALLOC;
    }
    if (!src.empty()) {
        SSR_HIP_CHECK(hipMemcpyAsync(*dst, src.data(), sizeof(T) * src.size(), hipMemcpyHostToDevice, ssr_stream()));
        SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
    }
    return SSR_OK;
}
}  // namespace

extern "C" {

int ssr_pattern_destroy(ssr_pattern *p)
{
    if (!p) return SSR_OK;
    (void)hipFree(p->d_row_ptr);
    (void)hipFree(p->d_col);
    (void)hipFree(p->d_emap);
    (void)hipFree(p->d_cptr);
    (void)hipFree(p->d_ccode);
    (void)hipFree(p->d_uptr);
    (void)hipFree(p->d_pslot);
    (void)hipFree(p->d_keblk);
    (void)hipFree(p->d_nptr);
    (void)hipFree(p->d_ncode);
    (void)hipFree(p->d_fe);
    (void)hipFree(p->d_tblk);
    (void)hipFree(p->d_ublk);
    (void)hipFree(p->d_diag);
    (void)hipFree(p->d_elem);
    (void)hipFree(p->d_brow);
    (void)hipFree(p->d_ref_pos);
    (void)hipFree(p->d_mask);
    delete p;
    return SSR_OK;
}

// elem: n_p x n_e (element-major, 0-based node ids); Q: dim x n_n (node-major,
// nonzero = free).  upload_device = 0 builds the host tables only (usable
// without a GPU).
int ssr_pattern_create(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const int64_t *elem,
                       const uint8_t *Q, int upload_device, ssr_pattern **out)
{
    SSR_REQUIRE(out, "pattern_create: null out");
    *out = nullptr;
    SSR_REQUIRE(dim == 2 || dim == 3, "pattern_create: dim must be 2 or 3");
    SSR_REQUIRE(n_n > 0 && n_p > 0 && n_e > 0 && elem && Q, "pattern_create: bad sizes");
    SSR_REQUIRE(n_n < (int64_t)1 << 31, "pattern_create: too many nodes for int32 indices");
    for (int64_t k = 0; k < n_p * n_e; ++k)
        SSR_REQUIRE(elem[k] >= 0 && elem[k] < n_n, "pattern_create: node id %lld out of range",
                    (long long)elem[k]);

    ssr_pattern *p = new ssr_pattern();
    p->dim = dim;
    p->n_n = n_n;
    p->n_p = n_p;
    p->n_e = n_e;
    // node -> elements
    std::vector<int64_t> ne_ptr(n_n + 1, 0), ne_idx(n_p * n_e);
    for (int64_t k = 0; k < n_p * n_e; ++k) ne_ptr[elem[k] + 1]++;
    for (int64_t i = 0; i < n_n; ++i) ne_ptr[i + 1] += ne_ptr[i];
    {
        std::vector<int64_t> fill(ne_ptr.begin(), ne_ptr.end() - 1);
        for (int64_t e = 0; e < n_e; ++e)
            for (int64_t a = 0; a < n_p; ++a) ne_idx[fill[elem[e * n_p + a]]++] = e;
    }
    // node adjacency (sorted, unique) with a marker array
    p->row_ptr.assign(n_n + 1, 0);
    std::vector<int32_t> mark(n_n, -1), nb;
    std::vector<int32_t> cols;
    cols.reserve((size_t)n_n * 30);
    for (int64_t r = 0; r < n_n; ++r) {
        nb.clear();
        for (int64_t q = ne_ptr[r]; q < ne_ptr[r + 1]; ++q) {
            const int64_t e = ne_idx[q];
            for (int64_t a = 0; a < n_p; ++a) {
                const int32_t c = (int32_t)elem[e * n_p + a];
                if (mark[c] != (int32_t)r) {
                    mark[c] = (int32_t)r;
                    nb.push_back(c);
                }
            }
        }
        std::sort(nb.begin(), nb.end());
        cols.insert(cols.end(), nb.begin(), nb.end());
        p->row_ptr[r + 1] = (int32_t)cols.size();
        SSR_REQUIRE(cols.size() < ((size_t)1 << 31), "pattern_create: too many blocks");
    }
    p->col.swap(cols);
    p->nnzb = (int64_t)p->col.size();
    auto find_blk = [&](int64_t r, int32_t c) -> int32_t {
        const int32_t *b = p->col.data() + p->row_ptr[r], *e = p->col.data() + p->row_ptr[r + 1];
        const int32_t *it = std::lower_bound(b, e, c);
        return (int32_t)(it - p->col.data());
    };
    p->diag.resize(n_n);
    for (int64_t r = 0; r < n_n; ++r) p->diag[r] = find_blk(r, (int32_t)r);
    p->brow.resize(p->nnzb);
    for (int64_t r = 0; r < n_n; ++r)
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k) p->brow[k] = (int32_t)r;
    p->tblk.resize(p->nnzb);
    for (int64_t r = 0; r < n_n; ++r)
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k) p->tblk[k] = find_blk(p->col[k], (int32_t)r);
    p->ublk.clear();
    for (int64_t r = 0; r < n_n; ++r)
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k)
            if (p->col[k] >= r) p->ublk.push_back(k);
    p->n_ublk = (int64_t)p->ublk.size();
    p->emap.resize(n_e * n_p * n_p);
    for (int64_t e = 0; e < n_e; ++e)
        for (int64_t a = 0; a < n_p; ++a)
            for (int64_t b = 0; b < n_p; ++b)
                p->emap[(e * n_p + a) * n_p + b] = find_blk(elem[e * n_p + a], (int32_t)elem[e * n_p + b]);

    // contributions per block for the gather-form assembly: each element
    // block pair (a <= b) is computed once; block (a,b) takes it as is and
    // block (b,a) transposed.  Pair numbering: row-major upper triangle.
    {
        const int64_t npair = n_p * (n_p + 1) / 2;
        if (n_e * npair * 2 < ((int64_t)1 << 31) && n_e * n_p * n_p < ((int64_t)1 << 31)) {
            p->cptr.assign(p->nnzb + 1, 0);
            for (int64_t k = 0; k < n_e * n_p * n_p; ++k) p->cptr[p->emap[k] + 1]++;
            for (int64_t b = 0; b < p->nnzb; ++b) p->cptr[b + 1] += p->cptr[b];
            p->ccode.resize(n_e * n_p * n_p);
            std::vector<int32_t> fill(p->cptr.begin(), p->cptr.end() - 1);
            for (int64_t e = 0; e < n_e; ++e)
                for (int64_t a = 0; a < n_p; ++a)
                    for (int64_t b = 0; b < n_p; ++b) {
                        const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
                        const int64_t pr = lo * n_p - lo * (lo - 1) / 2 + (hi - lo);
                        const int32_t blk = p->emap[(e * n_p + a) * n_p + b];
                        p->ccode[fill[blk]++] = (int32_t)((e * npair + pr) * 2 + (a > b ? 1 : 0));
                    }
        }
    }

    // slot lists of the upper blocks
    {
        const int64_t npair = n_p * (n_p + 1) / 2;
        if (n_e * npair < ((int64_t)1 << 31)) {
            std::vector<int32_t> uidx(p->nnzb, -1);
            for (size_t u = 0; u < p->ublk.size(); ++u) uidx[p->ublk[u]] = (int32_t)u;
            const int64_t nu = (int64_t)p->ublk.size();
            p->uptr.assign(nu + 1, 0);
            p->pslot.resize(n_e * npair);
            auto upper_of = [&](int64_t e, int64_t a, int64_t b) -> int32_t {
                const int64_t v = elem[e * n_p + a], w = elem[e * n_p + b];
                const int64_t lo = v <= w ? a : b, hi = v <= w ? b : a;
                return uidx[p->emap[(e * n_p + lo) * n_p + hi]];
            };
            for (int64_t e = 0; e < n_e; ++e)
                for (int64_t a = 0; a < n_p; ++a)
                    for (int64_t b = a; b < n_p; ++b) p->uptr[upper_of(e, a, b) + 1]++;
            for (int64_t u = 0; u < nu; ++u) p->uptr[u + 1] += p->uptr[u];
            std::vector<int32_t> fill(p->uptr.begin(), p->uptr.end() - 1);
            int64_t k = 0;
            for (int64_t e = 0; e < n_e; ++e)
                for (int64_t a = 0; a < n_p; ++a)
                    for (int64_t b = a; b < n_p; ++b) p->pslot[k++] = fill[upper_of(e, a, b)]++;
        }
    }

    // node incidences for the gather-form internal force
    if (n_e * n_p < ((int64_t)1 << 31)) {
        p->nptr.assign(n_n + 1, 0);
        for (int64_t k = 0; k < n_e * n_p; ++k) p->nptr[elem[k] + 1]++;
        for (int64_t v = 0; v < n_n; ++v) p->nptr[v + 1] += p->nptr[v];
        p->ncode.resize(n_e * n_p);
        std::vector<int32_t> fill(p->nptr.begin(), p->nptr.end() - 1);
        for (int64_t k = 0; k < n_e * n_p; ++k) p->ncode[fill[elem[k]]++] = (int32_t)k;
    }

    // free DOFs and the reference-order extraction map
    p->mask.assign(Q, Q + dim * n_n);
    for (auto &m : p->mask) m = m ? 1 : 0;
    p->n_free = 0;
    for (auto m : p->mask) p->n_free += m;
    const int bs = dim * dim;
    p->ref_pos.clear();
    p->ref_pos.reserve((size_t)p->nnzb * bs);
    // column-major find over the free pattern: outer loop = column (free dof
    // in increasing global order), inner = row ascending.  Entry (row I, col J)
    // sits in block (node(I), node(J)) at (comp(I), comp(J)).
    for (int64_t cn = 0; cn < n_n; ++cn)
        for (int cj = 0; cj < dim; ++cj) {
            if (!p->mask[dim * cn + cj]) continue;
            for (int32_t k = p->row_ptr[cn]; k < p->row_ptr[cn + 1]; ++k) {
                const int32_t rn = p->col[k];  // symmetric pattern: rows of column cn = neighbours
                const int32_t blk = p->tblk[k];  // block (rn, cn)
                for (int ri = 0; ri < dim; ++ri)
                    if (p->mask[dim * rn + ri]) p->ref_pos.push_back((int64_t)blk * bs + ri * dim + cj);
            }
        }
    p->nnz_free = (int64_t)p->ref_pos.size();

    if (upload_device) {
        int rc;
        std::vector<int32_t> el32(elem, elem + n_p * n_e);
        std::vector<double> md(p->mask.begin(), p->mask.end());
        if ((rc = upload(&p->d_row_ptr, p->row_ptr)) || (rc = upload(&p->d_col, p->col)) ||
            (rc = upload(&p->d_emap, p->emap)) || (rc = upload(&p->d_tblk, p->tblk)) ||
            (rc = upload(&p->d_ublk, p->ublk)) ||
            (rc = upload(&p->d_diag, p->diag)) || (rc = upload(&p->d_elem, el32)) ||
            (rc = upload(&p->d_ref_pos, p->ref_pos)) || (rc = upload(&p->d_mask, md)) ||
            (rc = upload(&p->d_brow, p->brow)) ||
            (!p->cptr.empty() && ((rc = upload(&p->d_cptr, p->cptr)) || (rc = upload(&p->d_ccode, p->ccode)))) ||
            (!p->pslot.empty() && ((rc = upload(&p->d_uptr, p->uptr)) || (rc = upload(&p->d_pslot, p->pslot)))) ||
            (!p->nptr.empty() && ((rc = upload(&p->d_nptr, p->nptr)) || (rc = upload(&p->d_ncode, p->ncode))))) {
            ssr_pattern_destroy(p);
            return rc;
        }
    }
    *out = p;
    return SSR_OK;
}

int ssr_pattern_info(const ssr_pattern *p, int64_t *info /* 8 */)
{
    SSR_REQUIRE(p && info, "pattern_info: null");
    info[0] = p->dim;
    info[1] = p->n_n;
    info[2] = p->n_p;
    info[3] = p->n_e;
    info[4] = p->nnzb;
    info[5] = p->n_free;
    info[6] = p->nnz_free;
    info[7] = p->d_row_ptr != nullptr;
    return SSR_OK;
}

// Copy host tables out: which = 0 row_ptr (n_n+1 int32), 1 col (nnzb int32),
// 2 emap (n_e*n_p*n_p int32), 3 tblk (nnzb int32), 4 diag (n_n int32),
// 5 ref_pos (nnz_free int64), 6 mask (dim*n_n uint8).
int ssr_pattern_get(const ssr_pattern *p, int which, void *dst)
{
    SSR_REQUIRE(p && dst, "pattern_get: null");
    if (p->device_built && which != 6) {
        // tables built on the device: copy the device table out
        const void *src = nullptr;
        size_t bytes = 0;
        switch (which) {
        case 0: src = p->d_row_ptr; bytes = sizeof(int32_t) * (p->n_n + 1); break;
        case 1: src = p->d_col; bytes = sizeof(int32_t) * p->nnzb; break;
        case 2: src = p->d_emap; bytes = sizeof(int32_t) * p->n_e * p->n_p * p->n_p; break;
        case 3: src = p->d_tblk; bytes = sizeof(int32_t) * p->nnzb; break;
        case 4: src = p->d_diag; bytes = sizeof(int32_t) * p->n_n; break;
        case 5: src = p->d_ref_pos; bytes = sizeof(int64_t) * p->nnz_free; break;
        default: SSR_REQUIRE(false, "pattern_get: bad table %d", which);
        }
        if (bytes) {
            SSR_HIP_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ssr_stream()));
            SSR_HIP_CHECK(hipStreamSynchronize(ssr_stream()));
        }
        return SSR_OK;
    }
    switch (which) {
    case 0: memcpy(dst, p->row_ptr.data(), sizeof(int32_t) * p->row_ptr.size()); break;
    case 1: memcpy(dst, p->col.data(), sizeof(int32_t) * p->col.size()); break;
    case 2: memcpy(dst, p->emap.data(), sizeof(int32_t) * p->emap.size()); break;
    case 3: memcpy(dst, p->tblk.data(), sizeof(int32_t) * p->tblk.size()); break;
    case 4: memcpy(dst, p->diag.data(), sizeof(int32_t) * p->diag.size()); break;
    case 5: memcpy(dst, p->ref_pos.data(), sizeof(int64_t) * p->ref_pos.size()); break;
    case 6: memcpy(dst, p->mask.data(), p->mask.size()); break;
    default: SSR_REQUIRE(false, "pattern_get: bad table %d", which);
    }
    return SSR_OK;
}

}  // extern "C"

// accessors for the kernels in other translation units
const int32_t *ssr_pattern_d_row_ptr(const ssr_pattern *p) { return p->d_row_ptr; }
const int32_t *ssr_pattern_d_col(const ssr_pattern *p) { return p->d_col; }
const int32_t *ssr_pattern_d_emap(const ssr_pattern *p) { return p->d_emap; }
// Pattern of a given node-block CSR (no elements): rows sorted, every
// diagonal block present, all DOFs free.  Used by the AMG instances built
// from COO triplets (amg_ijv.cpp); the assembly tables stay empty.
int ssr_pattern_create_blocks(int dim, int64_t n_n, std::vector<int32_t> row_ptr, std::vector<int32_t> col,
                              ssr_pattern **out)
{
    SSR_REQUIRE(out && (dim == 2 || dim == 3) && n_n > 0 && (int64_t)row_ptr.size() == n_n + 1,
                "pattern_create_blocks: bad arguments");
    ssr_pattern *p = new ssr_pattern();
    p->dim = dim;
    p->n_n = n_n;
    p->row_ptr.swap(row_ptr);
    p->col.swap(col);
    p->nnzb = (int64_t)p->col.size();
    auto find_blk = [&](int64_t r, int32_t c) -> int32_t {
        const int32_t *b = p->col.data() + p->row_ptr[r], *e = p->col.data() + p->row_ptr[r + 1];
        const int32_t *it = std::lower_bound(b, e, c);
        return (it != e && *it == c) ? (int32_t)(it - p->col.data()) : -1;
    };
    p->diag.resize(n_n);
    p->brow.resize(p->nnzb);
    p->tblk.resize(p->nnzb);
    for (int64_t r = 0; r < n_n; ++r) {
        p->diag[r] = find_blk(r, (int32_t)r);
        if (p->diag[r] < 0) {
            delete p;
            ssr_set_error("pattern_create_blocks: block row %lld has no diagonal block", (long long)r);
            return SSR_ERR_ARG;
        }
        for (int32_t k = p->row_ptr[r]; k < p->row_ptr[r + 1]; ++k) {
            p->brow[k] = (int32_t)r;
            p->tblk[k] = find_blk(p->col[k], (int32_t)r);
        }
    }
    p->mask.assign((size_t)dim * n_n, 1);
    p->n_free = dim * n_n;
    std::vector<double> md(p->mask.begin(), p->mask.end());
    int rc;
    if ((rc = upload(&p->d_row_ptr, p->row_ptr)) || (rc = upload(&p->d_col, p->col)) ||
        (rc = upload(&p->d_diag, p->diag)) || (rc = upload(&p->d_brow, p->brow)) ||
        (rc = upload(&p->d_tblk, p->tblk)) || (rc = upload(&p->d_mask, md))) {
        ssr_pattern_destroy(p);
        return rc;
    }
    *out = p;
    return SSR_OK;
}

const int32_t *ssr_pattern_d_tblk(const ssr_pattern *p) { return p->d_tblk; }
const int32_t *ssr_pattern_d_ublk(const ssr_pattern *p) { return p->d_ublk; }
const int32_t *ssr_pattern_d_uptr(const ssr_pattern *p) { return p->d_uptr; }
const int32_t *ssr_pattern_d_pslot(const ssr_pattern *p) { return p->d_pslot; }
int64_t ssr_pattern_n_ublk(const ssr_pattern *p) { return p->n_ublk; }
const int32_t *ssr_pattern_d_diag(const ssr_pattern *p) { return p->d_diag; }
const int32_t *ssr_pattern_d_elem(const ssr_pattern *p) { return p->d_elem; }
const int32_t *ssr_pattern_d_brow(const ssr_pattern *p) { return p->d_brow; }
const int64_t *ssr_pattern_d_ref_pos(const ssr_pattern *p) { return p->d_ref_pos; }
const double *ssr_pattern_d_mask(const ssr_pattern *p) { return p->d_mask; }
const int32_t *ssr_pattern_d_cptr(const ssr_pattern *p) { return p->d_cptr; }
const int32_t *ssr_pattern_d_ccode(const ssr_pattern *p) { return p->d_ccode; }
// element-block scratch of the gather-form assembly (allocated once)
int ssr_pattern_keblk(const ssr_pattern *p, double **out)
{
    if (!p->d_keblk) {
        const int64_t len = p->n_e * (p->n_p * (p->n_p + 1) / 2) * p->dim * p->dim;
        SSR_HIP_CHECK(hipMalloc((void **)&p->d_keblk, sizeof(double) * (len > 0 ? len : 1)));
    }
    *out = p->d_keblk;
    return SSR_OK;
}
const int32_t *ssr_pattern_d_nptr(const ssr_pattern *p) { return p->d_nptr; }
const int32_t *ssr_pattern_d_ncode(const ssr_pattern *p) { return p->d_ncode; }
// element-node force scratch of the gather-form internal force (allocated once)
int ssr_pattern_fe(const ssr_pattern *p, double **out)
{
    if (!p->d_fe) {
        const int64_t len = p->n_e * p->n_p * p->dim;
        SSR_HIP_CHECK(hipMalloc((void **)&p->d_fe, sizeof(double) * (len > 0 ? len : 1)));
    }
    *out = p->d_fe;
    return SSR_OK;
}
void ssr_pattern_sizes(const ssr_pattern *p, int *dim, int64_t *n_n, int64_t *n_p, int64_t *n_e,
                       int64_t *nnzb, int64_t *nnz_free)
{
    *dim = p->dim;
    *n_n = p->n_n;
    *n_p = p->n_p;
    *n_e = p->n_e;
    *nnzb = p->nnzb;
    *nnz_free = p->nnz_free;
}

// Device table addresses: which = 0 row_ptr, 1 col, 2 emap, 3 tblk, 4 diag,
// 5 elem (int32 n_p x n_e), 6 ref_pos, 7 mask (double), 8 brow, 9 ublk,
// 10 uptr, 11 pslot, 12 cptr, 13 ccode, 14 nptr, 15 ncode (NULL when absent).
extern "C" int ssr_pattern_dev_ptr(const ssr_pattern *p, int which, void **out)
{
    SSR_REQUIRE(p && out, "pattern_dev_ptr: null");
    switch (which) {
    case 0: *out = p->d_row_ptr; break;
    case 1: *out = p->d_col; break;
    case 2: *out = p->d_emap; break;
    case 3: *out = p->d_tblk; break;
    case 4: *out = p->d_diag; break;
    case 5: *out = p->d_elem; break;
    case 6: *out = p->d_ref_pos; break;
    case 7: *out = p->d_mask; break;
    case 8: *out = p->d_brow; break;
    case 9: *out = p->d_ublk; break;
    case 10: *out = p->d_uptr; break;
    case 11: *out = p->d_pslot; break;
    case 12: *out = p->d_cptr; break;
    case 13: *out = p->d_ccode; break;
    case 14: *out = p->d_nptr; break;
    case 15: *out = p->d_ncode; break;
    default: SSR_REQUIRE(false, "pattern_dev_ptr: bad table %d", which);
    }
    return SSR_OK;
}
