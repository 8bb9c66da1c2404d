// Setup path on the device: P2 gradients and integration weights, the volume
// force, per-point material fields, and every table of the fixed pattern.
//
// Reference behaviour followed here:
//  * DPhi and WEIGHT: elastic_stiffness_matrix_3D.m:60-96,148 and
//    elastic_stiffness_matrix_2D.m:50-80,118 -- the Jacobian at each
//    integration point as the sum over the element's nodes of coordinate times
//    reference derivative, its determinant and inverse by cofactors, the
//    physical gradients as Jinv * reference gradients, |det J| * w_q.  The
//    arithmetic is evaluated in the reference's order with FP contraction off.
//  * P2 basis and quadrature: local_basis_volume_3D.m:60-106,
//    local_basis_volume_2D.m:40-47, quadrature_volume_3D.m:39-42,
//    quadrature_volume_2D.m:28-33 (the same constant tables).
//  * f_V = sum over points of HatP * WEIGHT * f_V_int (vector_volume_3D.m:39-62,
//    vector_volume_2D.m:37-51), one ordered sum per node.
//  * per-point material fields from a per-element material id
//    (heterogenous_materials.m:35-87).
//  * the pattern tables of CONSTITUTIVE.init_K_r_pattern (CONSTITUTIVE.m:393-464)
//    and build_scatter_map (:582-635) in this build's node-block form -- the
//    same tables, bit for bit, as the host builder in pattern.cpp.
//
// The pattern is built from radix sorts and prefix sums: the node adjacency is
// the sorted, deduplicated list of element node pairs; every per-block or
// per-node contribution list whose order matters (element order inside a
// list) comes from a stable sort of the contributions, which are generated in
// that order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <stdint.h>

#include <vector>

#include "devutil.h"
#include "p2geom.h"
#include "pattern.h"
#include "ssr_common.h"

extern "C" int ssr_pattern_destroy(ssr_pattern *p);

using namespace ssr_dev;

namespace {

// ---- P2 reference element ---------------------------------------------------
// (tables, reference derivatives and the per-point Jacobian: p2geom.h)
using p2geom::kWF2;
using p2geom::kWF3;
using p2geom::kXi2;
using p2geom::kXi3;

#pragma clang fp contract(off)

// P2 tetrahedron values (local_basis_volume_3D.m:64-73)
__device__ inline double p2_tet_hat(int a, double x1, double x2, double x3)
{
    const double x0 = 1 - x1 - x2 - x3;
    switch (a) {
    case 0: return x0 * (2 * x0 - 1);
    case 1: return x1 * (2 * x1 - 1);
    case 2: return x2 * (2 * x2 - 1);
    case 3: return x3 * (2 * x3 - 1);
    case 4: return 4 * x0 * x1;
    case 5: return 4 * x1 * x2;
    case 6: return 4 * x0 * x2;
    case 7: return 4 * x1 * x3;
    case 8: return 4 * x2 * x3;
    default: return 4 * x0 * x3;
    }
}
// P2 triangle values (local_basis_volume_2D.m:42-43)
__device__ inline double p2_tri_hat(int a, double x1, double x2)
{
    const double x0 = 1 - x1 - x2;
    switch (a) {
    case 0: return x0 * (2 * x0 - 1);
    case 1: return x1 * (2 * x1 - 1);
    case 2: return x2 * (2 * x2 - 1);
    case 3: return 4 * x1 * x2;
    case 4: return 4 * x0 * x2;
    default: return 4 * x0 * x1;
    }
}

// One lane per integration point p = e * 11 + q.  coord: 3 x n_n node-major;
// DPhi_k: 10 x n_int point-major (elastic_stiffness_matrix_3D.m:60-96,148).
__global__ void __launch_bounds__(256) k_fe_grad3(int64_t n_int, const double *__restrict__ coord,
                                                  const int32_t *__restrict__ elem, double *__restrict__ DP1,
                                                  double *__restrict__ DP2, double *__restrict__ DP3,
                                                  double *__restrict__ W)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_int) return;
    const int64_t e = p / 11;
    const int q = (int)(p - e * 11);
    double H1[10], H2[10], H3[10], iv[9], DET;
    p2geom::tet_dhat(q, H1, H2, H3);
    p2geom::tet_jinv(coord, elem + e * 10, H1, H2, H3, iv, DET);
#pragma unroll
    for (int a = 0; a < 10; ++a) {
        DP1[p * 10 + a] = p2geom::tet_grad(iv, H1, H2, H3, 0, a);
        DP2[p * 10 + a] = p2geom::tet_grad(iv, H1, H2, H3, 1, a);
        DP3[p * 10 + a] = p2geom::tet_grad(iv, H1, H2, H3, 2, a);
    }
    W[p] = fabs(DET) * kWF3[q];
}

// 2D P2 triangles, p = e * 7 + q (elastic_stiffness_matrix_2D.m:64-80,118)
__global__ void __launch_bounds__(256) k_fe_grad2(int64_t n_int, const double *__restrict__ coord,
                                                  const int32_t *__restrict__ elem, double *__restrict__ DP1,
                                                  double *__restrict__ DP2, double *__restrict__ W)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_int) return;
    const int64_t e = p / 7;
    const int q = (int)(p - e * 7);
    double H1[6], H2[6], iv[4], DET;
    p2geom::tri_dhat(q, H1, H2);
    p2geom::tri_jinv(coord, elem + e * 6, H1, H2, iv, DET);
#pragma unroll
    for (int a = 0; a < 6; ++a) {
        DP1[p * 6 + a] = p2geom::tri_grad(iv, H1, H2, 0, a);
        DP2[p * 6 + a] = p2geom::tri_grad(iv, H1, H2, 1, a);
    }
    W[p] = fabs(DET) * kWF2[q];
}

// f(:, v) = sum over the element nodes (e, a) at v, in element order, of
// sum_q HatP(a, q) * (WEIGHT(p) * f_int(:, p)).  f_rows == dim: f_int is
// dim x n_int point-major; f_rows == 1: f_int is a density along -y
// (f_V_int = [0; -gamma(; 0)], the slope drivers' gravity load).
template <int DIM, int NP, int NQ>
__global__ void __launch_bounds__(256) k_volume_force(int64_t n_n, const int32_t *__restrict__ nptr,
                                                      const int32_t *__restrict__ ncode,
                                                      const double *__restrict__ W,
                                                      const double *__restrict__ fint, int f_rows,
                                                      double *__restrict__ f)
{
    const int64_t v = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (v >= n_n) return;
    double acc[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) acc[c] = 0.0;
    for (int32_t t = nptr[v]; t < nptr[v + 1]; ++t) {
        const int32_t code = ncode[t];
        const int64_t e = code / NP;
        const int a = code - (int)(e * NP);
        for (int q = 0; q < NQ; ++q) {
            const int64_t p = e * NQ + q;
            const double h = DIM == 3 ? p2_tet_hat(a, kXi3[0][q], kXi3[1][q], kXi3[2][q])
                                      : p2_tri_hat(a, kXi2[0][q], kXi2[1][q]);
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                const double fc = f_rows == 1 ? (c == 1 ? -fint[p] : 0.0) : fint[p * DIM + c];
                acc[c] += h * (W[p] * fc);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < DIM; ++c) f[v * DIM + c] = acc[c];
}

#pragma clang fp contract(on)

// per-point fields from the per-element material id: out_k[p] = table[m * n_f + k]
__global__ void __launch_bounds__(256) k_materials(int64_t n_int, int n_q, int n_mat, int n_f,
                                                   const double *__restrict__ table,
                                                   const int32_t *__restrict__ mat_id, double *const *out,
                                                   int32_t *bad)
{
    const int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (p >= n_int) return;
    const int32_t m = mat_id[p / n_q];
    if (m < 0 || m >= n_mat) {
        *bad = 1;
        return;
    }
    for (int k = 0; k < n_f; ++k)
        if (out[k]) out[k][p] = table[(int64_t)m * n_f + k];
}

// ---- pattern ------------------------------------------------------------------
__device__ inline int32_t lower_bound_i32(const int32_t *a, int32_t lo, int32_t hi, int32_t x)
{
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (a[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

__global__ void k_check_elem(int64_t m, int64_t n_n, const int32_t *__restrict__ elem, int32_t *bad)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m && (elem[i] < 0 || elem[i] >= n_n)) *bad = 1;
}

// key = row * n_n + col for every element node pair (a, b), element-major
__global__ void k_pair_keys(int64_t n_e, int n_p, int64_t n_n, const int32_t *__restrict__ elem,
                            uint64_t *__restrict__ keys)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t npp = (int64_t)n_p * n_p;
    if (i >= n_e * npp) return;
    const int64_t e = i / npp;
    const int r = (int)(i - e * npp), a = r / n_p, b = r - a * n_p;
    keys[i] = (uint64_t)elem[e * n_p + a] * (uint64_t)n_n + (uint64_t)elem[e * n_p + b];
}

__global__ void k_first_flags(int64_t m, const uint64_t *__restrict__ keys, int32_t *__restrict__ flag)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) flag[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1 : 0;
}

__global__ void k_compact_blocks(int64_t m, int64_t n_n, const uint64_t *__restrict__ keys,
                                 const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                                 int32_t *__restrict__ col, int32_t *__restrict__ brow, int32_t *row_cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m || !flag[i]) return;
    const int32_t r = (int32_t)(keys[i] / (uint64_t)n_n), c = (int32_t)(keys[i] % (uint64_t)n_n);
    col[pos[i]] = c;
    brow[pos[i]] = r;
    atomicAdd(&row_cnt[r], 1);
}

// diagonal block of each row, transpose block of each block, upper flags
__global__ void k_block_tables(int64_t n_n, int64_t nnzb, const int32_t *__restrict__ row_ptr,
                               const int32_t *__restrict__ col, const int32_t *__restrict__ brow,
                               int32_t *__restrict__ diag, int32_t *__restrict__ tblk, int32_t *__restrict__ uflag)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < n_n) diag[k] = lower_bound_i32(col, row_ptr[k], row_ptr[k + 1], (int32_t)k);
    if (k < nnzb) {
        const int32_t r = brow[k], c = col[k];
        tblk[k] = lower_bound_i32(col, row_ptr[c], row_ptr[c + 1], r);
        uflag[k] = c >= r ? 1 : 0;
    }
}

__global__ void k_scatter_upper(int64_t nnzb, const int32_t *__restrict__ uflag, const int32_t *__restrict__ upos,
                                int32_t *__restrict__ ublk, int32_t *__restrict__ uidx)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    if (uflag[k]) {
        ublk[upos[k]] = (int32_t)k;
        uidx[k] = upos[k];
    } else {
        uidx[k] = -1;
    }
}

// emap, and for the gather form the block (key) and code of every element pair
__global__ void k_emap(int64_t n_e, int n_p, const int32_t *__restrict__ elem, const int32_t *__restrict__ row_ptr,
                       const int32_t *__restrict__ col, int32_t *__restrict__ emap, uint32_t *__restrict__ ckey,
                       int32_t *__restrict__ cval, int32_t *cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t npp = (int64_t)n_p * n_p;
    if (i >= n_e * npp) return;
    const int64_t e = i / npp;
    const int r = (int)(i - e * npp), a = r / n_p, b = r - a * n_p;
    const int32_t v = elem[e * n_p + a], w = elem[e * n_p + b];
    const int32_t blk = lower_bound_i32(col, row_ptr[v], row_ptr[v + 1], w);
    emap[i] = blk;
    if (ckey) {
        const int64_t npair = (int64_t)n_p * (n_p + 1) / 2;
        const int lo = a < b ? a : b, hi = a < b ? b : a;
        const int64_t pr = (int64_t)lo * n_p - (int64_t)lo * (lo - 1) / 2 + (hi - lo);
        ckey[i] = (uint32_t)blk;
        cval[i] = (int32_t)((e * npair + pr) * 2 + (a > b ? 1 : 0));
        atomicAdd(&cnt[blk], 1);
    }
}

// the upper block of every element pair (a <= b, row-major upper triangle)
__global__ void k_pair_upper(int64_t n_e, int n_p, const int32_t *__restrict__ elem, const int32_t *__restrict__ emap,
                             const int32_t *__restrict__ uidx, uint32_t *__restrict__ ukey,
                             int32_t *__restrict__ uval, int32_t *cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t npair = (int64_t)n_p * (n_p + 1) / 2;
    if (i >= n_e * npair) return;
    const int64_t e = i / npair;
    int pr = (int)(i - e * npair), a = 0;
    while (pr >= n_p - a) {
        pr -= n_p - a;
        ++a;
    }
    const int b = a + pr;
    const int32_t v = elem[e * n_p + a], w = elem[e * n_p + b];
    const int lo = v <= w ? a : b, hi = v <= w ? b : a;
    const int32_t u = uidx[emap[(e * n_p + lo) * n_p + hi]];
    ukey[i] = (uint32_t)u;
    uval[i] = (int32_t)i;
    atomicAdd(&cnt[u], 1);
}

__global__ void k_invert_order(int64_t m, const int32_t *__restrict__ sorted_val, int32_t *__restrict__ slot)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) slot[sorted_val[i]] = (int32_t)i;
}

// node incidences of the element nodes, generated in element order
__global__ void k_node_keys(int64_t m, const int32_t *__restrict__ elem, uint32_t *__restrict__ key,
                            int32_t *__restrict__ val, int32_t *cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m) return;
    key[i] = (uint32_t)elem[i];
    val[i] = (int32_t)i;
    atomicAdd(&cnt[elem[i]], 1);
}

__global__ void k_mask(int64_t n, const uint8_t *__restrict__ Q, double *__restrict__ mask)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) mask[i] = Q[i] ? 1.0 : 0.0;
}

// entries of K(Q,Q) in each free column (column-major find, CONSTITUTIVE.m:445)
__global__ void k_ref_count(int dim, int64_t n_n, const int32_t *__restrict__ row_ptr,
                            const int32_t *__restrict__ col, const uint8_t *__restrict__ Q, int64_t *__restrict__ cnt)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= dim * n_n) return;
    int64_t c = 0;
    if (Q[q]) {
        const int64_t cn = q / dim;
        for (int32_t k = row_ptr[cn]; k < row_ptr[cn + 1]; ++k)
            for (int ri = 0; ri < dim; ++ri) c += Q[(int64_t)dim * col[k] + ri] ? 1 : 0;
    }
    cnt[q] = c;
}

__global__ void k_ref_fill(int dim, int64_t n_n, const int32_t *__restrict__ row_ptr, const int32_t *__restrict__ col,
                           const int32_t *__restrict__ tblk, const uint8_t *__restrict__ Q,
                           const int64_t *__restrict__ off, int64_t *__restrict__ ref_pos)
{
    const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (q >= dim * n_n || !Q[q]) return;
    const int64_t cn = q / dim;
    const int cj = (int)(q - cn * dim);
    const int bs = dim * dim;
    int64_t o = off[q];
    for (int32_t k = row_ptr[cn]; k < row_ptr[cn + 1]; ++k) {
        const int32_t rn = col[k], blk = tblk[k];
        for (int ri = 0; ri < dim; ++ri)
            if (Q[(int64_t)dim * rn + ri]) ref_pos[o++] = (int64_t)blk * bs + ri * dim + cj;
    }
}

}  // namespace

extern "C" {

int ssr_dev_fe_gradients(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const double *coord, const int32_t *elem,
                         double *DPhi1, double *DPhi2, double *DPhi3, double *WEIGHT)
{
    SSR_REQUIRE((dim == 3 && n_p == 10) || (dim == 2 && n_p == 6),
                "fe_gradients: P2 elements only (n_p = 10 in 3D, 6 in 2D)");
    SSR_REQUIRE(n_n > 0 && n_e >= 0 && coord && elem && DPhi1 && DPhi2 && WEIGHT && (dim == 2 || DPhi3),
                "fe_gradients: bad arguments");
    hipStream_t s = ssr_stream();
    DevBuf bad;
    int rc = bad.alloc(sizeof(int32_t));
    if (rc) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), s));
    SSR_LAUNCH(k_check_elem, n_e * n_p, n_e * n_p, n_n, elem, bad.as<int32_t>());
    int32_t hb = 0;
    if ((rc = read_scalar(bad.as<int32_t>(), &hb, s))) return rc;
    SSR_REQUIRE(!hb, "fe_gradients: element node id out of range");
    if (dim == 3) {
        const int64_t n_int = n_e * 11;
        SSR_LAUNCH(k_fe_grad3, n_int, n_int, coord, elem, DPhi1, DPhi2, DPhi3, WEIGHT);
    } else {
        const int64_t n_int = n_e * 7;
        SSR_LAUNCH(k_fe_grad2, n_int, n_int, coord, elem, DPhi1, DPhi2, WEIGHT);
    }
    return SSR_OK;
}

int ssr_dev_volume_force(const ssr_pattern *P, int n_q, const double *WEIGHT, const double *f_int, int f_rows,
                         double *f)
{
    SSR_REQUIRE(P && WEIGHT && f_int && f, "volume_force: null argument");
    SSR_REQUIRE(P->d_nptr && P->d_ncode, "volume_force: the pattern has no node incidence table");
    SSR_REQUIRE(f_rows == 1 || f_rows == P->dim, "volume_force: f_rows must be 1 or dim");
    hipStream_t s = ssr_stream();
    if (P->dim == 3) {
        SSR_REQUIRE(P->n_p == 10 && n_q == 11, "volume_force: 3D P2 with 11 points only");
        SSR_LAUNCH((k_volume_force<3, 10, 11>), P->n_n, P->n_n, P->d_nptr, P->d_ncode, WEIGHT, f_int, f_rows, f);
    } else {
        SSR_REQUIRE(P->n_p == 6 && n_q == 7, "volume_force: 2D P2 with 7 points only");
        SSR_LAUNCH((k_volume_force<2, 6, 7>), P->n_n, P->n_n, P->d_nptr, P->d_ncode, WEIGHT, f_int, f_rows, f);
    }
    return SSR_OK;
}

int ssr_dev_materials(int64_t n_e, int n_q, int n_mat, int n_fields, const double *table, const int32_t *mat_id,
                      double *const *out_host)
{
    SSR_REQUIRE(n_q > 0 && n_mat > 0 && n_fields > 0 && n_fields <= 16 && table && mat_id && out_host,
                "materials: bad arguments");
    hipStream_t s = ssr_stream();
    DevBuf outs, bad;
    int rc;
    if ((rc = outs.alloc(sizeof(double *) * n_fields)) || (rc = bad.alloc(sizeof(int32_t)))) return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(outs.p, out_host, sizeof(double *) * n_fields, hipMemcpyHostToDevice, s));
    SSR_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), s));
    const int64_t n_int = n_e * n_q;
    SSR_LAUNCH(k_materials, n_int, n_int, n_q, n_mat, n_fields, table, mat_id, outs.as<double *>(),
               bad.as<int32_t>());
    int32_t hb = 0;
    if ((rc = read_scalar(bad.as<int32_t>(), &hb, s))) return rc;
    SSR_REQUIRE(!hb, "materials: material id out of range");
    return SSR_OK;
}

// Every table of ssr_pattern_create, built on the device from a device
// element array (n_p x n_e, int32, element-major) and a device mask Q
// (dim x n_n node-major, nonzero = free).
int ssr_pattern_create_dev(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const int32_t *elem, const uint8_t *Q,
                           ssr_pattern **out)
{
    SSR_REQUIRE(out, "pattern_create_dev: null out");
    *out = nullptr;
    SSR_REQUIRE(dim == 2 || dim == 3, "pattern_create_dev: dim must be 2 or 3");
    SSR_REQUIRE(n_n > 0 && n_p > 0 && n_e > 0 && elem && Q, "pattern_create_dev: bad sizes");
    SSR_REQUIRE(n_n < (int64_t)1 << 31, "pattern_create_dev: too many nodes for int32 indices");
    const int64_t npp = n_p * n_p, m = n_e * npp;
    SSR_REQUIRE(m < (int64_t)1 << 31, "pattern_create_dev: too many element node pairs for int32 indices");
    hipStream_t s = ssr_stream();
    int rc;
    const int64_t n_dof = dim * n_n;

    DevBuf bad;
    if ((rc = bad.alloc(sizeof(int32_t)))) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(bad.p, 0, sizeof(int32_t), s));
    SSR_LAUNCH(k_check_elem, n_e * n_p, n_e * n_p, n_n, elem, bad.as<int32_t>());
    int32_t hb = 0;
    if ((rc = read_scalar(bad.as<int32_t>(), &hb, s))) return rc;
    SSR_REQUIRE(!hb, "pattern_create_dev: element node id out of range");

    ssr_pattern *p = new ssr_pattern();
    p->dim = dim;
    p->n_n = n_n;
    p->n_p = n_p;
    p->n_e = n_e;
    p->device_built = true;
    auto fail = [&](int code) {
        ssr_pattern_destroy(p);
        return code;
    };

    // node adjacency: sorted unique element node pairs
    DevBuf keys, skeys, flag, pos, rcnt, rptr;
    int32_t nnzb = 0;
    {
        if ((rc = keys.alloc(sizeof(uint64_t) * m)) || (rc = skeys.alloc(sizeof(uint64_t) * m))) return fail(rc);
        SSR_LAUNCH(k_pair_keys, m, n_e, (int)n_p, n_n, elem, keys.as<uint64_t>());
        if ((rc = sort_keys(keys.as<uint64_t>(), skeys.as<uint64_t>(), m,
                            bits_for((uint64_t)n_n * (uint64_t)n_n - 1), s)))
            return fail(rc);
        (void)hipFree(keys.release());
        if ((rc = flag.alloc(sizeof(int32_t) * (m + 1))) || (rc = pos.alloc(sizeof(int32_t) * (m + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemsetAsync(flag.as<int32_t>() + m, 0, sizeof(int32_t), s));
        SSR_LAUNCH(k_first_flags, m, m, skeys.as<uint64_t>(), flag.as<int32_t>());
        if ((rc = exclusive_sum(flag.as<int32_t>(), pos.as<int32_t>(), m + 1, s)) ||
            (rc = read_scalar(pos.as<int32_t>() + m, &nnzb, s)))
            return fail(rc);
        p->nnzb = nnzb;
        DevBuf col, brow;
        if ((rc = col.alloc(sizeof(int32_t) * nnzb)) || (rc = brow.alloc(sizeof(int32_t) * nnzb)) ||
            (rc = rcnt.alloc(sizeof(int32_t) * (n_n + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemsetAsync(rcnt.p, 0, sizeof(int32_t) * (n_n + 1), s));
        SSR_LAUNCH(k_compact_blocks, m, m, n_n, skeys.as<uint64_t>(), flag.as<int32_t>(), pos.as<int32_t>(),
                   col.as<int32_t>(), brow.as<int32_t>(), rcnt.as<int32_t>());
        int32_t tot = 0;
        if ((rc = counts_to_ptr(rcnt, n_n, rptr, &tot, s))) return fail(rc);
        p->d_row_ptr = (int32_t *)rptr.release();
        p->d_col = (int32_t *)col.release();
        p->d_brow = (int32_t *)brow.release();
        (void)hipFree(skeys.release());
        (void)hipFree(flag.release());
        (void)hipFree(pos.release());
    }

    // diagonal, transpose and upper blocks
    DevBuf uidx;
    {
        DevBuf diag, tblk, uflag, upos, ublk;
        if ((rc = diag.alloc(sizeof(int32_t) * n_n)) || (rc = tblk.alloc(sizeof(int32_t) * nnzb)) ||
            (rc = uflag.alloc(sizeof(int32_t) * (nnzb + 1))) || (rc = upos.alloc(sizeof(int32_t) * (nnzb + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemsetAsync(uflag.as<int32_t>() + nnzb, 0, sizeof(int32_t), s));
        const int64_t nk = n_n > nnzb ? n_n : nnzb;
        SSR_LAUNCH(k_block_tables, nk, n_n, nnzb, p->d_row_ptr, p->d_col, p->d_brow, diag.as<int32_t>(),
                   tblk.as<int32_t>(), uflag.as<int32_t>());
        int32_t nu = 0;
        if ((rc = exclusive_sum(uflag.as<int32_t>(), upos.as<int32_t>(), nnzb + 1, s)) ||
            (rc = read_scalar(upos.as<int32_t>() + nnzb, &nu, s)))
            return fail(rc);
        p->n_ublk = nu;
        if ((rc = ublk.alloc(sizeof(int32_t) * nu)) || (rc = uidx.alloc(sizeof(int32_t) * nnzb))) return fail(rc);
        SSR_LAUNCH(k_scatter_upper, nnzb, nnzb, uflag.as<int32_t>(), upos.as<int32_t>(), ublk.as<int32_t>(),
                   uidx.as<int32_t>());
        p->d_diag = (int32_t *)diag.release();
        p->d_tblk = (int32_t *)tblk.release();
        p->d_ublk = (int32_t *)ublk.release();
    }

    // element block map and the gather-form contribution lists
    const int64_t npair = n_p * (n_p + 1) / 2;
    {
        const bool gather = n_e * npair * 2 < ((int64_t)1 << 31) && m < ((int64_t)1 << 31);
        DevBuf emap, ckey, cval, skey, sval, cnt, cptr;
        if ((rc = emap.alloc(sizeof(int32_t) * m))) return fail(rc);
        if (gather) {
            if ((rc = ckey.alloc(sizeof(uint32_t) * m)) || (rc = cval.alloc(sizeof(int32_t) * m)) ||
                (rc = cnt.alloc(sizeof(int32_t) * (nnzb + 1))))
                return fail(rc);
            SSR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (nnzb + 1), s));
        }
        SSR_LAUNCH(k_emap, m, n_e, (int)n_p, elem, p->d_row_ptr, p->d_col, emap.as<int32_t>(),
                   gather ? ckey.as<uint32_t>() : nullptr, gather ? cval.as<int32_t>() : nullptr,
                   gather ? cnt.as<int32_t>() : nullptr);
        p->d_emap = (int32_t *)emap.release();
        if (gather) {
            if ((rc = skey.alloc(sizeof(uint32_t) * m)) || (rc = sval.alloc(sizeof(int32_t) * m)) ||
                (rc = sort_pairs(ckey.as<uint32_t>(), skey.as<uint32_t>(), cval.as<int32_t>(), sval.as<int32_t>(), m,
                                 bits_for((uint64_t)(nnzb > 0 ? nnzb - 1 : 0)), s)))
                return fail(rc);
            int32_t tot = 0;
            if ((rc = counts_to_ptr(cnt, nnzb, cptr, &tot, s))) return fail(rc);
            p->d_cptr = (int32_t *)cptr.release();
            p->d_ccode = (int32_t *)sval.release();
        }
    }

    // slot lists of the upper blocks
    if (n_e * npair < ((int64_t)1 << 31)) {
        const int64_t mp = n_e * npair;
        DevBuf ukey, uval, skey, sval, cnt, uptr, pslot;
        if ((rc = ukey.alloc(sizeof(uint32_t) * mp)) || (rc = uval.alloc(sizeof(int32_t) * mp)) ||
            (rc = cnt.alloc(sizeof(int32_t) * (p->n_ublk + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (p->n_ublk + 1), s));
        SSR_LAUNCH(k_pair_upper, mp, n_e, (int)n_p, elem, p->d_emap, uidx.as<int32_t>(), ukey.as<uint32_t>(),
                   uval.as<int32_t>(), cnt.as<int32_t>());
        if ((rc = skey.alloc(sizeof(uint32_t) * mp)) || (rc = sval.alloc(sizeof(int32_t) * mp)) ||
            (rc = sort_pairs(ukey.as<uint32_t>(), skey.as<uint32_t>(), uval.as<int32_t>(), sval.as<int32_t>(), mp,
                             bits_for((uint64_t)(p->n_ublk > 0 ? p->n_ublk - 1 : 0)), s)))
            return fail(rc);
        if ((rc = pslot.alloc(sizeof(int32_t) * mp))) return fail(rc);
        SSR_LAUNCH(k_invert_order, mp, mp, sval.as<int32_t>(), pslot.as<int32_t>());
        int32_t tot = 0;
        if ((rc = counts_to_ptr(cnt, p->n_ublk, uptr, &tot, s))) return fail(rc);
        p->d_uptr = (int32_t *)uptr.release();
        p->d_pslot = (int32_t *)pslot.release();
    }
    (void)hipFree(uidx.release());

    // node incidences for the gather-form internal force
    if (n_e * n_p < ((int64_t)1 << 31)) {
        const int64_t mn = n_e * n_p;
        DevBuf key, val, skey, sval, cnt, nptr;
        if ((rc = key.alloc(sizeof(uint32_t) * mn)) || (rc = val.alloc(sizeof(int32_t) * mn)) ||
            (rc = cnt.alloc(sizeof(int32_t) * (n_n + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (n_n + 1), s));
        SSR_LAUNCH(k_node_keys, mn, mn, elem, key.as<uint32_t>(), val.as<int32_t>(), cnt.as<int32_t>());
        if ((rc = skey.alloc(sizeof(uint32_t) * mn)) || (rc = sval.alloc(sizeof(int32_t) * mn)) ||
            (rc = sort_pairs(key.as<uint32_t>(), skey.as<uint32_t>(), val.as<int32_t>(), sval.as<int32_t>(), mn,
                             bits_for((uint64_t)(n_n - 1)), s)))
            return fail(rc);
        int32_t tot = 0;
        if ((rc = counts_to_ptr(cnt, n_n, nptr, &tot, s))) return fail(rc);
        p->d_nptr = (int32_t *)nptr.release();
        p->d_ncode = (int32_t *)sval.release();
    }

    // element copy, free mask and the reference-order extraction map
    {
        DevBuf el, mk, cnt, off, ref;
        if ((rc = el.alloc(sizeof(int32_t) * n_e * n_p)) || (rc = mk.alloc(sizeof(double) * n_dof)) ||
            (rc = cnt.alloc(sizeof(int64_t) * (n_dof + 1))) || (rc = off.alloc(sizeof(int64_t) * (n_dof + 1))))
            return fail(rc);
        SSR_HIP_CHECK(hipMemcpyAsync(el.p, elem, sizeof(int32_t) * n_e * n_p, hipMemcpyDeviceToDevice, s));
        SSR_LAUNCH(k_mask, n_dof, n_dof, Q, mk.as<double>());
        SSR_HIP_CHECK(hipMemsetAsync(cnt.as<int64_t>() + n_dof, 0, sizeof(int64_t), s));
        SSR_LAUNCH(k_ref_count, n_dof, dim, n_n, p->d_row_ptr, p->d_col, Q, cnt.as<int64_t>());
        int64_t nf = 0;
        if ((rc = exclusive_sum(cnt.as<int64_t>(), off.as<int64_t>(), n_dof + 1, s)) ||
            (rc = read_scalar(off.as<int64_t>() + n_dof, &nf, s)))
            return fail(rc);
        p->nnz_free = nf;
        if ((rc = ref.alloc(sizeof(int64_t) * (nf > 0 ? nf : 1)))) return fail(rc);
        SSR_LAUNCH(k_ref_fill, n_dof, dim, n_n, p->d_row_ptr, p->d_col, p->d_tblk, Q, off.as<int64_t>(),
                   ref.as<int64_t>());
        p->d_elem = (int32_t *)el.release();
        p->d_mask = mk.as<double>();
        mk.release();
        p->d_ref_pos = (int64_t *)ref.release();
        // the host keeps the free mask (a byte per DOF)
        p->mask.resize(n_dof);
        SSR_HIP_CHECK(hipMemcpyAsync(p->mask.data(), Q, n_dof, hipMemcpyDeviceToHost, s));
        SSR_HIP_CHECK(hipStreamSynchronize(s));
        p->n_free = 0;
        for (auto &b : p->mask) {
            b = b ? 1 : 0;
            p->n_free += b;
        }
    }
    *out = p;
    return SSR_OK;
}

}  // extern "C"
