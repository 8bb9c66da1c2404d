// Smoothed-aggregation hierarchy built on the device.
//
// Reference counterpart: the BoomerAMG setup that the binding re-runs on
// every values update (hypre_boomeramg_mex.cpp:1151-1211, setup :1200) with
// the rigid-body near-null space of near_null_space_elasticity_3D.m:46-103.
// This is the algorithm of the host build in amg.cpp (strength on node
// blocks, greedy aggregation, tentative prolongator by a QR of the near-null
// space per aggregate, one damped-Jacobi smoothing step, Galerkin products)
// with every level held as block CSR: rows and columns are whole nodes at the
// finest level and whole aggregates below it, so the prolongator has dim x k
// blocks, the restriction k x dim and the coarse operators k x k.
//
// What runs where:
//   device  D^-1 and the Gershgorin bound, the power iteration for
//           rho(D^-1 A), block norms for the strength test, the tentative
//           prolongator (one wavefront per aggregate, CGS2 with the host's
//           dropping rule), the smoothed prolongator, A*P, P' and P'(A P);
//   host    the strong node graph from the block norms and the greedy
//           aggregation (sequential by nature; the same function as the host
//           build, so the aggregates are the host build's).
// Sparse products: the block pattern of each product is the sorted, unique
// list of candidate (row, column) pairs (radix sort); each output block is then
// summed by one lane over its contributions in a fixed order, so the
// hierarchy is the same bits on every run.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <chrono>
#include <vector>

#include "amg_device.h"
#include "devutil.h"
#include "linalg.h"
#include "ssr_common.h"

using namespace ssr_dev;

namespace ssr_amg {

void DevBlocks::release()
{
    (void)hipFree(brp);
    (void)hipFree(bcol);
    (void)hipFree(brow);
    (void)hipFree(bval);
    brp = bcol = brow = nullptr;
    bval = nullptr;
    nbr = nblk = 0;
}

void DevLevelOut::release()
{
    A.release();
    P.release();
    R.release();
    (void)hipFree(diag);
    (void)hipFree(dinv);
    diag = nullptr;
    dinv = nullptr;
}

std::vector<int32_t> aggregate_graph(int64_t nn, const std::vector<int64_t> &sp, const std::vector<int32_t> &sc,
                                     int32_t &n_agg)
{
    std::vector<int32_t> agg(nn, -1);
    n_agg = 0;
    // pass 1: roots whose strong neighbours are all free
    for (int64_t i = 0; i < nn; ++i) {
        if (agg[i] >= 0 || sp[i + 1] == sp[i]) continue;
        bool ok = true;
        for (int64_t t = sp[i]; t < sp[i + 1] && ok; ++t) ok = agg[sc[t]] < 0;
        if (!ok) continue;
        agg[i] = n_agg;
        for (int64_t t = sp[i]; t < sp[i + 1]; ++t) agg[sc[t]] = n_agg;
        n_agg++;
    }
    // pass 2: attach to a neighbouring aggregate (first aggregated strong neighbour)
    std::vector<int32_t> agg2 = agg;
    for (int64_t i = 0; i < nn; ++i) {
        if (agg[i] >= 0) continue;
        int32_t best = -1;
        for (int64_t t = sp[i]; t < sp[i + 1]; ++t)
            if (agg[sc[t]] >= 0) {
                best = agg[sc[t]];
                break;
            }
        agg2[i] = best;
    }
    agg.swap(agg2);
    // pass 3: leftovers with strong neighbours form new aggregates
    for (int64_t i = 0; i < nn; ++i) {
        if (agg[i] >= 0 || sp[i + 1] == sp[i]) continue;
        agg[i] = n_agg;
        for (int64_t t = sp[i]; t < sp[i + 1]; ++t)
            if (agg[sc[t]] < 0) agg[sc[t]] = n_agg;
        n_agg++;
    }
    return agg;
}

}  // namespace ssr_amg

namespace {

using ssr_amg::DevBlocks;

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__device__ inline int32_t find_col(const int32_t *col, int32_t lo, int32_t hi, int32_t x)
{
    while (lo < hi) {
        const int32_t mid = lo + ((hi - lo) >> 1);
        if (col[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// scalar D^-1 (1 where the diagonal is zero) and max_d |D^-1_d| sum_j |a_dj|
template <int RB>
__global__ void __launch_bounds__(256) k_dinv_rowsum(int64_t n_nodes, const int32_t *__restrict__ rp,
                                                     const int32_t *__restrict__ diag, const double *__restrict__ val,
                                                     double *__restrict__ dinv, unsigned long long *rmax)
{
    const int64_t d = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (d >= n_nodes * RB) return;
    const int64_t i = d / RB;
    const int r = (int)(d - i * RB);
    const double a = val[(int64_t)diag[i] * RB * RB + r * RB + r];
    const double di = a != 0.0 ? 1.0 / a : 1.0;
    dinv[d] = di;
    double s = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const double *b = val + (int64_t)k * RB * RB + r * RB;
#pragma unroll
        for (int c = 0; c < RB; ++c) s += fabs(b[c]);
    }
    const double g = s * fabs(di);
    atomicMax(rmax, (unsigned long long)__double_as_longlong(g));  // g >= 0: bits order as values
}

__global__ void k_block_sumsq(int64_t nblk, int bs, const double *__restrict__ val, double *__restrict__ s2)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    double s = 0.0;
    for (int t = 0; t < bs; ++t) {
        const double v = val[k * bs + t];
        s += v * v;
    }
    s2[k] = s;
}

// start vector of the power iteration (the host build's)
__global__ void k_power_init(int64_t n, double *__restrict__ x)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < n) x[i] = 1.0 + 0.1 * (double)((i * 7919) % 13);
}

// Tentative prolongator: one wavefront per aggregate.  The near-null space
// restricted to the aggregate's unknowns (nodes in node order, components in
// order) is orthonormalised column by column (CGS2); a column whose norm
// after projection is at most 1e-10 of its initial norm is dropped.  Q goes to
// T (node-indexed RB x K blocks), the K x K factor R to Rf.
template <int RB, int K>
__global__ void __launch_bounds__(64) k_tentative(int64_t n_agg, const int32_t *__restrict__ aptr,
                                                  const int32_t *__restrict__ anode, const double *__restrict__ B,
                                                  int64_t n, double *__restrict__ T, double *__restrict__ Rf,
                                                  double *__restrict__ v, int32_t *dropped)
{
    const int64_t a = blockIdx.x;
    if (a >= n_agg) return;
    const int lane = threadIdx.x;
    const int32_t t0 = aptr[a];
    const int m = (aptr[a + 1] - t0) * RB;
    // R (K x K, rows = kept columns) is written by lane 0 as it forms
    double *Ra = Rf + a * K * K;
    if (lane == 0)
        for (int x = 0; x < K * K; ++x) Ra[x] = 0.0;
    int kept = 0;
    for (int c = 0; c < K; ++c) {
        double s = 0.0;
        for (int q = lane; q < m; q += 64) {
            const int64_t dof = (int64_t)anode[t0 + q / RB] * RB + q % RB;
            const double x = B[(int64_t)c * n + dof];
            v[dof] = x;
            s += x * x;
        }
        const double n0 = sqrt(wave_sum(s));
        double rcol[K];
#pragma unroll
        for (int q = 0; q < K; ++q) rcol[q] = 0.0;
        for (int pass = 0; pass < 2; ++pass)
            for (int qq = 0; qq < kept; ++qq) {
                double d = 0.0;
                for (int q = lane; q < m; q += 64) {
                    const int64_t dof = (int64_t)anode[t0 + q / RB] * RB + q % RB;
                    d += T[dof * K + qq] * v[dof];
                }
                d = wave_sum(d);
                for (int q = lane; q < m; q += 64) {
                    const int64_t dof = (int64_t)anode[t0 + q / RB] * RB + q % RB;
                    v[dof] -= d * T[dof * K + qq];
                }
#pragma unroll
                for (int x = 0; x < K; ++x)
                    if (x == qq) rcol[x] += d;
            }
        double s2 = 0.0;
        for (int q = lane; q < m; q += 64) {
            const int64_t dof = (int64_t)anode[t0 + q / RB] * RB + q % RB;
            s2 += v[dof] * v[dof];
        }
        const double nv = sqrt(wave_sum(s2));
        if (lane == 0) {
#pragma unroll
            for (int q = 0; q < K; ++q)
                if (q < kept) Ra[q * K + c] = rcol[q];
        }
        if (nv > 1e-10 * n0 && nv > 1e-300) {
            for (int q = lane; q < m; q += 64) {
                const int64_t dof = (int64_t)anode[t0 + q / RB] * RB + q % RB;
                T[dof * K + kept] = v[dof] / nv;
            }
            if (lane == 0) Ra[kept * K + c] = nv;
            kept++;
        }
    }
    if (lane == 0 && kept < K) atomicAdd(dropped, 1);
}

// candidate blocks of P in node row i: its own aggregate and those of its
// neighbours (sentinel for unaggregated nodes)
__global__ void k_p_keys(int64_t n_nodes, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                         const int32_t *__restrict__ agg, int64_t n_agg, uint64_t *__restrict__ keys)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const uint64_t sentinel = (uint64_t)n_nodes * (uint64_t)n_agg;
    int64_t o = i + rp[i];
    keys[o++] = agg[i] >= 0 ? (uint64_t)i * n_agg + agg[i] : sentinel;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int32_t g = agg[col[k]];
        keys[o++] = g >= 0 ? (uint64_t)i * n_agg + g : sentinel;
    }
}

// sorted keys -> first-occurrence flags (sentinels excluded)
__global__ void k_unique_flags(int64_t m, uint64_t sentinel, const uint64_t *__restrict__ keys,
                               int32_t *__restrict__ flag)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i < m) flag[i] = (keys[i] != sentinel && (i == 0 || keys[i] != keys[i - 1])) ? 1 : 0;
}

__global__ void k_unique_compact(int64_t m, uint64_t n_cols, const uint64_t *__restrict__ keys,
                                 const int32_t *__restrict__ flag, const int32_t *__restrict__ pos,
                                 int32_t *__restrict__ bcol, int32_t *__restrict__ brow, int32_t *cnt)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= m || !flag[i]) return;
    const int32_t r = (int32_t)(keys[i] / n_cols);
    bcol[pos[i]] = (int32_t)(keys[i] % n_cols);
    brow[pos[i]] = r;
    atomicAdd(&cnt[r], 1);
}

// P = T - omega D^-1 A T, block (i, a): T_i when a is i's aggregate, minus the
// row-scaled sum over the neighbours j in a, in column order
template <int RB, int K>
__global__ void __launch_bounds__(256) k_p_values(int64_t nblk, const int32_t *__restrict__ brow,
                                                  const int32_t *__restrict__ bcol, const int32_t *__restrict__ rp,
                                                  const int32_t *__restrict__ col, const double *__restrict__ val,
                                                  const int32_t *__restrict__ agg, const double *__restrict__ T,
                                                  const double *__restrict__ dinv, double omega,
                                                  double *__restrict__ pval)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int64_t i = brow[t];
    const int32_t a = bcol[t];
    double acc[RB][K];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < K; ++c) acc[r][c] = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int32_t j = col[k];
        if (agg[j] != a) continue;
        const double *A = val + (int64_t)k * RB * RB;
        const double *Tj = T + (int64_t)j * RB * K;
#pragma unroll
        for (int r = 0; r < RB; ++r)
#pragma unroll
            for (int s = 0; s < RB; ++s) {
                const double ars = A[r * RB + s];
#pragma unroll
                for (int c = 0; c < K; ++c) acc[r][c] += ars * Tj[s * K + c];
            }
    }
    const bool own = agg[i] == a;
#pragma unroll
    for (int r = 0; r < RB; ++r) {
        const double sc = omega * dinv[i * RB + r];
#pragma unroll
        for (int c = 0; c < K; ++c) {
            const double base = own ? T[(i * RB + r) * K + c] : 0.0;
            pval[(t * RB + r) * K + c] = base - sc * acc[r][c];
        }
    }
}

// number of candidate blocks that block k (row i, column j) of the left
// factor contributes: the blocks of row j of the right factor
__global__ void k_rowlen_of_col(int64_t nblk, const int32_t *__restrict__ bcol, const int32_t *__restrict__ rrp,
                                int64_t *__restrict__ cnt)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k < nblk) cnt[k] = rrp[bcol[k] + 1] - rrp[bcol[k]];
}

__global__ void k_product_keys(int64_t nblk, const int32_t *__restrict__ brow, const int32_t *__restrict__ bcol,
                               const int32_t *__restrict__ rrp, const int32_t *__restrict__ rcol,
                               const int64_t *__restrict__ off, uint64_t n_cols, uint64_t *__restrict__ keys)
{
    const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (k >= nblk) return;
    const uint64_t i = (uint64_t)brow[k];
    const int32_t j = bcol[k];
    int64_t o = off[k];
    for (int32_t t = rrp[j]; t < rrp[j + 1]; ++t) keys[o++] = i * n_cols + (uint64_t)rcol[t];
}

// C = A B, block (i, b): sum over the blocks (i, j) of A, in column order, of
// A_ij B_jb (B_jb found by binary search in row j of B)
template <int M, int N, int L>
__global__ void __launch_bounds__(256) k_product_values(int64_t nblk, const int32_t *__restrict__ brow,
                                                        const int32_t *__restrict__ bcol,
                                                        const int32_t *__restrict__ arp,
                                                        const int32_t *__restrict__ acol,
                                                        const double *__restrict__ aval,
                                                        const int32_t *__restrict__ brp_,
                                                        const int32_t *__restrict__ bcol_,
                                                        const double *__restrict__ bval_, double *__restrict__ cval)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    const int32_t i = brow[t], b = bcol[t];
    double acc[M][L];
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int c = 0; c < L; ++c) acc[r][c] = 0.0;
    for (int32_t k = arp[i]; k < arp[i + 1]; ++k) {
        const int32_t j = acol[k];
        const int32_t lo = brp_[j], hi = brp_[j + 1];
        const int32_t s = find_col(bcol_, lo, hi, b);
        if (s >= hi || bcol_[s] != b) continue;
        const double *A = aval + (int64_t)k * M * N;
        const double *Bm = bval_ + (int64_t)s * N * L;
#pragma unroll
        for (int r = 0; r < M; ++r)
#pragma unroll
            for (int q = 0; q < N; ++q) {
                const double arq = A[r * N + q];
#pragma unroll
                for (int c = 0; c < L; ++c) acc[r][c] += arq * Bm[q * L + c];
            }
    }
#pragma unroll
    for (int r = 0; r < M; ++r)
#pragma unroll
        for (int c = 0; c < L; ++c) cval[((int64_t)t * M + r) * L + c] = acc[r][c];
}

// R = P': block u of R is block src[u] of P transposed; its column is that
// block's row
template <int RB, int K>
__global__ void k_transpose_blocks(int64_t nblk, const int32_t *__restrict__ src, const int32_t *__restrict__ pbrow,
                                   const double *__restrict__ pval, int32_t *__restrict__ rcol,
                                   double *__restrict__ rval)
{
    const int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (u >= nblk) return;
    const int32_t t = src[u];
    rcol[u] = pbrow[t];
#pragma unroll
    for (int r = 0; r < RB; ++r)
#pragma unroll
        for (int c = 0; c < K; ++c) rval[((int64_t)u * K + c) * RB + r] = pval[((int64_t)t * RB + r) * K + c];
}

__global__ void k_keys_of_cols(int64_t nblk, const int32_t *__restrict__ bcol, uint32_t *__restrict__ key,
                               int32_t *__restrict__ val, int32_t *cnt)
{
    const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (t >= nblk) return;
    key[t] = (uint32_t)bcol[t];
    val[t] = (int32_t)t;
    atomicAdd(&cnt[bcol[t]], 1);
}

__global__ void k_diag_index(int64_t n_rows, const int32_t *__restrict__ rp, const int32_t *__restrict__ col,
                             int32_t *__restrict__ diag, int32_t *missing)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_rows) return;
    const int32_t k = find_col(col, rp[i], rp[i + 1], (int32_t)i);
    if (k >= rp[i + 1] || col[k] != (int32_t)i) *missing = 1;
    diag[i] = k;
}

// near-null space of the next level (n_agg K x K column-major) from the
// factors R of the aggregates
__global__ void k_next_B(int64_t n_agg, int K, const double *__restrict__ Rf, double *__restrict__ Bc)
{
    const int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    const int64_t nc = n_agg * K;
    if (x >= nc * K) return;
    const int64_t c = x / nc, row = x - c * nc;
    Bc[x] = Rf[row * K + c];
}

// node-block Jacobi Gershgorin bound (the host build's block_gershgorin)
template <int DIM>
__global__ void __launch_bounds__(256) k_block_gershgorin(int64_t n_nodes, const int32_t *__restrict__ rp,
                                                          const int32_t *__restrict__ diag,
                                                          const double *__restrict__ val, unsigned long long *rmax)
{
    const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
    if (i >= n_nodes) return;
    const double *D = val + (int64_t)diag[i] * DIM * DIM;
    double Bi[DIM * DIM];
    if (DIM == 2) {
        const double det = D[0] * D[3] - D[1] * D[2];
        Bi[0] = D[3] / det;
        Bi[1] = -D[1] / det;
        Bi[2] = -D[2] / det;
        Bi[3] = D[0] / det;
    } else {
        const double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8], c02 = D[3] * D[7] - D[4] * D[6];
        const double id = 1.0 / (D[0] * c00 + D[1] * c01 + D[2] * c02);
        const double b[9] = {c00 * id, (D[2] * D[7] - D[1] * D[8]) * id, (D[1] * D[5] - D[2] * D[4]) * id,
                             c01 * id, (D[0] * D[8] - D[2] * D[6]) * id, (D[2] * D[3] - D[0] * D[5]) * id,
                             c02 * id, (D[1] * D[6] - D[0] * D[7]) * id, (D[0] * D[4] - D[1] * D[3]) * id};
#pragma unroll
        for (int t = 0; t < DIM * DIM; ++t) Bi[t] = b[t];
    }
    double acc[DIM];
#pragma unroll
    for (int r = 0; r < DIM; ++r) acc[r] = 0.0;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const double *A = val + (int64_t)k * DIM * DIM;
#pragma unroll
        for (int r = 0; r < DIM; ++r)
#pragma unroll
            for (int c = 0; c < DIM; ++c) {
                double v = 0.0;
#pragma unroll
                for (int q = 0; q < DIM; ++q) v += Bi[r * DIM + q] * A[q * DIM + c];
                acc[r] += fabs(v);
            }
    }
    double mx = 0.0;
#pragma unroll
    for (int r = 0; r < DIM; ++r) mx = acc[r] > mx ? acc[r] : mx;
    if (mx == mx) atomicMax(rmax, (unsigned long long)__double_as_longlong(mx));
}

// ---- host helpers --------------------------------------------------------------
using Clock = std::chrono::steady_clock;

double lap(Clock::time_point &t0)
{
    const auto t1 = Clock::now();
    const double s = std::chrono::duration<double>(t1 - t0).count();
    t0 = t1;
    return s;
}

// sorted unique (row, column) pairs of m candidate keys row * n_cols + col
// (sentinel n_rows * n_cols skipped) -> block CSR structure
int pairs_to_blocks(DevBuf &keys, int64_t m, int64_t n_rows, int64_t n_cols, DevBlocks &out, hipStream_t s)
{
    int rc;
    SSR_REQUIRE(m < ((int64_t)1 << 31), "AMG device setup: %lld candidate blocks exceed int32 indices",
                (long long)m);
    const uint64_t sentinel = (uint64_t)n_rows * (uint64_t)n_cols;
    DevBuf sk, flag, pos, cnt, brp, bcol, brow;
    if ((rc = sk.alloc(sizeof(uint64_t) * (m > 0 ? m : 1)))) return rc;
    if (m > 0 && (rc = sort_keys(keys.as<uint64_t>(), sk.as<uint64_t>(), m, bits_for(sentinel), s))) return rc;
    (void)hipFree(keys.release());
    if ((rc = flag.alloc(sizeof(int32_t) * (m + 1))) || (rc = pos.alloc(sizeof(int32_t) * (m + 1)))) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(flag.as<int32_t>() + m, 0, sizeof(int32_t), s));
    SSR_LAUNCH(k_unique_flags, m, m, sentinel, sk.as<uint64_t>(), flag.as<int32_t>());
    int32_t nblk = 0;
    if ((rc = exclusive_sum(flag.as<int32_t>(), pos.as<int32_t>(), m + 1, s)) ||
        (rc = read_scalar(pos.as<int32_t>() + m, &nblk, s)))
        return rc;
    if ((rc = bcol.alloc(sizeof(int32_t) * (nblk > 0 ? nblk : 1))) ||
        (rc = brow.alloc(sizeof(int32_t) * (nblk > 0 ? nblk : 1))) ||
        (rc = cnt.alloc(sizeof(int32_t) * (n_rows + 1))))
        return rc;
    SSR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (n_rows + 1), s));
    SSR_LAUNCH(k_unique_compact, m, m, (uint64_t)n_cols, sk.as<uint64_t>(), flag.as<int32_t>(), pos.as<int32_t>(),
               bcol.as<int32_t>(), brow.as<int32_t>(), cnt.as<int32_t>());
    int32_t tot = 0;
    if ((rc = counts_to_ptr(cnt, n_rows, brp, &tot, s))) return rc;
    out.release();
    out.nbr = n_rows;
    out.nblk = nblk;
    out.brp = (int32_t *)brp.release();
    out.bcol = (int32_t *)bcol.release();
    out.brow = (int32_t *)brow.release();
    return SSR_OK;
}

// structure of C = A B from the block structures of A (with block rows) and B
int product_structure(const DevBlocks &A, const DevBlocks &B, int64_t n_cols, DevBlocks &C, hipStream_t s)
{
    int rc;
    DevBuf cnt, off, keys;
    const int64_t nb = A.nblk;
    if ((rc = cnt.alloc(sizeof(int64_t) * (nb + 1))) || (rc = off.alloc(sizeof(int64_t) * (nb + 1)))) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(cnt.as<int64_t>() + nb, 0, sizeof(int64_t), s));
    SSR_LAUNCH(k_rowlen_of_col, nb, nb, A.bcol, B.brp, cnt.as<int64_t>());
    int64_t m = 0;
    if ((rc = exclusive_sum(cnt.as<int64_t>(), off.as<int64_t>(), nb + 1, s)) ||
        (rc = read_scalar(off.as<int64_t>() + nb, &m, s)))
        return rc;
    if ((rc = keys.alloc(sizeof(uint64_t) * (m > 0 ? m : 1)))) return rc;
    SSR_LAUNCH(k_product_keys, nb, nb, A.brow, A.bcol, B.brp, B.bcol, off.as<int64_t>(), (uint64_t)n_cols,
               keys.as<uint64_t>());
    (void)hipFree(cnt.release());
    (void)hipFree(off.release());
    return pairs_to_blocks(keys, m, A.nbr, n_cols, C, s);
}

template <int M, int N, int L>
int product_values(const DevBlocks &A, const DevBlocks &B, DevBlocks &C, hipStream_t s)
{
    SSR_HIP_CHECK(hipMalloc((void **)&C.bval, sizeof(double) * (C.nblk > 0 ? C.nblk : 1) * M * L));
    SSR_LAUNCH((k_product_values<M, N, L>), C.nblk, C.nblk, C.brow, C.bcol, A.brp, A.bcol, A.bval, B.brp, B.bcol,
               B.bval, C.bval);
    C.rb = M;
    C.cb = L;
    return SSR_OK;
}

// rho(D^-1 A) by 30 power iterations from the host build's start vector
int power_rho(int rb, const DevBlocks &A, bool fine, const double *dinv, int its, double *rho, hipStream_t s)
{
    const int64_t n = A.nbr * rb;
    DevBuf x, y, sc;
    int rc;
    if ((rc = x.alloc(sizeof(double) * n)) || (rc = y.alloc(sizeof(double) * n)) ||
        (rc = sc.alloc(sizeof(double) * 2)))
        return rc;
    SSR_LAUNCH(k_power_init, n, n, x.as<double>());
    double *nrm = sc.as<double>(), *lam = sc.as<double>() + 1;
    double *px = x.as<double>(), *py = y.as<double>();
    for (int k = 0; k < its; ++k) {
        if ((rc = ssr::dot<double>(n, px, px, nullptr, nrm)) ||
            (rc = ssr::scale_inv_sqrt<double>(n, nrm, 0.0, px)))
            return rc;
        if (fine)
            rc = ssr::spmv_bsr<double>(rb, A.nbr, A.brp, A.bcol, A.bval, px, py, 1.0, 0.0);
        else
            rc = ssr::spmv_brc<double>(A.rb, A.cb, A.nbr, A.brp, A.bcol, A.bval, px, py, 1.0, 0.0, nullptr, A.nblk);
        if (rc || (rc = ssr::mul<double>(n, dinv, py)) || (rc = ssr::dot<double>(n, px, py, nullptr, lam)))
            return rc;
        double *t = px;
        px = py;
        py = t;
    }
    double h = 0.0;
    if ((rc = read_scalar(lam, &h, s))) return rc;
    *rho = h > 0 ? h : 1.0;
    return SSR_OK;
}

template <int RB>
int dinv_rowsum(const DevBlocks &A, const int32_t *diag, double *dinv, double *rho_up, hipStream_t s)
{
    DevBuf mx;
    int rc;
    if ((rc = mx.alloc(sizeof(unsigned long long)))) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(mx.p, 0, sizeof(unsigned long long), s));
    SSR_LAUNCH(k_dinv_rowsum<RB>, A.nbr * RB, A.nbr, A.brp, diag, A.bval, dinv, mx.as<unsigned long long>());
    unsigned long long b = 0;
    if ((rc = read_scalar(mx.as<unsigned long long>(), &b, s))) return rc;
    double g;
    memcpy(&g, &b, sizeof(g));
    *rho_up = g > 0 ? g : 1.0;
    return SSR_OK;
}

// One level: A (rb x rb blocks, with block rows), its diag and near-null
// space B (n x K) -> out (dinv, rho, rho_up, P, R), the next operator Ac with
// its diag, and its near-null space Bc.  Returns kNotUniform when an aggregate
// drops a near-null vector, or 1 in *stop when the level is the coarsest.
template <int RB, int K>
int build_level(const DevBlocks &A, bool fine, const int32_t *diag, const std::vector<int32_t> &h_rp,
                const std::vector<int32_t> &h_col, const double *B, const ssr_amg::DevOptions &opt, bool last,
                ssr_amg::DevLevelOut &out, DevBlocks &Ac, int32_t **Ac_diag, double **Bc, int *stop, double *ph,
                hipStream_t s)
{
    int rc;
    const int64_t nn = A.nbr, n = nn * RB;
    auto t0 = Clock::now();
    *stop = 0;
    out.rb = RB;
    out.n_nodes = nn;
    SSR_HIP_CHECK(hipMalloc((void **)&out.dinv, sizeof(double) * n));
    if ((rc = dinv_rowsum<RB>(A, diag, out.dinv, &out.rho_up, s))) return rc;
    if ((rc = power_rho(RB, A, fine, out.dinv, 30, &out.rho, s))) return rc;
    ph[0] += lap(t0);
    if (n <= opt.max_coarse || last) {
        *stop = 1;
        return SSR_OK;
    }
    // strength: block norms -> host strong graph -> aggregation
    std::vector<double> s2(A.nblk);
    {
        DevBuf d2;
        if ((rc = d2.alloc(sizeof(double) * (A.nblk > 0 ? A.nblk : 1)))) return rc;
        SSR_LAUNCH(k_block_sumsq, A.nblk, A.nblk, RB * RB, A.bval, d2.as<double>());
        SSR_HIP_CHECK(hipMemcpyAsync(s2.data(), d2.p, sizeof(double) * A.nblk, hipMemcpyDeviceToHost, s));
        SSR_HIP_CHECK(hipStreamSynchronize(s));
    }
    std::vector<int64_t> sp(nn + 1, 0);
    std::vector<int32_t> sc;
    sc.reserve(A.nblk);
    {
        std::vector<double> dn(nn, 0.0);
        for (int64_t i = 0; i < nn; ++i)
            for (int32_t k = h_rp[i]; k < h_rp[i + 1]; ++k)
                if (h_col[k] == i) dn[i] = s2[k];
        const double th2 = opt.theta * opt.theta;
        for (int64_t i = 0; i < nn; ++i) {
            for (int32_t k = h_rp[i]; k < h_rp[i + 1]; ++k) {
                const int32_t j = h_col[k];
                if (j == i) continue;
                if (s2[k] > 0.0 && s2[k] >= th2 * sqrt(dn[i] * dn[j])) sc.push_back(j);
            }
            sp[i + 1] = (int64_t)sc.size();
        }
    }
    int32_t n_agg = 0;
    std::vector<int32_t> agg = ssr_amg::aggregate_graph(nn, sp, sc, n_agg);
    ph[1] += lap(t0);
    if (n_agg == 0 || (int64_t)n_agg * K >= 0.9 * n) {
        *stop = 1;
        return SSR_OK;
    }
    // aggregate node lists (node order) and the tentative prolongator
    std::vector<int32_t> aptr(n_agg + 1, 0), anode;
    for (int64_t i = 0; i < nn; ++i)
        if (agg[i] >= 0) aptr[agg[i] + 1]++;
    for (int32_t a = 0; a < n_agg; ++a) aptr[a + 1] += aptr[a];
    anode.resize(aptr[n_agg]);
    {
        std::vector<int32_t> f(aptr.begin(), aptr.end() - 1);
        for (int64_t i = 0; i < nn; ++i)
            if (agg[i] >= 0) anode[f[agg[i]]++] = (int32_t)i;
    }
    DevBuf d_agg, d_aptr, d_anode, T, Rf, v, dropped;
    if ((rc = d_agg.alloc(sizeof(int32_t) * nn)) || (rc = d_aptr.alloc(sizeof(int32_t) * (n_agg + 1))) ||
        (rc = d_anode.alloc(sizeof(int32_t) * (anode.empty() ? 1 : anode.size()))) ||
        (rc = T.alloc(sizeof(double) * n * K)) || (rc = Rf.alloc(sizeof(double) * n_agg * K * K)) ||
        (rc = v.alloc(sizeof(double) * n)) || (rc = dropped.alloc(sizeof(int32_t))))
        return rc;
    SSR_HIP_CHECK(hipMemcpyAsync(d_agg.p, agg.data(), sizeof(int32_t) * nn, hipMemcpyHostToDevice, s));
    SSR_HIP_CHECK(hipMemcpyAsync(d_aptr.p, aptr.data(), sizeof(int32_t) * (n_agg + 1), hipMemcpyHostToDevice, s));
    if (!anode.empty())
        SSR_HIP_CHECK(
            hipMemcpyAsync(d_anode.p, anode.data(), sizeof(int32_t) * anode.size(), hipMemcpyHostToDevice, s));
    SSR_HIP_CHECK(hipMemsetAsync(T.p, 0, sizeof(double) * n * K, s));
    SSR_HIP_CHECK(hipMemsetAsync(dropped.p, 0, sizeof(int32_t), s));
    k_tentative<RB, K><<<(unsigned)n_agg, 64, 0, s>>>(n_agg, d_aptr.as<int32_t>(), d_anode.as<int32_t>(), B, n,
                                                      T.as<double>(), Rf.as<double>(), v.as<double>(),
                                                      dropped.as<int32_t>());
    SSR_HIP_CHECK(hipGetLastError());
    int32_t nd = 0;
    if ((rc = read_scalar(dropped.as<int32_t>(), &nd, s))) return rc;
    ph[2] += lap(t0);
    if (nd > 0) return ssr_amg::kNotUniform;
    (void)hipFree(v.release());
    // smoothed prolongator P = T - omega D^-1 A T, omega = 4 / (3 rho)
    DevBlocks &P = out.P;
    {
        DevBuf keys;
        const int64_t m = nn + A.nblk;
        if ((rc = keys.alloc(sizeof(uint64_t) * m))) return rc;
        SSR_LAUNCH(k_p_keys, nn, nn, A.brp, A.bcol, d_agg.as<int32_t>(), (int64_t)n_agg, keys.as<uint64_t>());
        if ((rc = pairs_to_blocks(keys, m, nn, n_agg, P, s))) return rc;
        SSR_HIP_CHECK(hipMalloc((void **)&P.bval, sizeof(double) * (P.nblk > 0 ? P.nblk : 1) * RB * K));
        const double omega = 4.0 / (3.0 * out.rho);
        SSR_LAUNCH((k_p_values<RB, K>), P.nblk, P.nblk, P.brow, P.bcol, A.brp, A.bcol, A.bval, d_agg.as<int32_t>(),
                   T.as<double>(), out.dinv, omega, P.bval);
        P.rb = RB;
        P.cb = K;
    }
    (void)hipFree(T.release());
    SSR_HIP_CHECK(hipStreamSynchronize(s));
    ph[3] += lap(t0);
    // A P
    DevBlocks AP;
    if ((rc = product_structure(A, P, n_agg, AP, s)) || (rc = product_values<RB, RB, K>(A, P, AP, s))) {
        AP.release();
        return rc;
    }
    SSR_HIP_CHECK(hipStreamSynchronize(s));
    ph[4] += lap(t0);
    // R = P' (rows: aggregates; the nodes of each row ascending)
    DevBlocks &R = out.R;
    {
        DevBuf key, val, skey, sval, cnt, brp;
        if ((rc = key.alloc(sizeof(uint32_t) * P.nblk)) || (rc = val.alloc(sizeof(int32_t) * P.nblk)) ||
            (rc = skey.alloc(sizeof(uint32_t) * P.nblk)) || (rc = sval.alloc(sizeof(int32_t) * P.nblk)) ||
            (rc = cnt.alloc(sizeof(int32_t) * (n_agg + 1)))) {
            AP.release();
            return rc;
        }
        SSR_HIP_CHECK(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * (n_agg + 1), s));
        SSR_LAUNCH(k_keys_of_cols, P.nblk, P.nblk, P.bcol, key.as<uint32_t>(), val.as<int32_t>(), cnt.as<int32_t>());
        int32_t tot = 0;
        if ((rc = sort_pairs(key.as<uint32_t>(), skey.as<uint32_t>(), val.as<int32_t>(), sval.as<int32_t>(), P.nblk,
                             bits_for((uint64_t)(n_agg > 0 ? n_agg - 1 : 0)), s)) ||
            (rc = counts_to_ptr(cnt, n_agg, brp, &tot, s))) {
            AP.release();
            return rc;
        }
        R.release();
        R.rb = K;
        R.cb = RB;
        R.nbr = n_agg;
        R.nblk = P.nblk;
        R.brp = (int32_t *)brp.release();
        R.brow = (int32_t *)skey.release();  // the sorted keys are the rows
        SSR_HIP_CHECK(hipMalloc((void **)&R.bcol, sizeof(int32_t) * (R.nblk > 0 ? R.nblk : 1)));
        SSR_HIP_CHECK(hipMalloc((void **)&R.bval, sizeof(double) * (R.nblk > 0 ? R.nblk : 1) * K * RB));
        SSR_LAUNCH((k_transpose_blocks<RB, K>), R.nblk, R.nblk, sval.as<int32_t>(), P.brow, P.bval, R.bcol, R.bval);
    }
    SSR_HIP_CHECK(hipStreamSynchronize(s));
    ph[5] += lap(t0);
    // Ac = R (A P)
    if ((rc = product_structure(R, AP, n_agg, Ac, s)) || (rc = product_values<K, RB, K>(R, AP, Ac, s))) {
        AP.release();
        return rc;
    }
    AP.release();
    {
        DevBuf missing;
        if ((rc = missing.alloc(sizeof(int32_t)))) return rc;
        SSR_HIP_CHECK(hipMemsetAsync(missing.p, 0, sizeof(int32_t), s));
        SSR_HIP_CHECK(hipMalloc((void **)Ac_diag, sizeof(int32_t) * n_agg));
        SSR_LAUNCH(k_diag_index, n_agg, n_agg, Ac.brp, Ac.bcol, *Ac_diag, missing.as<int32_t>());
        SSR_HIP_CHECK(hipMalloc((void **)Bc, sizeof(double) * n_agg * K * K));
        SSR_LAUNCH(k_next_B, (int64_t)n_agg * K * K, (int64_t)n_agg, K, Rf.as<double>(), *Bc);
        int32_t miss = 0;
        if ((rc = read_scalar(missing.as<int32_t>(), &miss, s))) return rc;
        SSR_REQUIRE(!miss, "AMG device setup: a coarse operator row has no diagonal block");
    }
    ph[6] += lap(t0);
    return SSR_OK;
}

}  // namespace

namespace ssr_amg {

int amg_device_build(int dim, int64_t n_nodes, const int32_t *d_rp, const int32_t *d_col, const int32_t *d_brow,
                     const int32_t *d_diag, const double *d_val, const std::vector<int32_t> &h_rp,
                     const std::vector<int32_t> &h_col, const double *d_B, int k, const DevOptions &opt,
                     std::vector<DevLevelOut> &out, double *phase_s)
{
    hipStream_t s = ssr_stream();
    for (auto &l : out) l.release();
    out.clear();
    if (!((dim == 3 && (k == 6 || k == 3)) || (dim == 2 && k == 3))) return kNotUniform;
    int rc = SSR_OK;
    // level 0: the caller's fine operator (borrowed)
    DevBlocks A0;
    A0.rb = A0.cb = dim;
    A0.nbr = n_nodes;
    A0.nblk = (int64_t)h_col.size();
    A0.brp = const_cast<int32_t *>(d_rp);
    A0.bcol = const_cast<int32_t *>(d_col);
    A0.brow = const_cast<int32_t *>(d_brow);
    A0.bval = const_cast<double *>(d_val);
    const double *B = d_B;
    double *B_owned = nullptr;
    DevBlocks next_A;  // operator of the level about to be built (owned by it once built)
    int32_t *next_diag = nullptr;
    std::vector<int32_t> hrp, hcol;  // host pattern of the current coarse level
    const std::vector<int32_t> *prp = &h_rp, *pcol = &h_col;
    for (int lev = 0;; ++lev) {
        out.emplace_back();
        DevLevelOut &lv = out.back();
        if (lev > 0) {
            lv.A = next_A;
            lv.diag = next_diag;
            next_A = DevBlocks();
            next_diag = nullptr;
        }
        const DevBlocks &A = lev == 0 ? A0 : lv.A;
        const int32_t *diag = lev == 0 ? d_diag : lv.diag;
        DevBlocks Ac;
        int32_t *Ac_diag = nullptr;
        double *Bc = nullptr;
        int stop = 0;
        const bool last = lev + 1 >= opt.max_levels;
        const int rb = lev == 0 ? dim : k;
        if (rb == 3 && k == 6)
            rc = build_level<3, 6>(A, lev == 0, diag, *prp, *pcol, B, opt, last, lv, Ac, &Ac_diag, &Bc, &stop,
                                   phase_s, s);
        else if (rb == 6 && k == 6)
            rc = build_level<6, 6>(A, lev == 0, diag, *prp, *pcol, B, opt, last, lv, Ac, &Ac_diag, &Bc, &stop,
                                   phase_s, s);
        else if (rb == 2 && k == 3)
            rc = build_level<2, 3>(A, lev == 0, diag, *prp, *pcol, B, opt, last, lv, Ac, &Ac_diag, &Bc, &stop,
                                   phase_s, s);
        else
            rc = build_level<3, 3>(A, lev == 0, diag, *prp, *pcol, B, opt, last, lv, Ac, &Ac_diag, &Bc, &stop,
                                   phase_s, s);
        if (rc || stop) {
            Ac.release();
            (void)hipFree(Ac_diag);
            (void)hipFree(Bc);
            break;
        }
        (void)hipFree(B_owned);
        B_owned = Bc;
        B = Bc;
        next_A = Ac;
        next_diag = Ac_diag;
        hrp.resize(Ac.nbr + 1);
        hcol.resize(Ac.nblk);
        SSR_HIP_CHECK(hipMemcpyAsync(hrp.data(), Ac.brp, sizeof(int32_t) * (Ac.nbr + 1), hipMemcpyDeviceToHost, s));
        SSR_HIP_CHECK(hipMemcpyAsync(hcol.data(), Ac.bcol, sizeof(int32_t) * Ac.nblk, hipMemcpyDeviceToHost, s));
        SSR_HIP_CHECK(hipStreamSynchronize(s));
        prp = &hrp;
        pcol = &hcol;
    }
    (void)hipFree(B_owned);
    if (rc) {
        for (auto &l : out) l.release();
        out.clear();
    }
    return rc;
}

int fine_block_gershgorin(int dim, int64_t n_nodes, const int32_t *d_rp, const int32_t *d_diag, const double *d_val,
                          double *out)
{
    hipStream_t s = ssr_stream();
    DevBuf mx;
    int rc;
    if ((rc = mx.alloc(sizeof(unsigned long long)))) return rc;
    SSR_HIP_CHECK(hipMemsetAsync(mx.p, 0, sizeof(unsigned long long), s));
    if (dim == 3)
        SSR_LAUNCH(k_block_gershgorin<3>, n_nodes, n_nodes, d_rp, d_diag, d_val, mx.as<unsigned long long>());
    else
        SSR_LAUNCH(k_block_gershgorin<2>, n_nodes, n_nodes, d_rp, d_diag, d_val, mx.as<unsigned long long>());
    unsigned long long b = 0;
    if ((rc = read_scalar(mx.as<unsigned long long>(), &b, s))) return rc;
    double g;
    memcpy(&g, &b, sizeof(g));
    *out = g > 0 ? g : 1.0;
    return SSR_OK;
}

}  // namespace ssr_amg
