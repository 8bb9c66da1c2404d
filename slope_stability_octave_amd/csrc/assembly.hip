// Element-level operators for P2 elements on gfx950:
//   strain gather   E = B U                 (CONSTITUTIVE.m:218,251)
//   internal force  F = B' (WEIGHT .* S)    (CONSTITUTIVE.m:281)
//   tangent values  sum_q B_q'(w_q D_q)B_q  (assemble_K_tangent_vals.c:85-171,
//                                            assemble_K_tangent_vals_2D.c:88-160)
//   K_r = r K_el + (1-r) K_t, symmetrised  (CONSTITUTIVE.m:515-532)
//
// B is never formed: each integration point carries the physical gradients of
// the n_p basis functions (DPhi1..3, n_p x n_int, as produced by
// elastic_stiffness_matrix_{2D,3D}.m), and the Voigt rows of B are applied
// structurally (3 nonzeros per column in 3D, 2 in 2D).
//
// Tangent kernel: one 64-lane wavefront per element (3D; three elements per
// wavefront in 2D).  The element's gradients, w*D for all points and, per
// point, the products D*B_b of every node are staged in LDS; each lane owns
// one node pair (a,b) with a <= b (55 pairs for a 10-node tetrahedron) and
// accumulates its dim x dim block over the points.  Ke is symmetric, so the
// lane adds its block at (a,b) and its transpose at (b,a).  Accumulation into
// the global values uses hardware FP64 atomics.
#include "p2geom.h"
#include "ssr_common.h"

#include <algorithm>

struct ssr_pattern;
const int32_t *ssr_pattern_d_row_ptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_col(const ssr_pattern *p);
const int32_t *ssr_pattern_d_emap(const ssr_pattern *p);
const int32_t *ssr_pattern_d_tblk(const ssr_pattern *p);
const int32_t *ssr_pattern_d_ublk(const ssr_pattern *p);
const int32_t *ssr_pattern_d_uptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_pslot(const ssr_pattern *p);
int64_t ssr_pattern_n_ublk(const ssr_pattern *p);
const int32_t *ssr_pattern_d_diag(const ssr_pattern *p);
const int32_t *ssr_pattern_d_elem(const ssr_pattern *p);
const int32_t *ssr_pattern_d_brow(const ssr_pattern *p);
const int64_t *ssr_pattern_d_ref_pos(const ssr_pattern *p);
const double *ssr_pattern_d_mask(const ssr_pattern *p);
const int32_t *ssr_pattern_d_cptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_ccode(const ssr_pattern *p);
int ssr_pattern_keblk(const ssr_pattern *p, double **out);
const int32_t *ssr_pattern_d_nptr(const ssr_pattern *p);
const int32_t *ssr_pattern_d_ncode(const ssr_pattern *p);
int ssr_pattern_fe(const ssr_pattern *p, double **out);
void ssr_pattern_sizes(const ssr_pattern *p, int *dim, int64_t *n_n, int64_t *n_p, int64_t *n_e,
                       int64_t *nnzb, int64_t *nnz_free);

namespace {

template <int DIM> struct Voigt;
// 3D: strain rows (e11,e22,e33,g12,g23,g13); column of node displacement
// component c has entries (row, gradient component):
//   c=x: (0,x) (3,y) (5,z);  c=y: (1,y) (3,x) (4,z);  c=z: (2,z) (4,y) (5,x)
template <> struct Voigt<3> {
    static constexpr int NS = 6, NNZ = 3;
    __device__ static constexpr int row(int c, int k)
    {
        return c == 0 ? (k == 0 ? 0 : (k == 1 ? 3 : 5)) : c == 1 ? (k == 0 ? 1 : (k == 1 ? 3 : 4))
                                                                  : (k == 0 ? 2 : (k == 1 ? 4 : 5));
    }
    __device__ static constexpr int grad(int c, int k)
    {
        return c == 0 ? (k == 0 ? 0 : (k == 1 ? 1 : 2)) : c == 1 ? (k == 0 ? 1 : (k == 1 ? 0 : 2))
                                                                  : (k == 0 ? 2 : (k == 1 ? 1 : 0));
    }
};
// 2D: rows (e11,e22,g12); c=x: (0,x) (2,y); c=y: (1,y) (2,x)
template <> struct Voigt<2> {
    static constexpr int NS = 3, NNZ = 2;
    __device__ static constexpr int row(int c, int k) { return c == 0 ? (k == 0 ? 0 : 2) : (k == 0 ? 1 : 2); }
    __device__ static constexpr int grad(int c, int k) { return c == 0 ? (k == 0 ? 0 : 1) : (k == 0 ? 1 : 0); }
};

__device__ __forceinline__ int symidx(int i, int j, int n)
{
    if (i > j) { int t = i; i = j; j = t; }
    return i * n - (i * (i - 1)) / 2 + (j - i);
}

__device__ __forceinline__ void atomic_add_f64(double *p, double v)
{
    // hardware global_atomic_add_f64 (built with -munsafe-fp-atomics)
    atomicAdd(p, v);
}

// ---------------------------------------------------------------------------
// Tangent assembly
// ---------------------------------------------------------------------------
// DSL: 0 = MEX layout (ns*ns column-major per point), 1 = packed symmetric.
// SCAT: 0 = node-block CSR via emap (atomic adds), 1 = reference int64
// scatter map (atomic adds), 2 = plain stores of the element's pair blocks
// to out[(e * NPAIR + pair) * dim^2] for the gather pass.
template <int DIM, int NP, int NQ, int DSL, int SCAT>
__global__ __launch_bounds__(64) void k_assemble(
    int64_t n_e, const double *__restrict__ D1, const double *__restrict__ D2,
    const double *__restrict__ D3, const double *__restrict__ DS, const double *__restrict__ W,
    const int32_t *__restrict__ emap, const int64_t *__restrict__ smap, double *__restrict__ out)
{
    using V = Voigt<DIM>;
    constexpr int NS = V::NS, NSYM = NS * (NS + 1) / 2, NPAIR = NP * (NP + 1) / 2;
    constexpr int EPB = 64 / NPAIR;  // elements per wavefront
    constexpr int NLD = DIM * NP;
    __shared__ double sG[EPB][NQ][NP][DIM];
    __shared__ double sD[EPB][NQ][NSYM];
    __shared__ double sDB[EPB][NP][NS][DIM];

    const int lane = threadIdx.x;
    const int le = lane / NPAIR;          // element slot of this lane
    const int pr = lane - le * NPAIR;     // pair index
    // pair -> (a,b), a <= b, row-major over the upper triangle
    int pa = 0, pb = 0;
    {
        int k = pr, a = 0;
        while (k >= NP - a) { k -= NP - a; ++a; }
        pa = a; pb = a + k;
    }
    const bool active = le < EPB;

    for (int64_t e0 = (int64_t)blockIdx.x * EPB; e0 < n_e; e0 += (int64_t)gridDim.x * EPB) {
        // stage gradients and weighted tangents
        for (int t = lane; t < EPB * NQ * NP; t += 64) {
            const int s = t / (NQ * NP), r = t - s * NQ * NP;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int64_t g = e * NQ * NP + r;  // (point, node) flattened
                const int q = r / NP, a = r - q * NP;
                sG[s][q][a][0] = D1[g];
                sG[s][q][a][1] = D2[g];
                if (DIM == 3) sG[s][q][a][DIM - 1] = D3[g];
            }
        }
        for (int t = lane; t < EPB * NQ * NSYM; t += 64) {
            const int s = t / (NQ * NSYM), r = t - s * NQ * NSYM;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int q = r / NSYM, k = r - q * NSYM;
                // packed index k -> (i,j), i <= j
                int i = 0, kk = k;
                while (kk >= NS - i) { kk -= NS - i; ++i; }
                const int j = i + kk;
                const int64_t g = e * NQ + q;
                const double w = W[g];
                const double d = DSL == 0 ? DS[g * NS * NS + i + NS * j] : DS[g * NSYM + k];
                sD[s][q][k] = w * d;
            }
        }
        __syncthreads();

        double acc[DIM][DIM];
#pragma unroll
        for (int i = 0; i < DIM; ++i)
#pragma unroll
            for (int j = 0; j < DIM; ++j) acc[i][j] = 0.0;

        for (int q = 0; q < NQ; ++q) {
            // DB_b = (wD) B_b for every node b of every element slot
            for (int t = lane; t < EPB * NP * NS * DIM; t += 64) {
                const int s = t / (NP * NS * DIM);
                int r = t - s * NP * NS * DIM;
                const int b = r / (NS * DIM);
                r -= b * NS * DIM;
                const int row = r / DIM, c = r - row * DIM;
                double v = 0.0;
#pragma unroll
                for (int k = 0; k < V::NNZ; ++k)
                    v += sD[s][q][symidx(row, V::row(c, k), NS)] * sG[s][q][b][V::grad(c, k)];
                sDB[s][b][row][c] = v;
            }
            __syncthreads();
            if (active) {
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k)
                            v += sG[le][q][pa][V::grad(i, k)] * sDB[le][pb][V::row(i, k)][j];
                        acc[i][j] += v;
                    }
            }
            __syncthreads();
        }

        const int64_t e = e0 + le;
        if (active && e < n_e) {
            if (SCAT == 2) {
                double *o = out + ((int64_t)e * NPAIR + pr) * DIM * DIM;
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) o[i * DIM + j] = acc[i][j];
            } else if (SCAT == 0) {
                const int32_t bab = emap[(e * NP + pa) * NP + pb];
                double *o = out + (int64_t)bab * DIM * DIM;
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) atomic_add_f64(o + i * DIM + j, acc[i][j]);
                if (pa != pb) {
                    const int32_t bba = emap[(e * NP + pb) * NP + pa];
                    double *o2 = out + (int64_t)bba * DIM * DIM;
#pragma unroll
                    for (int i = 0; i < DIM; ++i)
#pragma unroll
                        for (int j = 0; j < DIM; ++j) atomic_add_f64(o2 + j * DIM + i, acc[i][j]);
                }
            } else {
                const int64_t *m = smap + e * (int64_t)NLD * NLD;
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) {
                        const int ra = DIM * pa + i, cb = DIM * pb + j;
                        const int64_t k1 = m[ra + NLD * cb];
                        if (k1 > 0) atomic_add_f64(out + k1 - 1, acc[i][j]);
                        if (pa != pb) {
                            const int64_t k2 = m[cb + NLD * ra];
                            if (k2 > 0) atomic_add_f64(out + k2 - 1, acc[i][j]);
                        }
                    }
            }
        }
        __syncthreads();
    }
}

// Element pair blocks without per-point barriers: the element's gradients and
// w*D of all its points are staged in LDS once (4.5 kB for a P2
// tetrahedron), then every lane forms its own D*B_b (54 FMA per point in 3D)
// and B_a'(D B_b) (27 FMA) straight from LDS; the w*D reads are the same
// address for every lane of the element (broadcast).  The k_assemble form
// shares D*B_b between the lanes of one node through LDS instead, at the cost
// of two barriers and its LDS round trip per point.  The products are the
// same expressions in the same order.
// With SLOT, each pair block goes to its slot in the upper block's
// contribution list (pslot), transposed when its first node has the larger
// global id, so that the slot holds the block as the upper block takes it.
template <int DIM, int NP, int NQ, int DSL, bool SLOT = false>
__global__ __launch_bounds__(64) void k_assemble_pairs(int64_t n_e, const double *__restrict__ D1,
                                                       const double *__restrict__ D2, const double *__restrict__ D3,
                                                       const double *__restrict__ DS, const double *__restrict__ W,
                                                       double *__restrict__ out, const int32_t *__restrict__ elem,
                                                       const int32_t *__restrict__ pslot)
{
    using V = Voigt<DIM>;
    constexpr int NS = V::NS, NSYM = NS * (NS + 1) / 2, NPAIR = NP * (NP + 1) / 2;
    constexpr int EPB = 64 / NPAIR;
    __shared__ double sG[EPB][NQ][NP][DIM];
    __shared__ double sD[EPB][NQ][NSYM];
    const int lane = threadIdx.x;
    const int le = lane / NPAIR;
    const int pr = lane - le * NPAIR;
    int pa = 0, pb = 0;
    {
        int k = pr, a = 0;
        while (k >= NP - a) { k -= NP - a; ++a; }
        pa = a; pb = a + k;
    }
    const bool active = le < EPB;
    for (int64_t e0 = (int64_t)blockIdx.x * EPB; e0 < n_e; e0 += (int64_t)gridDim.x * EPB) {
        for (int t = lane; t < EPB * NQ * NP; t += 64) {
            const int s = t / (NQ * NP), r = t - s * NQ * NP;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int64_t g = e * NQ * NP + r;
                const int q = r / NP, a = r - q * NP;
                sG[s][q][a][0] = D1[g];
                sG[s][q][a][1] = D2[g];
                if (DIM == 3) sG[s][q][a][DIM - 1] = D3[g];
            }
        }
        for (int t = lane; t < EPB * NQ * NSYM; t += 64) {
            const int s = t / (NQ * NSYM), r = t - s * NQ * NSYM;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int q = r / NSYM, k = r - q * NSYM;
                int i = 0, kk = k;
                while (kk >= NS - i) { kk -= NS - i; ++i; }
                const int j = i + kk;
                const int64_t g = e * NQ + q;
                const double w = W[g];
                const double d = DSL == 0 ? DS[g * NS * NS + i + NS * j] : DS[g * NSYM + k];
                sD[s][q][k] = w * d;
            }
        }
        __syncthreads();
        const int64_t e = e0 + le;
        if (active && e < n_e) {
            double acc[DIM][DIM];
#pragma unroll
            for (int i = 0; i < DIM; ++i)
#pragma unroll
                for (int j = 0; j < DIM; ++j) acc[i][j] = 0.0;
            for (int q = 0; q < NQ; ++q) {
                double ga[DIM], gb[DIM];
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    ga[c] = sG[le][q][pa][c];
                    gb[c] = sG[le][q][pb][c];
                }
                // t = (wD) B_b, NS x DIM
                double t[NS][DIM];
#pragma unroll
                for (int row = 0; row < NS; ++row)
#pragma unroll
                    for (int c = 0; c < DIM; ++c) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) v += sD[le][q][symidx(row, V::row(c, k), NS)] * gb[V::grad(c, k)];
                        t[row][c] = v;
                    }
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) v += ga[V::grad(i, k)] * t[V::row(i, k)][j];
                        acc[i][j] += v;
                    }
            }
            if (SLOT) {
                double *o = out + (int64_t)pslot[(int64_t)e * NPAIR + pr] * DIM * DIM;
                if (elem[e * NP + pa] > elem[e * NP + pb]) {
#pragma unroll
                    for (int i = 0; i < DIM; ++i)
#pragma unroll
                        for (int j = 0; j < DIM; ++j) o[j * DIM + i] = acc[i][j];
                } else {
#pragma unroll
                    for (int i = 0; i < DIM; ++i)
#pragma unroll
                        for (int j = 0; j < DIM; ++j) o[i * DIM + j] = acc[i][j];
                }
            } else {
                double *o = out + ((int64_t)e * NPAIR + pr) * DIM * DIM;
#pragma unroll
                for (int i = 0; i < DIM; ++i)
#pragma unroll
                    for (int j = 0; j < DIM; ++j) o[i * DIM + j] = acc[i][j];
            }
        }
        __syncthreads();
    }
}

// 3D pair blocks, two pairs of one column node per lane: the lane of node b
// forms t = (wD) B_b once per point and uses it for the pairs (a1, b) and
// (a2, b), a1 < a2 <= b, so the 54 FMA of D*B_b and the 21 broadcast LDS
// reads of w*D serve two blocks (30 lanes per element, ceil((b+1)/2) for
// node b; a column of odd length leaves one lane with a single pair, which
// repeats its first pair and stores it once).  Two elements per wavefront,
// one per 32-lane half.  The block arithmetic is k_assemble_pairs' term for
// term, so the two kernels give the same bits.
template <int DSL, bool SLOT>
__global__ __launch_bounds__(64) void k_assemble_pairs2(int64_t n_e, const double *__restrict__ D1,
                                                        const double *__restrict__ D2,
                                                        const double *__restrict__ D3,
                                                        const double *__restrict__ DS,
                                                        const double *__restrict__ W, double *__restrict__ out,
                                                        const int32_t *__restrict__ elem,
                                                        const int32_t *__restrict__ pslot)
{
    using V = Voigt<3>;
    constexpr int NP = 10, NQ = 11, NS = 6, NSYM = 21, NPAIR = 55, LPE = 30;
    __shared__ double sG[2][NQ][NP][3];
    __shared__ double sD[2][NQ][NSYM];
    const int lane = threadIdx.x, le = lane >> 5, li = lane & 31;
    int b = 0, a1 = 0, a2 = -1;
    {
        int k = li < LPE ? li : LPE - 1;
        for (int bb = 0; bb < NP; ++bb) {
            const int ng = (bb + 2) / 2;
            if (k < ng) {
                b = bb;
                a1 = 2 * k;
                a2 = 2 * k + 1 <= bb ? 2 * k + 1 : -1;
                break;
            }
            k -= ng;
        }
    }
    const int a2c = a2 >= 0 ? a2 : a1;
    const int pr1 = a1 * NP - (a1 * (a1 - 1)) / 2 + (b - a1);
    const int pr2 = a2c * NP - (a2c * (a2c - 1)) / 2 + (b - a2c);
    for (int64_t e0 = (int64_t)blockIdx.x * 2; e0 < n_e; e0 += (int64_t)gridDim.x * 2) {
        for (int t = lane; t < 2 * NQ * NP; t += 64) {
            const int s = t / (NQ * NP), r = t - s * NQ * NP;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int64_t g = e * NQ * NP + r;
                const int q = r / NP, a = r - q * NP;
                sG[s][q][a][0] = D1[g];
                sG[s][q][a][1] = D2[g];
                sG[s][q][a][2] = D3[g];
            }
        }
        for (int t = lane; t < 2 * NQ * NSYM; t += 64) {
            const int s = t / (NQ * NSYM), r = t - s * NQ * NSYM;
            const int64_t e = e0 + s;
            if (e < n_e) {
                const int q = r / NSYM, k = r - q * NSYM;
                int i = 0, kk = k;
                while (kk >= NS - i) { kk -= NS - i; ++i; }
                const int j = i + kk;
                const int64_t g = e * NQ + q;
                const double w = W[g];
                const double d = DSL == 0 ? DS[g * NS * NS + i + NS * j] : DS[g * NSYM + k];
                sD[s][q][k] = w * d;
            }
        }
        __syncthreads();
        const int64_t e = e0 + le;
        if (li < LPE && e < n_e) {
            double acc1[3][3], acc2[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc1[i][j] = acc2[i][j] = 0.0;
            for (int q = 0; q < NQ; ++q) {
                double g1[3], g2[3], gb[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    g1[c] = sG[le][q][a1][c];
                    g2[c] = sG[le][q][a2c][c];
                    gb[c] = sG[le][q][b][c];
                }
                double t[NS][3];
#pragma unroll
                for (int row = 0; row < NS; ++row)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) v += sD[le][q][symidx(row, V::row(c, k), NS)] * gb[V::grad(c, k)];
                        t[row][c] = v;
                    }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double v1 = 0.0, v2 = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) {
                            v1 += g1[V::grad(i, k)] * t[V::row(i, k)][j];
                            v2 += g2[V::grad(i, k)] * t[V::row(i, k)][j];
                        }
                        acc1[i][j] += v1;
                        acc2[i][j] += v2;
                    }
            }
            const int32_t nb = elem[e * NP + b];
            if (SLOT) {
                double *o = out + (int64_t)pslot[e * NPAIR + pr1] * 9;
                if (elem[e * NP + a1] > nb) {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) o[j * 3 + i] = acc1[i][j];
                } else {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) o[i * 3 + j] = acc1[i][j];
                }
                if (a2 >= 0) {
                    double *o2 = out + (int64_t)pslot[e * NPAIR + pr2] * 9;
                    if (elem[e * NP + a2] > nb) {
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int j = 0; j < 3; ++j) o2[j * 3 + i] = acc2[i][j];
                    } else {
#pragma unroll
                        for (int i = 0; i < 3; ++i)
#pragma unroll
                            for (int j = 0; j < 3; ++j) o2[i * 3 + j] = acc2[i][j];
                    }
                }
            } else {
                double *o = out + (e * NPAIR + pr1) * 9;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) o[i * 3 + j] = acc1[i][j];
                if (a2 >= 0) {
                    double *o2 = out + (e * NPAIR + pr2) * 9;
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) o2[i * 3 + j] = acc2[i][j];
                }
            }
        }
        __syncthreads();
    }
}

// Loads of one element pair of k_assemble_pairs2p into registers.  Chunk
// indices past the end of the pair (the last pair may hold one element) are
// clamped to its last chunk, so every lane loads and stores without a branch:
// duplicates load and store the same values.  (Branches here make the
// compiler wait on each load in turn.)
// A native two-double vector: its loads, stores and copies stay in registers.
typedef double dvec2 __attribute__((ext_vector_type(2)));

template <int NGI, int NDI, bool SLOT>
__device__ __forceinline__ void pairs2p_prefetch(int64_t n_e, int64_t e0, int lane, const double *__restrict__ D1,
                                                 const double *__restrict__ D2, const double *__restrict__ D3,
                                                 const double *__restrict__ DS, const double *__restrict__ W,
                                                 const int32_t *__restrict__ elem, const int32_t *__restrict__ pslot,
                                                 int b, int a1, int a2c, int pr1, int pr2, dvec2 (&pg0)[NGI],
                                                 dvec2 (&pg1)[NGI], dvec2 (&pg2)[NGI], dvec2 (&pd)[NDI],
                                                 double (&pw)[NDI], int32_t (&pm)[5])
{
    constexpr int NP = 10, NQ = 11, GE = 110, DE = 396, NPAIR = 55;
    const int ne = n_e - e0 >= 2 ? 2 : 1;
    // the lane's element (its half of the pair, or the first when the pair
    // holds one) and where its two blocks go
    const int64_t ec = e0 + ((lane >> 5) < ne ? (lane >> 5) : 0);
    if (SLOT) {
        pm[0] = elem[ec * NP + b];
        pm[1] = elem[ec * NP + a1];
        pm[2] = elem[ec * NP + a2c];
        pm[3] = pslot[ec * NPAIR + pr1];
        pm[4] = pslot[ec * NPAIR + pr2];
    }
    const int gmax = ne * (GE / 2) - 1, dmax = ne * (DE / 2) - 1;
#pragma unroll
    for (int m = 0; m < NGI; ++m) {
        const int ch = min(lane + 64 * m, gmax);
        pg0[m] = *reinterpret_cast<const dvec2 *>(D1 + e0 * GE + 2 * ch);
        pg1[m] = *reinterpret_cast<const dvec2 *>(D2 + e0 * GE + 2 * ch);
        pg2[m] = *reinterpret_cast<const dvec2 *>(D3 + e0 * GE + 2 * ch);
    }
#pragma unroll
    for (int m = 0; m < NDI; ++m) {
        const int ch = min(lane + 64 * m, dmax);
        pd[m] = *reinterpret_cast<const dvec2 *>(DS + e0 * DE + 2 * ch);
        const int o = 2 * ch, se = o / DE, q = (o - se * DE) / 36;
        pw[m] = W[(e0 + se) * NQ + q];
    }
}

// Software-pipelined k_assemble_pairs2 (MEX-layout DS, 16-byte aligned
// inputs): a persistent grid, and while a wavefront forms the pair blocks of
// one element pair from LDS, the gradients, tangents and weights of its next
// element pair are in flight into registers (16-byte loads), moved to LDS
// once the current blocks are stored.  The memory latency of each element pair
// is hidden behind the previous pair's compute instead of exposed before its
// own.  The staged values and the block arithmetic are k_assemble_pairs2's,
// so the results are the same bits.
template <bool SLOT>
__global__ __launch_bounds__(64) void k_assemble_pairs2p(int64_t n_e, const double *__restrict__ D1,
                                                         const double *__restrict__ D2,
                                                         const double *__restrict__ D3,
                                                         const double *__restrict__ DS,
                                                         const double *__restrict__ W, double *__restrict__ out,
                                                         const int32_t *__restrict__ elem,
                                                         const int32_t *__restrict__ pslot)
{
    using V = Voigt<3>;
    constexpr int NP = 10, NQ = 11, NS = 6, NSYM = 21, NPAIR = 55, LPE = 30;
    constexpr int GE = NQ * NP;           // gradients per element and component
    constexpr int DE = NQ * NS * NS;      // raw tangent entries per element
    constexpr int NGI = (GE + 63) / 64;   // 16-byte gradient chunks per lane and component (2 elements)
    constexpr int NDI = (DE + 63) / 64;   // 16-byte tangent chunks per lane (2 elements)
    __shared__ __align__(16) double sG[3][2 * GE];
    __shared__ double sD[2][NQ][NSYM + 1];
    __shared__ double sO[SLOT ? 2 : 1][NPAIR][9];    // the pair's blocks on their way to the slots
    __shared__ int64_t sSlot[SLOT ? 2 : 1][NPAIR];
    const int lane = threadIdx.x, le = lane >> 5, li = lane & 31;
    int b = 0, a1 = 0, a2 = -1;
    {
        int k = li < LPE ? li : LPE - 1;
        for (int bb = 0; bb < NP; ++bb) {
            const int ng = (bb + 2) / 2;
            if (k < ng) {
                b = bb;
                a1 = 2 * k;
                a2 = 2 * k + 1 <= bb ? 2 * k + 1 : -1;
                break;
            }
            k -= ng;
        }
    }
    const int a2c = a2 >= 0 ? a2 : a1;
    const int pr1 = a1 * NP - (a1 * (a1 - 1)) / 2 + (b - a1);
    const int pr2 = a2c * NP - (a2c * (a2c - 1)) / 2 + (b - a2c);
    dvec2 pg0[NGI], pg1[NGI], pg2[NGI], pd[NDI];
    double pw[NDI];
    int32_t pm[5] = {0, 0, 0, 0, 0};
    const int64_t stride = (int64_t)gridDim.x * 2;
    int64_t e0 = (int64_t)blockIdx.x * 2;
    if (e0 < n_e)
        pairs2p_prefetch<NGI, NDI, SLOT>(n_e, e0, lane, D1, D2, D3, DS, W, elem, pslot, b, a1, a2c, pr1, pr2, pg0, pg1,
                                         pg2, pd, pw, pm);
    for (; e0 < n_e; e0 += stride) {
        {
            const int ne = n_e - e0 >= 2 ? 2 : 1;
            const int gmax = ne * (GE / 2) - 1, dmax = ne * (DE / 2) - 1;
#pragma unroll
            for (int m = 0; m < NGI; ++m) {
                const int ch = min(lane + 64 * m, gmax);
                *reinterpret_cast<dvec2 *>(&sG[0][2 * ch]) = pg0[m];
                *reinterpret_cast<dvec2 *>(&sG[1][2 * ch]) = pg1[m];
                *reinterpret_cast<dvec2 *>(&sG[2][2 * ch]) = pg2[m];
            }
#pragma unroll
            for (int m = 0; m < NDI; ++m) {
                const int ch = min(lane + 64 * m, dmax);
                const int o = 2 * ch, se = o / DE, r = o - se * DE, q = r / (NS * NS), idx = r - q * NS * NS;
                const int i = idx % NS, j = idx / NS;  // idx even: entries (i, j), (i + 1, j)
                // entries below the diagonal go to a spare slot (index NSYM) instead of a branch
                const int k0 = i <= j ? i * NS - (i * (i - 1)) / 2 + (j - i) : NSYM;
                const int k1 = i + 1 <= j ? (i + 1) * NS - ((i + 1) * i) / 2 + (j - i - 1) : NSYM;
                sD[se][q][k0] = pw[m] * pd[m].x;
                sD[se][q][k1] = pw[m] * pd[m].y;
            }
        }
        // the store targets of this pair came with its prefetch: loaded before
        // the next pair's prefetch, so (the load counter being in order)
        // waiting for them at the stores does not wait for that prefetch
        const int64_t e = e0 + le;
        const bool act = li < LPE && e < n_e;
        const bool tr1 = pm[1] > pm[0], tr2 = pm[2] > pm[0];
        const int64_t s1 = SLOT ? (int64_t)pm[3] : e * NPAIR + pr1;
        const int64_t s2 = SLOT ? (int64_t)pm[4] : e * NPAIR + pr2;
        __syncthreads();
        // unconditional (the last iteration reloads its own pair), so that no
        // branch makes the stores below wait for these loads
        pairs2p_prefetch<NGI, NDI, SLOT>(n_e, e0 + stride < n_e ? e0 + stride : e0, lane, D1, D2, D3, DS, W, elem, pslot,
                                         b, a1, a2c, pr1, pr2, pg0, pg1, pg2, pd, pw, pm);
        if (act) {
            double acc1[3][3], acc2[3][3];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) acc1[i][j] = acc2[i][j] = 0.0;
            for (int q = 0; q < NQ; ++q) {
                double g1[3], g2[3], gb[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    g1[c] = sG[c][le * GE + q * NP + a1];
                    g2[c] = sG[c][le * GE + q * NP + a2c];
                    gb[c] = sG[c][le * GE + q * NP + b];
                }
                double t[NS][3];
#pragma unroll
                for (int row = 0; row < NS; ++row)
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        double v = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) v += sD[le][q][symidx(row, V::row(c, k), NS)] * gb[V::grad(c, k)];
                        t[row][c] = v;
                    }
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        double v1 = 0.0, v2 = 0.0;
#pragma unroll
                        for (int k = 0; k < V::NNZ; ++k) {
                            v1 += g1[V::grad(i, k)] * t[V::row(i, k)][j];
                            v2 += g2[V::grad(i, k)] * t[V::row(i, k)][j];
                        }
                        acc1[i][j] += v1;
                        acc2[i][j] += v2;
                    }
            }
            if (SLOT) {
                // through LDS, so that the stores below write each slot with
                // consecutive lanes
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) sO[le][pr1][tr1 ? j * 3 + i : i * 3 + j] = acc1[i][j];
                sSlot[le][pr1] = s1;
                if (a2 >= 0) {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) sO[le][pr2][tr2 ? j * 3 + i : i * 3 + j] = acc2[i][j];
                    sSlot[le][pr2] = s2;
                }
            } else {
                double *o1 = out + s1 * 9;
                double *o2 = out + s2 * 9;
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) o1[i * 3 + j] = acc1[i][j];
                if (a2 >= 0) {
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) o2[i * 3 + j] = acc2[i][j];
                }
            }
        }
        if (SLOT) {
            // Each 72-byte slot as four 16-byte pieces and one 8-byte piece,
            // five consecutive lanes per slot (a slot at an even index starts
            // 16-byte aligned and ends with the 8-byte piece, one at an odd
            // index starts with it): about two memory requests per slot
            // instead of nine scattered 8-byte stores.
            __syncthreads();
            const int ne = n_e - e0 >= 2 ? 2 : 1;
            for (int p = lane; p < ne * NPAIR * 5; p += 64) {
                const int sl = p / 5, piece = p - sl * 5;
                const int se = sl / NPAIR, pr = sl - se * NPAIR;
                const int64_t dst = sSlot[se][pr];
                const double *v = sO[se][pr];
                double *o = out + dst * 9;
                const int first = (dst & 1) == 0 ? 2 * piece : (piece == 0 ? 0 : 2 * piece - 1);
                const bool wide = (dst & 1) == 0 ? piece < 4 : piece > 0;
                if (wide) {
                    dvec2 w2;
                    w2.x = v[first];
                    w2.y = v[first + 1];
                    *reinterpret_cast<dvec2 *>(o + first) = w2;
                } else {
                    o[first] = v[first];
                }
            }
        }
        __syncthreads();
    }
}

// Upper blocks from their slot lists (contiguous, in element order): entry q
// of upper block b is the plain sum of entry q of its slots, the same terms in
// the same order as the per-block gather of k_gather_blocks, so the values
// are the same bits.  REG = false writes K_t into b and, transposed, its
// mirror tb; REG = true writes the regularised, symmetrised, masked K_r as
// k_gather_regularize does.
template <int DIM, bool REG>
__global__ __launch_bounds__(256) void k_gather_slots(int64_t n_ub, const int32_t *__restrict__ ublk,
                                                      const int32_t *__restrict__ uptr,
                                                      const double *__restrict__ slots,
                                                      const int32_t *__restrict__ brow,
                                                      const int32_t *__restrict__ col,
                                                      const int32_t *__restrict__ tblk,
                                                      const double *__restrict__ mask, double r,
                                                      const double *__restrict__ Kel, double *__restrict__ out)
{
    constexpr int DD = DIM * DIM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t u = t / DD;
    if (u >= n_ub) return;
    const int q = (int)(t - u * DD), i = q / DIM, j = q - i * DIM, qt = j * DIM + i;
    const int64_t b = ublk[u];
    const int64_t tb = tblk[b];
    const int32_t c0 = uptr[u], c1 = uptr[u + 1];
    // the operands of the epilogue are loaded before the slot stream
    double kb = 0.0, ktb = 0.0, mro = 1.0, mco = 1.0;
    int64_t ro = 0, co = 0;
    if (REG) {
        kb = Kel[b * DD + q];
        ktb = Kel[tb * DD + qt];
        ro = (int64_t)brow[b] * DIM + i;
        co = (int64_t)col[b] * DIM + j;
        mro = mask[ro];
        mco = mask[co];
    }
    // four slot loads in flight per step, added in list order (the same sum)
    const double *sp = slots + (int64_t)c0 * DD;
    const int32_t n = c1 - c0;
    double s = 0.0;
    int32_t c = 0;
    for (; c + 4 <= n; c += 4) {
        const double x0 = sp[(int64_t)c * DD + q], x1 = sp[(int64_t)(c + 1) * DD + q];
        const double x2 = sp[(int64_t)(c + 2) * DD + q], x3 = sp[(int64_t)(c + 3) * DD + q];
        s += x0;
        s += x1;
        s += x2;
        s += x3;
    }
    for (; c < n; ++c) s += sp[(int64_t)c * DD + q];
    double st = s;
    if (tb == b) {
        st = 0.0;
        for (c = 0; c + 4 <= n; c += 4) {
            const double x0 = sp[(int64_t)c * DD + qt], x1 = sp[(int64_t)(c + 1) * DD + qt];
            const double x2 = sp[(int64_t)(c + 2) * DD + qt], x3 = sp[(int64_t)(c + 3) * DD + qt];
            st += x0;
            st += x1;
            st += x2;
            st += x3;
        }
        for (; c < n; ++c) st += sp[(int64_t)c * DD + qt];
    }
    if (!REG) {
        out[b * DD + q] = s;
        if (tb != b) out[tb * DD + qt] = s;
        return;
    }
    double v;
    {
#pragma clang fp contract(off)
        const double v1 = r * kb + (1.0 - r) * s;
        const double v2 = r * ktb + (1.0 - r) * st;
        v = 0.5 * (v1 + v2);
    }
    if (mro == 0.0 || mco == 0.0) v = (ro == co) ? 1.0 : 0.0;
    out[b * DD + q] = v;
    if (tb != b) out[tb * DD + qt] = v;
}

// K_r = r K_el + (1-r) sum_e Ke, symmetrised and Dirichlet-masked, in one pass
// over the blocks on and above the block diagonal (CONSTITUTIVE.m:515-532):
// each such block b = (v, w) sums its element contributions once and writes
// both b and, transposed, its mirror tb = (w, v).  The mirror's sum is the
// same sum bit for bit (the contribution lists of b and tb hold the same
// element pair blocks in the same element order, b's as is where tb's are
// transposed), so the result equals the two-pass path (gather, then
// k_regularize) exactly, and neither K_t nor the transposed reads of it exist.
// Diagonal blocks (tb == b) also form the transposed entry's sum.
template <int DIM>
__global__ __launch_bounds__(256) void k_gather_regularize(int64_t n_ub, const int32_t *__restrict__ ublk,
                                                           const int32_t *__restrict__ brow,
                                                           const int32_t *__restrict__ col,
                                                           const int32_t *__restrict__ tblk,
                                                           const int32_t *__restrict__ cptr,
                                                           const int32_t *__restrict__ code,
                                                           const double *__restrict__ ke,
                                                           const double *__restrict__ mask, double r,
                                                           const double *__restrict__ Kel, double *__restrict__ Kr)
{
    constexpr int DD = DIM * DIM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t u = t / DD;
    if (u >= n_ub) return;
    const int q = (int)(t - u * DD), i = q / DIM, j = q - i * DIM, qt = j * DIM + i;
    const int64_t b = ublk[u];
    const int64_t tb = tblk[b];
    const int32_t c0 = cptr[b], c1 = cptr[b + 1];
    double s = 0.0;
    for (int32_t c = c0; c < c1; ++c) {
        const int32_t cd = code[c];
        s += ke[(int64_t)(cd >> 1) * DD + ((cd & 1) ? qt : q)];
    }
    double st = s;
    if (tb == b) {
        st = 0.0;
        for (int32_t c = c0; c < c1; ++c) {
            const int32_t cd = code[c];
            st += ke[(int64_t)(cd >> 1) * DD + ((cd & 1) ? q : qt)];
        }
    }
    double v;
    {
#pragma clang fp contract(off)
        const double v1 = r * Kel[b * DD + q] + (1.0 - r) * s;
        const double v2 = r * Kel[tb * DD + qt] + (1.0 - r) * st;
        v = 0.5 * (v1 + v2);
    }
    const int64_t ro = (int64_t)brow[b] * DIM + i, co = (int64_t)col[b] * DIM + j;
    if (mask[ro] == 0.0 || mask[co] == 0.0) v = (ro == co) ? 1.0 : 0.0;
    Kr[b * DD + q] = v;
    if (tb != b) Kr[tb * DD + qt] = v;
}

// ---------------------------------------------------------------------------
// Strain gather: one thread per integration point.  U is dim x n_n
// (node-major), E is point-major (ns per point).  GEO: the point's gradients
// are computed from the node coordinates (p2geom.h, the arithmetic that
// produced the stored DPhi on the device) instead of read from D1..D3.
// ---------------------------------------------------------------------------
template <int DIM, int NP, int NQ, bool GEO = false>
__global__ __launch_bounds__(256) void k_strain(int64_t n_e, const int32_t *__restrict__ elem,
                                                const double *__restrict__ D1,
                                                const double *__restrict__ D2,
                                                const double *__restrict__ D3,
                                                const double *__restrict__ U, double *__restrict__ E,
                                                const double *__restrict__ coord = nullptr)
{
    using V = Voigt<DIM>;
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= n_e * NQ) return;
    const int64_t e = g / NQ;
    double G[DIM][NP];
    if (GEO) p2geom::point_gradients<DIM, NP>(coord, elem + e * NP, (int)(g - e * NQ), G);
    double eps[V::NS];
#pragma unroll
    for (int r = 0; r < V::NS; ++r) eps[r] = 0.0;
#pragma unroll
    for (int a = 0; a < NP; ++a) {
        const int64_t node = elem[e * NP + a];
        double gr[DIM];
        if (GEO) {
#pragma unroll
            for (int c = 0; c < DIM; ++c) gr[c] = G[c][a];
        } else {
            gr[0] = D1[g * NP + a];
            gr[1] = D2[g * NP + a];
            if (DIM == 3) gr[DIM - 1] = D3[g * NP + a];
        }
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            const double u = U[node * DIM + c];
#pragma unroll
            for (int k = 0; k < V::NNZ; ++k) eps[V::row(c, k)] += gr[V::grad(c, k)] * u;
        }
    }
#pragma unroll
    for (int r = 0; r < V::NS; ++r) E[g * V::NS + r] = eps[r];
}

// Internal force: one thread per (element, node); S has s_rows rows per
// point (6 in 3D; 4 in 2D where sigma33 is not used).  GATHER: the thread
// writes its node force to out[(e*NP + a)*DIM + c] for k_gather_force;
// otherwise it adds it to F (dim x n_n, zeroed first) with atomics.
template <int DIM, int NP, int NQ, bool GATHER>
__global__ __launch_bounds__(256) void k_force(int64_t n_e, const int32_t *__restrict__ elem,
                                               const double *__restrict__ D1,
                                               const double *__restrict__ D2,
                                               const double *__restrict__ D3,
                                               const double *__restrict__ S, int s_rows,
                                               const double *__restrict__ W, double *__restrict__ F)
{
    // one thread per (element, node): the NP threads of an element read its
    // point gradients as contiguous spans and share the stress rows
    using V = Voigt<DIM>;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t e = t / NP;
    const int a = (int)(t - e * NP);
    if (e >= n_e) return;
    double f[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) f[c] = 0.0;
    for (int q = 0; q < NQ; ++q) {
        const int64_t g = e * NQ + q;
        double ws[V::NS];
#pragma unroll
        for (int r = 0; r < V::NS; ++r) ws[r] = W[g] * S[g * s_rows + r];
        double gr[DIM];
        gr[0] = D1[g * NP + a];
        gr[1] = D2[g * NP + a];
        if (DIM == 3) gr[DIM - 1] = D3[g * NP + a];
#pragma unroll
        for (int c = 0; c < DIM; ++c) {
            double v = 0.0;
#pragma unroll
            for (int k = 0; k < V::NNZ; ++k) v += gr[V::grad(c, k)] * ws[V::row(c, k)];
            f[c] += v;
        }
    }
    if (GATHER) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) F[t * DIM + c] = f[c];
    } else {
        const int64_t node = elem[e * NP + a];
#pragma unroll
        for (int c = 0; c < DIM; ++c) atomic_add_f64(F + node * DIM + c, f[c]);
    }
}

// k_force with the gradients computed from the node coordinates: EB elements
// per block.  Each lane of the first phase takes one point, computes its
// gradients (p2geom.h) and the products v(a, c) of every node a, which the
// second phase, one lane per (element, node), sums over the points in order.
// The products and the sums are k_force's, term for term, so the results are
// the same bits as k_force on the stored gradients.
template <int DIM, int NP, int NQ, int EB, bool GATHER>
__global__ __launch_bounds__(256) void k_force_geo(int64_t n_e, const int32_t *__restrict__ elem,
                                                   const double *__restrict__ coord,
                                                   const double *__restrict__ S, int s_rows,
                                                   const double *__restrict__ W, double *__restrict__ F)
{
    using V = Voigt<DIM>;
    __shared__ double sV[EB][NQ][NP][DIM];
    const int t = threadIdx.x;
    const int64_t e0 = (int64_t)blockIdx.x * EB;
    if (t < EB * NQ) {
        const int el = t / NQ, q = t - (t / NQ) * NQ;
        const int64_t e = e0 + el;
        if (e < n_e) {
            const int64_t g = e * NQ + q;
            double G[DIM][NP];
            p2geom::point_gradients<DIM, NP>(coord, elem + e * NP, q, G);
            double ws[V::NS];
#pragma unroll
            for (int r = 0; r < V::NS; ++r) ws[r] = W[g] * S[g * s_rows + r];
#pragma unroll
            for (int a = 0; a < NP; ++a) {
                double gr[DIM];
#pragma unroll
                for (int c = 0; c < DIM; ++c) gr[c] = G[c][a];
#pragma unroll
                for (int c = 0; c < DIM; ++c) {
                    double v = 0.0;
#pragma unroll
                    for (int k = 0; k < V::NNZ; ++k) v += gr[V::grad(c, k)] * ws[V::row(c, k)];
                    sV[el][q][a][c] = v;
                }
            }
        }
    }
    __syncthreads();
    if (t >= EB * NP) return;
    const int el = t / NP, a = t - (t / NP) * NP;
    const int64_t e = e0 + el;
    if (e >= n_e) return;
    double f[DIM];
#pragma unroll
    for (int c = 0; c < DIM; ++c) f[c] = 0.0;
    for (int q = 0; q < NQ; ++q)
#pragma unroll
        for (int c = 0; c < DIM; ++c) f[c] += sV[el][q][a][c];
    if (GATHER) {
#pragma unroll
        for (int c = 0; c < DIM; ++c) F[(e * NP + a) * DIM + c] = f[c];
    } else {
        const int64_t node = elem[e * NP + a];
#pragma unroll
        for (int c = 0; c < DIM; ++c) atomic_add_f64(F + node * DIM + c, f[c]);
    }
}

// F(c, v) = sum of the element-node forces fe of node v in element order
// (nptr/ncode from the pattern): the same sum, bit for bit, on every run.
// One thread per (node, component).
template <int DIM>
__global__ __launch_bounds__(256) void k_gather_force(int64_t n_n, const int32_t *__restrict__ nptr,
                                                      const int32_t *__restrict__ ncode,
                                                      const double *__restrict__ fe, double *__restrict__ F)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n_n * DIM) return;
    const int64_t v = t / DIM;
    const int c = (int)(t - v * DIM);
    double s = 0.0;
    for (int32_t k = nptr[v]; k < nptr[v + 1]; ++k) s += fe[(int64_t)ncode[k] * DIM + c];
    F[t] = s;
}

// K_r = r K_el + (1-r) K_t, symmetrised through the transpose block, then
// Dirichlet rows/columns masked (unit diagonal).  One thread per block entry.
template <int DIM>
__global__ void k_regularize(int64_t nnzb, const int32_t *__restrict__ brow,
                             const int32_t *__restrict__ col, const int32_t *__restrict__ tblk,
                             const double *__restrict__ mask, double r,
                             const double *__restrict__ Kel, const double *__restrict__ Kt,
                             double *__restrict__ Kr)
{
    constexpr int BS = DIM * DIM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nnzb * BS) return;
    const int64_t b = t / BS;
    const int ij = (int)(t - b * BS), i = ij / DIM, j = ij - (ij / DIM) * DIM;
    const int64_t bt = tblk[b];
    const int64_t tt = bt * BS + j * DIM + i;
    double v;
    {
        // no FMA contraction: the same rounding as the reference's
        // r*V_elast + (1-r)*V_tang (CONSTITUTIVE.m:515-532) and as k_gather_regularize
#pragma clang fp contract(off)
        const double v1 = r * Kel[t] + (1.0 - r) * Kt[t];
        const double v2 = r * Kel[tt] + (1.0 - r) * Kt[tt];
        v = 0.5 * (v1 + v2);
    }
    const int64_t ro = (int64_t)brow[b] * DIM + i, co = (int64_t)col[b] * DIM + j;
    if (mask[ro] == 0.0 || mask[co] == 0.0) v = (ro == co) ? 1.0 : 0.0;
    Kr[t] = v;
}

__global__ void k_gather_ref(int64_t nnz, const int64_t *__restrict__ pos, const double *__restrict__ src,
                             double *__restrict__ dst)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < nnz) dst[k] = src[pos[k]];
}

// Elastic DS per point: 2*DEV*shear + VOL*bulk with the operators of
// CONSTITUTIVE.m:413-422, MEX layout.  There IOTA = [1;1;1;zeros(n_strain-3,1)],
// so in 2D (n_strain = 3) VOL couples all three strain rows, shear included;
// that operator is what the reference regularises with, and it is kept.
__global__ void k_ds_elastic(int64_t n, int ns, const double *__restrict__ shear,
                             const double *__restrict__ bulk, double *__restrict__ DS)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * ns * ns) return;
    const int64_t p = t / (ns * ns);
    const int k = (int)(t - p * ns * ns), i = k % ns, j = k / ns;
    const int nn = ns == 6 ? 3 : 2;  // normal components of the Voigt identity
    const double vol = (i < 3 && j < 3) ? 1.0 : 0.0;
    const double idv = (i == j) ? (i < nn ? 1.0 : 0.5) : 0.0;
    const double dev = idv - vol / 3.0;
    DS[t] = 2.0 * dev * shear[p] + vol * bulk[p];
}

// Block values as the sum of their element blocks, in element order (no
// atomics, so the result is the same on every run): dim^2 lanes per block,
// lane q taking entry q of each contribution (entry q^T of a transposed one),
// so a contribution is read as one contiguous 8*dim^2-byte span.
template <int DIM>
__global__ __launch_bounds__(256) void k_gather_blocks(int64_t nnzb, const int32_t *__restrict__ cptr,
                                                       const int32_t *__restrict__ code,
                                                       const double *__restrict__ ke, double *__restrict__ out)
{
    constexpr int DD = DIM * DIM;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t b = t / DD;
    if (b >= nnzb) return;
    const int q = (int)(t - b * DD), i = q / DIM, j = q - i * DIM, qt = j * DIM + i;
    double s = 0.0;
    for (int32_t c = cptr[b]; c < cptr[b + 1]; ++c) {
        const int32_t cd = code[c];
        s += ke[(int64_t)(cd >> 1) * DD + ((cd & 1) ? qt : q)];
    }
    out[t] = s;
}

}  // namespace

// ---------------------------------------------------------------------------
// Entry points
// ---------------------------------------------------------------------------
static bool aligned16(const void *a, const void *b, const void *c, const void *d)
{
    return ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b) | reinterpret_cast<uintptr_t>(c) |
             reinterpret_cast<uintptr_t>(d)) & 15u) == 0;
}

// Persistent grid of the pipelined element kernel: two resident wavefronts
// per SIMD (its register use), each looping over element pairs.
static unsigned pipelined_grid(int64_t n_e)
{
    return (unsigned)std::min<int64_t>((n_e + 1) / 2, (int64_t)ssr_cu_count() * 8);
}

static int assemble_launch(int dim, int64_t n_p, int64_t n_e, int n_q, int dsl, int scat,
                           const double *D1, const double *D2, const double *D3, const double *DS,
                           const double *W, const int32_t *emap, const int64_t *smap, double *out)
{
    if (n_e == 0) return SSR_OK;
    const int64_t want = 4096;  // resident wavefronts; grid-stride beyond
    if (dim == 3) {
        SSR_REQUIRE(n_p == 10 && n_q == 11, "tangent assembly (3D): P2 tetrahedra with 11 points supported (got n_p=%lld, n_q=%d)",
                    (long long)n_p, n_q);
        const unsigned grid = (unsigned)std::min<int64_t>(n_e, want * 16);
        const unsigned grid2 = (unsigned)std::min<int64_t>((n_e + 1) / 2, INT32_MAX);  // two elements per group
#define L3(DSL_, SC_) hipLaunchKernelGGL((k_assemble<3, 10, 11, DSL_, SC_>), dim3(grid), dim3(64), 0, ssr_stream(), \
                                         n_e, D1, D2, D3, DS, W, emap, smap, out)
#define P3(DSL_) hipLaunchKernelGGL((k_assemble_pairs<3, 10, 11, DSL_>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, \
                                    D1, D2, D3, DS, W, out, nullptr, nullptr)
#define U3(DSL_) hipLaunchKernelGGL((k_assemble_pairs2<DSL_, false>), dim3(grid2), dim3(64), 0, ssr_stream(), \
                                    n_e, D1, D2, D3, DS, W, out, nullptr, nullptr)
        const int ek = ssr_tuning(SSR_TUNE_ELEMENT_KERNEL);
        if (scat == 2 && ek == 0 && dsl == 0 && aligned16(D1, D2, D3, DS)) {
            hipLaunchKernelGGL((k_assemble_pairs2p<false>), dim3(pipelined_grid(n_e)), dim3(64), 0, ssr_stream(), n_e,
                               D1, D2, D3, DS, W, out, nullptr, nullptr);
        } else if (scat == 2 && ek == 0) {
            if (dsl == 0) U3(0);
            else U3(1);
        } else if (scat == 2 && ek == 2) {
            if (dsl == 0) P3(0);
            else P3(1);
        } else if (scat == 2) {
            if (dsl == 0) L3(0, 2);
            else L3(1, 2);
        } else if (dsl == 0 && scat == 0) L3(0, 0);
        else if (dsl == 0) L3(0, 1);
        else if (scat == 0) L3(1, 0);
        else L3(1, 1);
#undef L3
#undef P3
#undef U3
    } else {
        SSR_REQUIRE(n_p == 6 && n_q == 7, "tangent assembly (2D): P2 triangles with 7 points supported (got n_p=%lld, n_q=%d)",
                    (long long)n_p, n_q);
        const int64_t groups = (n_e + 2) / 3;
        const unsigned grid = (unsigned)std::min<int64_t>(groups, want * 16);
#define L2(DSL_, SC_) hipLaunchKernelGGL((k_assemble<2, 6, 7, DSL_, SC_>), dim3(grid), dim3(64), 0, ssr_stream(), \
                                         n_e, D1, D2, D3, DS, W, emap, smap, out)
#define P2(DSL_) hipLaunchKernelGGL((k_assemble_pairs<2, 6, 7, DSL_>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, \
                                    D1, D2, D3, DS, W, out, nullptr, nullptr)
        if (scat == 2 && ssr_tuning(SSR_TUNE_ELEMENT_KERNEL) != 1) {
            if (dsl == 0) P2(0);
            else P2(1);
        } else if (scat == 2) {
            if (dsl == 0) L2(0, 2);
            else L2(1, 2);
        } else if (dsl == 0 && scat == 0) L2(0, 0);
        else if (dsl == 0) L2(0, 1);
        else if (scat == 0) L2(1, 0);
        else L2(1, 1);
#undef L2
#undef P2
    }
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// Element pair blocks into the slot lists, then the upper-block gather
// (REG: K_r from K_elast and r; else K_t).  Needs the slot tables.
static int slot_assembly(const ssr_pattern *P, int n_q, const double *D1, const double *D2, const double *D3,
                         const double *DS, int dsl, const double *W, bool reg, double r, const double *Kel,
                         double *out)
{
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    if (n_e == 0) return SSR_OK;
    double *slots = nullptr;
    int rc = ssr_pattern_keblk(P, &slots);
    if (rc) return rc;
    const int32_t *elem = ssr_pattern_d_elem(P), *pslot = ssr_pattern_d_pslot(P);
    const int64_t want = 4096;
    if (dim == 3) {
        SSR_REQUIRE(n_p == 10 && n_q == 11, "tangent assembly (3D): P2 tetrahedra with 11 points supported");
        const unsigned grid = (unsigned)std::min<int64_t>(n_e, want * 16);
        // two elements per group, one group per element pair: the elements in
        // flight stay a contiguous range, so their slot writes share lines
        const unsigned grid2 = (unsigned)std::min<int64_t>((n_e + 1) / 2, INT32_MAX);
        const bool two = ssr_tuning(SSR_TUNE_ELEMENT_KERNEL) == 0;
        if (two && dsl == 0 && aligned16(D1, D2, D3, DS))
            hipLaunchKernelGGL((k_assemble_pairs2p<true>), dim3(pipelined_grid(n_e)), dim3(64), 0, ssr_stream(), n_e,
                               D1, D2, D3, DS, W, slots, elem, pslot);
        else if (two && dsl == 0)
            hipLaunchKernelGGL((k_assemble_pairs2<0, true>), dim3(grid2), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
        else if (two)
            hipLaunchKernelGGL((k_assemble_pairs2<1, true>), dim3(grid2), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
        else if (dsl == 0)
            hipLaunchKernelGGL((k_assemble_pairs<3, 10, 11, 0, true>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
        else
            hipLaunchKernelGGL((k_assemble_pairs<3, 10, 11, 1, true>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
    } else {
        SSR_REQUIRE(n_p == 6 && n_q == 7, "tangent assembly (2D): P2 triangles with 7 points supported");
        const unsigned grid = (unsigned)std::min<int64_t>((n_e + 2) / 3, want * 16);
        if (dsl == 0)
            hipLaunchKernelGGL((k_assemble_pairs<2, 6, 7, 0, true>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
        else
            hipLaunchKernelGGL((k_assemble_pairs<2, 6, 7, 1, true>), dim3(grid), dim3(64), 0, ssr_stream(), n_e, D1,
                               D2, D3, DS, W, slots, elem, pslot);
    }
    SSR_HIP_CHECK(hipGetLastError());
    const int64_t n_ub = ssr_pattern_n_ublk(P);
    const int64_t threads = n_ub * dim * dim;
    if (threads == 0) return SSR_OK;
#define GS(D_, R_)                                                                                                  \
    hipLaunchKernelGGL((k_gather_slots<D_, R_>), dim3(ssr_blocks(threads, 256)), dim3(256), 0, ssr_stream(), n_ub,  \
                       ssr_pattern_d_ublk(P), ssr_pattern_d_uptr(P), slots, ssr_pattern_d_brow(P),                  \
                       ssr_pattern_d_col(P), ssr_pattern_d_tblk(P), ssr_pattern_d_mask(P), r, Kel, out)
    if (dim == 3) {
        if (reg) GS(3, true);
        else GS(3, false);
    } else {
        if (reg) GS(2, true);
        else GS(2, false);
    }
#undef GS
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

static bool slot_form(const ssr_pattern *P)
{
    return ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) == 0 && ssr_tuning(SSR_TUNE_ELEMENT_KERNEL) != 1 &&
           ssr_pattern_d_pslot(P) && ssr_pattern_d_uptr(P) && ssr_pattern_d_ublk(P);
}

extern "C" {

// Reference-layout tangent values on device pointers: V (nnz) is zeroed and
// filled through the int64 scatter map (1-based, 0 = skip).
int ssr_dev_assemble_K_tangent_vals(int dim, int64_t n_p, int64_t n_int, int64_t n_e, int n_q,
                                    const double *DPhi1, const double *DPhi2, const double *DPhi3,
                                    const double *DS, const double *WEIGHT, const int64_t *smap,
                                    int64_t nnz, double *V)
{
    SSR_REQUIRE(dim == 2 || dim == 3, "assemble_K_tangent_vals: dim must be 2 or 3");
    SSR_REQUIRE(n_e * n_q == n_int, "assemble_K_tangent_vals: n_int (%lld) != n_e*n_q (%lld)",
                (long long)n_int, (long long)(n_e * n_q));
    SSR_REQUIRE(nnz >= 0, "assemble_K_tangent_vals: nnz_out must be >= 0");
    if (nnz) SSR_HIP_CHECK(hipMemsetAsync(V, 0, sizeof(double) * nnz, ssr_stream()));
    return assemble_launch(dim, n_p, n_e, n_q, 0, 1, DPhi1, DPhi2, DPhi3, DS, WEIGHT, nullptr, smap, V);
}

// Node-block values of sum_e Ke on the pattern (zeroed first).  ds_layout:
// 0 = MEX layout DS (ns*ns per point), 1 = packed symmetric (ns(ns+1)/2).
int ssr_dev_assemble_bsr(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                         const double *DPhi3, const double *DS, int ds_layout, const double *WEIGHT,
                         double *Kvals)
{
    SSR_REQUIRE(P, "assemble_bsr: null pattern");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    SSR_REQUIRE(ssr_pattern_d_emap(P), "assemble_bsr: pattern has no device tables");
    if (slot_form(P)) return slot_assembly(P, n_q, DPhi1, DPhi2, DPhi3, DS, ds_layout, WEIGHT, false, 0.0, nullptr,
                                           Kvals);
    if (ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) == 0 && ssr_pattern_d_cptr(P)) {
        double *ke = nullptr;
        int rc = ssr_pattern_keblk(P, &ke);
        if (rc) return rc;
        if ((rc = assemble_launch(dim, n_p, n_e, n_q, ds_layout, 2, DPhi1, DPhi2, DPhi3, DS, WEIGHT, nullptr,
                                  nullptr, ke)))
            return rc;
        const int64_t threads = nnzb * dim * dim;
        if (dim == 3)
            hipLaunchKernelGGL(k_gather_blocks<3>, dim3(ssr_blocks(threads, 256)), dim3(256), 0, ssr_stream(), nnzb,
                               ssr_pattern_d_cptr(P), ssr_pattern_d_ccode(P), ke, Kvals);
        else
            hipLaunchKernelGGL(k_gather_blocks<2>, dim3(ssr_blocks(threads, 256)), dim3(256), 0, ssr_stream(), nnzb,
                               ssr_pattern_d_cptr(P), ssr_pattern_d_ccode(P), ke, Kvals);
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
    SSR_HIP_CHECK(hipMemsetAsync(Kvals, 0, sizeof(double) * nnzb * dim * dim, ssr_stream()));
    return assemble_launch(dim, n_p, n_e, n_q, ds_layout, 0, DPhi1, DPhi2, DPhi3, DS, WEIGHT,
                           ssr_pattern_d_emap(P), nullptr, Kvals);
}

int ssr_dev_strain(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                   const double *DPhi3, const double *U, double *E)
{
    SSR_REQUIRE(P, "strain: null pattern");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    const int64_t n = n_e * n_q;
    const int th = 256;
    if (dim == 3) {
        SSR_REQUIRE(n_p == 10 && n_q == 11, "strain: P2 tetrahedra with 11 points supported");
        hipLaunchKernelGGL((k_strain<3, 10, 11>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n_e,
                           ssr_pattern_d_elem(P), DPhi1, DPhi2, DPhi3, U, E);
    } else {
        SSR_REQUIRE(n_p == 6 && n_q == 7, "strain: P2 triangles with 7 points supported");
        hipLaunchKernelGGL((k_strain<2, 6, 7>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n_e,
                           ssr_pattern_d_elem(P), DPhi1, DPhi2, DPhi3, U, E);
    }
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_force(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                  const double *DPhi3, const double *S, int s_rows, const double *WEIGHT, double *F)
{
    SSR_REQUIRE(P, "force: null pattern");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    if (dim == 3)
        SSR_REQUIRE(n_p == 10 && n_q == 11 && s_rows >= 6, "force: P2 tetrahedra with 11 points supported");
    else
        SSR_REQUIRE(n_p == 6 && n_q == 7 && s_rows >= 3, "force: P2 triangles with 7 points supported");
    const int th = 256;
    const unsigned grid = ssr_blocks(n_e * n_p, th);
    const int32_t *elem = ssr_pattern_d_elem(P);
    if (ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) == 0 && ssr_pattern_d_nptr(P)) {
        // deterministic: element-node forces, then one ordered sum per node
        double *fe = nullptr;
        int rc = ssr_pattern_fe(P, &fe);
        if (rc) return rc;
        if (dim == 3)
            hipLaunchKernelGGL((k_force<3, 10, 11, true>), dim3(grid), dim3(th), 0, ssr_stream(), n_e, elem, DPhi1,
                               DPhi2, DPhi3, S, s_rows, WEIGHT, fe);
        else
            hipLaunchKernelGGL((k_force<2, 6, 7, true>), dim3(grid), dim3(th), 0, ssr_stream(), n_e, elem, DPhi1,
                               DPhi2, DPhi3, S, s_rows, WEIGHT, fe);
        SSR_HIP_CHECK(hipGetLastError());
        const unsigned g2 = ssr_blocks(n_n * dim, th);
        if (dim == 3)
            hipLaunchKernelGGL(k_gather_force<3>, dim3(g2), dim3(th), 0, ssr_stream(), n_n, ssr_pattern_d_nptr(P),
                               ssr_pattern_d_ncode(P), fe, F);
        else
            hipLaunchKernelGGL(k_gather_force<2>, dim3(g2), dim3(th), 0, ssr_stream(), n_n, ssr_pattern_d_nptr(P),
                               ssr_pattern_d_ncode(P), fe, F);
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
    SSR_HIP_CHECK(hipMemsetAsync(F, 0, sizeof(double) * dim * n_n, ssr_stream()));
    if (dim == 3)
        hipLaunchKernelGGL((k_force<3, 10, 11, false>), dim3(grid), dim3(th), 0, ssr_stream(), n_e, elem, DPhi1,
                           DPhi2, DPhi3, S, s_rows, WEIGHT, F);
    else
        hipLaunchKernelGGL((k_force<2, 6, 7, false>), dim3(grid), dim3(th), 0, ssr_stream(), n_e, elem, DPhi1,
                           DPhi2, DPhi3, S, s_rows, WEIGHT, F);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// ssr_dev_strain with each point's gradients computed from the node
// coordinates (dim x n_n node-major) instead of read from DPhi.
int ssr_dev_strain_geo(const ssr_pattern *P, int n_q, const double *coord, const double *U, double *E)
{
    SSR_REQUIRE(P && coord && U && E, "strain_geo: null argument");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    const int64_t n = n_e * n_q;
    const int th = 256;
    if (dim == 3) {
        SSR_REQUIRE(n_p == 10 && n_q == 11, "strain_geo: P2 tetrahedra with 11 points supported");
        hipLaunchKernelGGL((k_strain<3, 10, 11, true>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n_e,
                           ssr_pattern_d_elem(P), nullptr, nullptr, nullptr, U, E, coord);
    } else {
        SSR_REQUIRE(n_p == 6 && n_q == 7, "strain_geo: P2 triangles with 7 points supported");
        hipLaunchKernelGGL((k_strain<2, 6, 7, true>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n_e,
                           ssr_pattern_d_elem(P), nullptr, nullptr, nullptr, U, E, coord);
    }
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// ssr_dev_force with each point's gradients computed from the node
// coordinates; the same forms (ordered per-node sums or atomics) and bits.
int ssr_dev_force_geo(const ssr_pattern *P, int n_q, const double *coord, const double *S, int s_rows,
                      const double *WEIGHT, double *F)
{
    SSR_REQUIRE(P && coord && S && WEIGHT && F, "force_geo: null argument");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    if (dim == 3)
        SSR_REQUIRE(n_p == 10 && n_q == 11 && s_rows >= 6, "force_geo: P2 tetrahedra with 11 points supported");
    else
        SSR_REQUIRE(n_p == 6 && n_q == 7 && s_rows >= 3, "force_geo: P2 triangles with 7 points supported");
    constexpr int EB3 = 23, EB2 = 36;  // elements per 256-lane block: 253 and 252 points
    const unsigned grid = (unsigned)((n_e + (dim == 3 ? EB3 : EB2) - 1) / (dim == 3 ? EB3 : EB2));
    const int32_t *elem = ssr_pattern_d_elem(P);
    const bool gather = ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) == 0 && ssr_pattern_d_nptr(P);
    if (gather) {
        double *fe = nullptr;
        int rc = ssr_pattern_fe(P, &fe);
        if (rc) return rc;
        if (dim == 3)
            hipLaunchKernelGGL((k_force_geo<3, 10, 11, EB3, true>), dim3(grid), dim3(256), 0, ssr_stream(), n_e, elem,
                               coord, S, s_rows, WEIGHT, fe);
        else
            hipLaunchKernelGGL((k_force_geo<2, 6, 7, EB2, true>), dim3(grid), dim3(256), 0, ssr_stream(), n_e, elem,
                               coord, S, s_rows, WEIGHT, fe);
        SSR_HIP_CHECK(hipGetLastError());
        const unsigned g2 = ssr_blocks(n_n * dim, 256);
        if (dim == 3)
            hipLaunchKernelGGL(k_gather_force<3>, dim3(g2), dim3(256), 0, ssr_stream(), n_n, ssr_pattern_d_nptr(P),
                               ssr_pattern_d_ncode(P), fe, F);
        else
            hipLaunchKernelGGL(k_gather_force<2>, dim3(g2), dim3(256), 0, ssr_stream(), n_n, ssr_pattern_d_nptr(P),
                               ssr_pattern_d_ncode(P), fe, F);
        SSR_HIP_CHECK(hipGetLastError());
        return SSR_OK;
    }
    SSR_HIP_CHECK(hipMemsetAsync(F, 0, sizeof(double) * dim * n_n, ssr_stream()));
    if (dim == 3)
        hipLaunchKernelGGL((k_force_geo<3, 10, 11, EB3, false>), dim3(grid), dim3(256), 0, ssr_stream(), n_e, elem,
                           coord, S, s_rows, WEIGHT, F);
    else
        hipLaunchKernelGGL((k_force_geo<2, 6, 7, EB2, false>), dim3(grid), dim3(256), 0, ssr_stream(), n_e, elem,
                           coord, S, s_rows, WEIGHT, F);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_regularize(const ssr_pattern *P, double r, const double *Kel, const double *Kt, double *Kr)
{
    const int32_t *brow = ssr_pattern_d_brow(P);
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    const int64_t n = nnzb * dim * dim;
    if (dim == 3)
        hipLaunchKernelGGL(k_regularize<3>, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), nnzb, brow,
                           ssr_pattern_d_col(P), ssr_pattern_d_tblk(P), ssr_pattern_d_mask(P), r, Kel, Kt, Kr);
    else
        hipLaunchKernelGGL(k_regularize<2>, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), nnzb, brow,
                           ssr_pattern_d_col(P), ssr_pattern_d_tblk(P), ssr_pattern_d_mask(P), r, Kel, Kt, Kr);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// K_r = r K_elast + (1-r) sum_e B'(w DS)B, symmetrised and Dirichlet-masked,
// written once (CONSTITUTIVE.build_K_r_QQ_vals, CONSTITUTIVE.m:487-532):
// element pair blocks, then one gather-and-regularise pass over the upper
// blocks.  Equal bit for bit to ssr_dev_assemble_bsr followed by
// ssr_dev_regularize in the gather form; the atomic form (tuning key 5) and
// patterns without contribution lists take that two-pass path.
int ssr_dev_assemble_K_r(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                         const double *DPhi3, const double *DS, int ds_layout, const double *WEIGHT, double r,
                         const double *K_elast, double *K_r)
{
    SSR_REQUIRE(P && K_elast && K_r && K_elast != K_r, "assemble_K_r: bad arguments");
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    SSR_REQUIRE(ssr_pattern_d_emap(P), "assemble_K_r: pattern has no device tables");
    int rc;
    if (slot_form(P)) return slot_assembly(P, n_q, DPhi1, DPhi2, DPhi3, DS, ds_layout, WEIGHT, true, r, K_elast, K_r);
    if (ssr_tuning(SSR_TUNE_ASSEMBLY_FORM) != 0 || !ssr_pattern_d_cptr(P) || !ssr_pattern_d_ublk(P)) {
        double *Kt = nullptr;
        if ((rc = ssr_pool_malloc((void **)&Kt, sizeof(double) * nnzb * dim * dim))) return rc;
        rc = ssr_dev_assemble_bsr(P, n_q, DPhi1, DPhi2, DPhi3, DS, ds_layout, WEIGHT, Kt);
        if (!rc) rc = ssr_dev_regularize(P, r, K_elast, Kt, K_r);
        ssr_pool_free(Kt);
        return rc;
    }
    double *ke = nullptr;
    if ((rc = ssr_pattern_keblk(P, &ke))) return rc;
    if ((rc = assemble_launch(dim, n_p, n_e, n_q, ds_layout, 2, DPhi1, DPhi2, DPhi3, DS, WEIGHT, nullptr, nullptr,
                              ke)))
        return rc;
    const int64_t n_ub = ssr_pattern_n_ublk(P);
    const int64_t threads = n_ub * dim * dim;
    if (threads == 0) return SSR_OK;
#define GR(D_)                                                                                                    \
    hipLaunchKernelGGL(k_gather_regularize<D_>, dim3(ssr_blocks(threads, 256)), dim3(256), 0, ssr_stream(), n_ub, \
                       ssr_pattern_d_ublk(P), ssr_pattern_d_brow(P), ssr_pattern_d_col(P), ssr_pattern_d_tblk(P),   \
                       ssr_pattern_d_cptr(P), ssr_pattern_d_ccode(P), ke, ssr_pattern_d_mask(P), r, K_elast, K_r)
    if (dim == 3) GR(3);
    else GR(2);
#undef GR
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// V_ref (nnz_free) = block values gathered into the reference's order.
int ssr_dev_extract_ref(const ssr_pattern *P, const double *Kvals, double *Vref)
{
    int dim;
    int64_t n_n, n_p, n_e, nnzb, nnzf;
    ssr_pattern_sizes(P, &dim, &n_n, &n_p, &n_e, &nnzb, &nnzf);
    if (nnzf == 0) return SSR_OK;
    hipLaunchKernelGGL(k_gather_ref, dim3(ssr_blocks(nnzf, 256)), dim3(256), 0, ssr_stream(), nnzf,
                       ssr_pattern_d_ref_pos(P), Kvals, Vref);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

int ssr_dev_ds_elastic(int dim, int64_t n_int, const double *shear, const double *bulk, double *DS)
{
    const int ns = dim == 3 ? 6 : 3;
    const int64_t n = n_int * ns * ns;
    if (n == 0) return SSR_OK;
    hipLaunchKernelGGL(k_ds_elastic, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n_int, ns, shear,
                       bulk, DS);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

}  // extern "C"
