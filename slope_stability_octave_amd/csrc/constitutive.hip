// Mohr-Coulomb return mapping (stress S and consistent tangent DS per
// integration point) for gfx950.
//
// Reference behaviour: constitutive_3D_kernel.h:101-410 (3D) and
// constitutive_2D_kernel.h:36-369 (plane strain), driven per point by
// constitutive_problem_{2D,3D}_{S,SDS}_mex.c.
//
// Design (one lane per point):
//  * Branch selection (elastic / smooth / left edge / right edge / apex) uses
//    the same invariants, trigonometric eigenvalues and tests as the
//    reference, evaluated with FP contraction OFF so that the decisions see
//    the reference's rounding.
//  * Every branch is then reduced to a handful of scalars: S and DS are
//    linear combinations of the eigenprojections P1..P3, the identity vector
//    iota, the trial strain t, the operator d(T^2)/dT and the Voigt identity.
//    One shared epilogue evaluates DS = a*dT2 + b*Id + sum C_uv u (x) v, so
//    lanes of a wavefront that took different branches do not serialise over
//    five separate 6x6 loops.  DS is symmetric (associated flow), so only the
//    21 upper entries are formed.
//  * The MEX-layout entry points read/write point-major (AoS) arrays;
//    packed/SoA variants serve the device-resident Newton path.
#include "ssr_common.h"

namespace {

constexpr double kTwoPiOver3 = 2.0 * 3.14159265358979323846 / 3.0;

// d(T^2)/dT entry (row i, col j) in the mixed Voigt convention of
// constitutive_3D_kernel.h:48-75 (symmetric).
__host__ __device__ __forceinline__ double dsq3(int i, int j, const double t[6])
{
    if (i > j) { int k = i; i = j; j = k; }
    if (i == j) {
        if (i < 3) return 2.0 * t[i];
        if (i == 3) return 0.5 * (t[0] + t[1]);
        if (i == 4) return 0.5 * (t[1] + t[2]);
        return 0.5 * (t[0] + t[2]);
    }
    if ((i == 0 && j == 3) || (i == 1 && j == 3)) return t[3];
    if ((i == 0 && j == 5) || (i == 2 && j == 5)) return t[5];
    if ((i == 1 && j == 4) || (i == 2 && j == 4)) return t[4];
    if (i == 3 && j == 4) return 0.5 * t[5];
    if (i == 3 && j == 5) return 0.5 * t[4];
    if (i == 4 && j == 5) return 0.5 * t[3];
    return 0.0;
}

// Branch-reduced coefficients for one 3D point.
struct Coef3 {
    double sP1, sP3, sI, sT;                  // S = sP1 P1 + sP3 P3 + sI iota + sT t
    double a, b;                              // DS += a dT2 + b Id
    double c11, c33, c13, cII, c1I, c3I, cIT, c1T, c3T;  // symmetric C over (P1,P3,iota,t)
    double g1, g3;                            // 1/((e_k-e_m)(e_k-e_n)) of the projections in use
    bool use1, use3;
};

// Returns branch code 0..4 (elastic, smooth, left edge, right edge, apex).
__host__ __device__ __forceinline__ int mc3d_coef(const double t[6], double cb, double sphi, double G,
                                         double lam, double &e1, double &e2, double &e3, Coef3 &k)
{
#pragma clang fp contract(off)
    const double spsi = sphi;  // associated flow
    const double I1 = t[0] + t[1] + t[2];
    const double I2 = t[0] * t[1] + t[0] * t[2] + t[1] * t[2] - t[3] * t[3] - t[4] * t[4] - t[5] * t[5];
    const double I3 = t[0] * t[1] * t[2] - t[2] * t[3] * t[3] - t[1] * t[5] * t[5] - t[0] * t[4] * t[4] +
                      2.0 * t[3] * t[4] * t[5];
    double Q = (I1 * I1 - 3.0 * I2) / 9.0;
    Q = Q < 0.0 ? 0.0 : Q;
    const double R = (-2.0 * I1 * I1 * I1 + 9.0 * I1 * I2 - 27.0 * I3) / 54.0;
    const double rq = sqrt(Q);
    double th = 0.0;
    if (!(Q < 1e-30)) {
        double c = R / (rq * Q);
        c = c < -1.0 ? -1.0 : (c > 1.0 ? 1.0 : c);
        th = acos(c) / 3.0;
    }
    e1 = -2.0 * rq * cos(th + kTwoPiOver3) + I1 / 3.0;
    e2 = -2.0 * rq * cos(th - kTwoPiOver3) + I1 / 3.0;
    e3 = -2.0 * rq * cos(th) + I1 / 3.0;

    // field by field: an aggregate `k = Coef3{}` was left partly in scratch
    // memory (32 bytes per lane stored and reloaded through HBM)
    k.sP1 = k.sP3 = k.sI = k.sT = 0.0;
    k.a = k.b = 0.0;
    k.c11 = k.c33 = k.c13 = k.cII = k.c1I = k.c3I = k.cIT = k.c1T = k.c3T = 0.0;
    k.g1 = k.g3 = 0.0;
    k.use1 = k.use3 = false;
    const double f_tr = 2.0 * G * ((1.0 + sphi) * e1 - (1.0 - sphi) * e3) + 2.0 * lam * sphi * I1 - cb;
    if (f_tr <= 0.0) {
        k.sI = lam * I1;
        k.sT = 2.0 * G;
        k.b = 2.0 * G;
        k.cII = lam;
        return 0;
    }
    const double g_sl = (e1 - e2) / (1.0 + spsi);
    const double g_sr = (e2 - e3) / (1.0 - spsi);
    const double den_s = 4.0 * lam * sphi * spsi + 4.0 * G * (1.0 + sphi * spsi);
    const double lam_s = f_tr / den_s;
    if (lam_s <= g_sl && lam_s <= g_sr) {
        const double g1 = 1.0 / ((e1 - e2) * (e1 - e3));
        const double g2 = 1.0 / ((e2 - e1) * (e2 - e3));  // used in A2 only
        const double g3 = 1.0 / ((e3 - e1) * (e3 - e2));
        const double s1 = lam * I1 + 2.0 * G * e1 - lam_s * (2.0 * lam * spsi + 2.0 * G * (1.0 + spsi));
        const double s2 = lam * I1 + 2.0 * G * e2 - lam_s * 2.0 * lam * spsi;
        const double s3 = lam * I1 + 2.0 * G * e3 - lam_s * (2.0 * lam * spsi - 2.0 * G * (1.0 - spsi));
        const double A1 = s1 * g1, A2 = s2 * g2, A3 = s3 * g3;
        const double k1a = 2.0 * e1 - e2 - e3, k1b = e2 - e3;
        const double k2a = 2.0 * e2 - e1 - e3, k2b = e1 - e3;
        const double k3a = 2.0 * e3 - e1 - e2, k3b = e1 - e2;
        const double r = 1.0 / den_s;
        const double al1 = 2.0 * G * (1.0 + sphi), al3 = -2.0 * G * (1.0 - sphi), alI = 2.0 * lam * sphi;
        // P2 = iota - P1 - P3 (the eigenprojections sum to the identity), so
        // the P2 terms of S and of c22 P2 (x) P2 fold into the P1/P3/iota ones.
        const double c22 = -A1 * k1b - A2 * k2a + A3 * k3b + 2.0 * G;
        k.sP1 = s1 - s2; k.sP3 = s3 - s2; k.sI = s2;
        k.a = A1 + A2 + A3;
        k.b = -(A1 * (e2 + e3) + A2 * (e1 + e3) + A3 * (e1 + e2));
        k.c11 = -A1 * k1a - A2 * k2b - A3 * k3b + 2.0 * G - r * al1 * al1 + c22;
        k.c33 = A1 * k1b + A2 * k2b - A3 * k3a + 2.0 * G - r * al3 * al3 + c22;
        k.c13 = -r * al1 * al3 + c22;
        k.cII = lam - r * alI * alI + c22;
        k.c1I = -r * al1 * alI - c22;
        k.c3I = -r * al3 * alI - c22;
        k.use1 = k.use3 = true;
        k.g1 = g1; k.g3 = g3;
        return 1;
    }
    if (g_sl < g_sr) {
        const double den_l = 4.0 * lam * sphi * spsi + G * (1.0 + sphi) * (1.0 + spsi) +
                             2.0 * G * (1.0 - sphi) * (1.0 - spsi);
        const double lam_l = (G * ((1.0 + sphi) * (e1 + e2) - 2.0 * (1.0 - sphi) * e3) + 2.0 * lam * sphi * I1 - cb) / den_l;
        const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - spsi);
        if (lam_l >= g_sl && lam_l <= g_la) {
            const double g3 = 1.0 / ((e3 - e1) * (e3 - e2));
            const double s1 = lam * I1 + G * (e1 + e2) - lam_l * (2.0 * lam * spsi + G * (1.0 + spsi));
            const double s3 = lam * I1 + 2.0 * G * e3 - lam_l * (2.0 * lam * spsi - 2.0 * G * (1.0 - spsi));
            const double A = (s3 - s1) * g3;
            const double a12 = e1 + e2;
            const double r = 1.0 / den_l;
            const double bI = G * (1.0 + sphi) + 2.0 * lam * sphi;
            const double b3 = -(G * (1.0 + sphi) + 2.0 * G * (1.0 - sphi));
            k.sI = s1; k.sP3 = s3 - s1;
            k.a = A;
            k.b = -A * a12;
            k.cII = A * a12 + lam + G - r * bI * bI;
            k.c3I = -A * a12 + A * e3 - G - r * b3 * bI;
            k.c33 = A * (2.0 * a12 - 4.0 * e3) + 3.0 * G - r * b3 * b3;
            k.cIT = -A;
            k.c3T = A;
            k.use3 = true;
            k.g3 = g3;
            return 2;
        }
    }
    if (g_sl > g_sr) {
        const double den_r = 4.0 * lam * sphi * spsi + 2.0 * G * (1.0 + sphi) * (1.0 + spsi) +
                             G * (1.0 - sphi) * (1.0 - spsi);
        const double lam_r = (G * (2.0 * (1.0 + sphi) * e1 - (1.0 - sphi) * (e2 + e3)) + 2.0 * lam * sphi * I1 - cb) / den_r;
        const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + spsi);
        if (lam_r >= g_sr && lam_r <= g_ra) {
            const double g1 = 1.0 / ((e1 - e2) * (e1 - e3));
            const double s1 = lam * I1 + 2.0 * G * e1 - lam_r * (2.0 * lam * spsi + 2.0 * G * (1.0 + spsi));
            const double s3 = lam * I1 + G * (e2 + e3) - lam_r * (2.0 * lam * spsi - G * (1.0 - spsi));
            const double A = (s1 - s3) * g1;
            const double a23 = e2 + e3;
            const double r = 1.0 / den_r;
            const double b1 = 2.0 * G * (1.0 + sphi) + G * (1.0 - sphi);
            const double bI = 2.0 * lam * sphi - G * (1.0 - sphi);
            k.sI = s3; k.sP1 = s1 - s3;
            k.a = A;
            k.b = -A * a23;
            k.cII = A * a23 + lam + G - r * bI * bI;
            k.c1I = -A * a23 + A * e1 - G - r * b1 * bI;
            k.c11 = A * (2.0 * a23 - 4.0 * e1) + 3.0 * G - r * b1 * b1;
            k.cIT = -A;
            k.c1T = A;
            k.use1 = true;
            k.g1 = g1;
            return 3;
        }
    }
    k.sI = cb / (2.0 * sphi);
    return 4;
}

// Eigenprojection (T^2 - (e_m+e_n) T + e_m e_n I) * g, g = 1/((e_k-e_m)(e_k-e_n))
// (the inverse gap already formed for the branch coefficients).
__host__ __device__ __forceinline__ void proj3(const double t2[6], const double t[6], double g, double em,
                                      double en, bool use, double P[6])
{
    const double sm = em + en, pm = em * en;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        double v = g * (t2[i] - sm * t[i] + (i < 3 ? pm : 0.0));
        P[i] = use ? v : 0.0;
    }
}

// Full point evaluation.  S[6] always, D21 (upper triangle, row-major packed)
// when WITH_DS.
template <bool WITH_DS>
__host__ __device__ __forceinline__ int mc3d_point(const double eps[6], double cb, double sphi, double G,
                                          double lam, double S[6], double D21[21])
{
    double t[6] = {eps[0], eps[1], eps[2], 0.5 * eps[3], 0.5 * eps[4], 0.5 * eps[5]};
    double e1, e2, e3;
    Coef3 k;
    const int br = mc3d_coef(t, cb, sphi, G, lam, e1, e2, e3, k);

    double t2[6];
    t2[0] = t[0] * t[0] + t[3] * t[3] + t[5] * t[5];
    t2[1] = t[1] * t[1] + t[3] * t[3] + t[4] * t[4];
    t2[2] = t[2] * t[2] + t[4] * t[4] + t[5] * t[5];
    t2[3] = t[0] * t[3] + t[1] * t[3] + t[4] * t[5];
    t2[4] = t[3] * t[5] + t[1] * t[4] + t[2] * t[4];
    t2[5] = t[0] * t[5] + t[3] * t[4] + t[2] * t[5];
    double P1[6], P3[6];
    proj3(t2, t, k.g1, e2, e3, k.use1, P1);
    proj3(t2, t, k.g3, e1, e2, k.use3, P3);

#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const double io = i < 3 ? 1.0 : 0.0;
        S[i] = k.sP1 * P1[i] + k.sP3 * P3[i] + k.sI * io + k.sT * t[i];
    }
    if (WITH_DS) {
        // W_u(j) = sum_v C_uv v(j) for u in (P1,P3,iota,t)
        double W1[6], W3[6], WI[6], WT[6];
#pragma unroll
        for (int j = 0; j < 6; ++j) {
            const double io = j < 3 ? 1.0 : 0.0;
            W1[j] = k.c11 * P1[j] + k.c13 * P3[j] + k.c1I * io + k.c1T * t[j];
            W3[j] = k.c13 * P1[j] + k.c33 * P3[j] + k.c3I * io + k.c3T * t[j];
            WI[j] = k.c1I * P1[j] + k.c3I * P3[j] + k.cII * io + k.cIT * t[j];
            WT[j] = k.c1T * P1[j] + k.c3T * P3[j] + k.cIT * io;
        }
        int q = 0;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
#pragma unroll
            for (int j = i; j < 6; ++j) {
                const double id = (i == j) ? (i < 3 ? 1.0 : 0.5) : 0.0;
                double v = P1[i] * W1[j] + P3[i] * W3[j] + t[i] * WT[j];
                if (i < 3) v += WI[j];
                v += k.a * dsq3(i, j, t) + k.b * id;
                D21[q++] = v;
            }
        }
    }
    return br;
}

// Packed upper index of (i,j), i<=j, for 6x6.
__host__ __device__ constexpr int sym6(int i, int j) { return i * 6 - (i * (i - 1)) / 2 + (j - i); }

// ---------------------------------------------------------------------------
// MEX layout kernels: E (6 x n), S (6 x n), DS (36 x n, each 6x6 column-major)
// point-major, material arrays 1 x n.
// ---------------------------------------------------------------------------
template <bool WITH_DS>
__global__ __launch_bounds__(256) void k_mc3d_aos(int64_t n, const double *__restrict__ E,
                                                  const double *__restrict__ cbar,
                                                  const double *__restrict__ sphi,
                                                  const double *__restrict__ shear,
                                                  const double *__restrict__ lame,
                                                  double *__restrict__ S, double *__restrict__ DS,
                                                  int8_t *__restrict__ branch)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double2 *e2 = reinterpret_cast<const double2 *>(E + 6 * p);
    double2 a = e2[0], b = e2[1], c = e2[2];
    double eps[6] = {a.x, a.y, b.x, b.y, c.x, c.y};
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cbar[p], sphi[p], shear[p], lame[p], s, d);
    double2 *s2 = reinterpret_cast<double2 *>(S + 6 * p);
    s2[0] = make_double2(s[0], s[1]);
    s2[1] = make_double2(s[2], s[3]);
    s2[2] = make_double2(s[4], s[5]);
    if (WITH_DS) {
        double2 *o = reinterpret_cast<double2 *>(DS + 36 * p);
#pragma unroll
        for (int k = 0; k < 18; ++k) {
            const int i0 = (2 * k) % 6, j0 = (2 * k) / 6;  // column-major (row i, col j)
            const int i1 = (2 * k + 1) % 6, j1 = (2 * k + 1) / 6;
            const double v0 = d[i0 <= j0 ? sym6(i0, j0) : sym6(j0, i0)];
            const double v1 = d[i1 <= j1 ? sym6(i1, j1) : sym6(j1, i1)];
            o[k] = make_double2(v0, v1);
        }
    }
    if (branch) branch[p] = (int8_t)br;
}

// Same contract, one 64-lane wavefront per block, with every global access of
// the point-major arrays coalesced through LDS: the wave loads its 64x6 strain
// block as one contiguous span, and writes S (64x6) and DS (64x36) as
// contiguous spans of 16-byte stores.  DS is staged as the 21 unique entries
// per point (stride 21 doubles spreads the LDS banks) and expanded to the full
// column-major 6x6 block on the way out.
typedef double dvec2 __attribute__((ext_vector_type(2)));

// 16-byte store; NT selects the non-temporal (streaming) form, which does not
// keep the written lines in the caches.
template <bool NT>
__device__ __forceinline__ void store2(double *p, double a, double b)
{
    dvec2 v = {a, b};
    if (NT) __builtin_nontemporal_store(v, reinterpret_cast<dvec2 *>(p));
    else *reinterpret_cast<dvec2 *>(p) = v;
}

template <bool WITH_DS, bool NT = false>
__global__ __launch_bounds__(64) void k_mc3d_aos_lds(int64_t n, const double *__restrict__ E,
                                                     const double *__restrict__ cbar,
                                                     const double *__restrict__ sphi,
                                                     const double *__restrict__ shear,
                                                     const double *__restrict__ lame,
                                                     double *__restrict__ S, double *__restrict__ DS,
                                                     int8_t *__restrict__ branch)
{
    constexpr int W = 64;
    __shared__ double sbuf[WITH_DS ? W * 21 : W * 6];
    const int lane = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * W;
    const int cnt = (int)((n - p0) < W ? (n - p0) : W);
    const int64_t p = p0 + lane;
    {
        const double2 *Eg = reinterpret_cast<const double2 *>(E + 6 * p0);
        double2 *s2 = reinterpret_cast<double2 *>(sbuf);
        for (int t = lane; t < cnt * 3; t += W) s2[t] = Eg[t];
    }
    __syncthreads();
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double cb = 1.0, sp = 0.5, G = 1.0, L = 1.0;
    if (lane < cnt) {
#pragma unroll
        for (int i = 0; i < 6; ++i) eps[i] = sbuf[lane * 6 + i];
        cb = cbar[p];
        sp = sphi[p];
        G = shear[p];
        L = lame[p];
    }
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
    __syncthreads();
    {
        double *Sg = S + 6 * p0;
        for (int t = lane; t < cnt * 3; t += W) store2<NT>(Sg + 2 * t, sbuf[2 * t], sbuf[2 * t + 1]);
    }
    if (WITH_DS) {
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 21; ++k) sbuf[lane * 21 + k] = d[k];
        __syncthreads();
        double *Dg = DS + 36 * p0;
        for (int t = lane; t < cnt * 18; t += W) {
            const int o = 2 * t, q = o / 36, c = o - 36 * q;  // c even: c, c+1 in the same point
            const int i0 = c % 6, j0 = c / 6, i1 = (c + 1) % 6, j1 = (c + 1) / 6;
            const double v0 = sbuf[q * 21 + (i0 <= j0 ? sym6(i0, j0) : sym6(j0, i0))];
            const double v1 = sbuf[q * 21 + (i1 <= j1 ? sym6(i1, j1) : sym6(j1, i1))];
            store2<NT>(Dg + o, v0, v1);
        }
    }
    if (branch && lane < cnt) branch[p] = (int8_t)br;
}

// Same as k_mc3d_aos_lds with half the LDS: DS is staged and written out in
// two phases of 32 points (5.4 kB per wavefront instead of 10.75 kB), which
// roughly doubles the wavefronts an LDS-limited compute unit can hold.
template <bool WITH_DS, bool NT, int MINW = 1>
__global__ __launch_bounds__(64, MINW) void k_mc3d_aos_half(int64_t n, const double *__restrict__ E,
                                                      const double *__restrict__ cbar,
                                                      const double *__restrict__ sphi,
                                                      const double *__restrict__ shear,
                                                      const double *__restrict__ lame,
                                                      double *__restrict__ S, double *__restrict__ DS,
                                                      int8_t *__restrict__ branch)
{
    constexpr int W = 64, H = 32;
    __shared__ double sbuf[H * 21];  // >= W*6 for the E / S staging
    const int lane = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * W;
    const int cnt = (int)((n - p0) < W ? (n - p0) : W);
    const int64_t p = p0 + lane;
    {
        const double2 *Eg = reinterpret_cast<const double2 *>(E + 6 * p0);
        double2 *s2 = reinterpret_cast<double2 *>(sbuf);
        for (int t = lane; t < cnt * 3; t += W) s2[t] = Eg[t];
    }
    __syncthreads();
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double cb = 1.0, sp = 0.5, G = 1.0, L = 1.0;
    if (lane < cnt) {
#pragma unroll
        for (int i = 0; i < 6; ++i) eps[i] = sbuf[lane * 6 + i];
        cb = cbar[p];
        sp = sphi[p];
        G = shear[p];
        L = lame[p];
    }
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
    __syncthreads();
    {
        double *Sg = S + 6 * p0;
        for (int t = lane; t < cnt * 3; t += W) store2<NT>(Sg + 2 * t, sbuf[2 * t], sbuf[2 * t + 1]);
    }
    if (WITH_DS) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ch = cnt - h * H < H ? (cnt - h * H > 0 ? cnt - h * H : 0) : H;
            __syncthreads();
            if (lane >= h * H && lane < h * H + H) {
#pragma unroll
                for (int k = 0; k < 21; ++k) sbuf[(lane - h * H) * 21 + k] = d[k];
            }
            __syncthreads();
            double *Dg = DS + 36 * (p0 + h * H);
            for (int t = lane; t < ch * 18; t += W) {
                const int o = 2 * t, q = o / 36, c = o - 36 * q;
                const int i0 = c % 6, j0 = c / 6, i1 = (c + 1) % 6, j1 = (c + 1) / 6;
                const double v0 = sbuf[q * 21 + (i0 <= j0 ? sym6(i0, j0) : sym6(j0, i0))];
                const double v1 = sbuf[q * 21 + (i1 <= j1 ? sym6(i1, j1) : sym6(j1, i1))];
                store2<NT>(Dg + o, v0, v1);
            }
        }
    }
    if (branch && lane < cnt) branch[p] = (int8_t)br;
}

// Half-staged form whose LDS rows hold each point's full column-major 6x6
// block (36 doubles) instead of the 21 packed entries: the symmetric
// expansion happens once per lane at compile-time offsets, and the write-out
// is a plain 16-byte copy of the contiguous span (no per-element index
// arithmetic, which dominated the VALU count of the packed form).
template <bool WITH_DS, bool NT>
__global__ __launch_bounds__(64) void k_mc3d_x36(int64_t n, const double *__restrict__ E,
                                                 const double *__restrict__ cbar,
                                                 const double *__restrict__ sphi,
                                                 const double *__restrict__ shear,
                                                 const double *__restrict__ lame,
                                                 double *__restrict__ S, double *__restrict__ DS,
                                                 int8_t *__restrict__ branch)
{
    constexpr int W = 64, H = 32;
    __shared__ __align__(16) double sbuf[WITH_DS ? H * 36 : W * 6];  // H*36 >= W*6
    const int lane = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * W;
    const int cnt = (int)((n - p0) < W ? (n - p0) : W);
    const int64_t p = p0 + lane;
    {
        const double2 *Eg = reinterpret_cast<const double2 *>(E + 6 * p0);
        double2 *s2 = reinterpret_cast<double2 *>(sbuf);
        for (int t = lane; t < cnt * 3; t += W) s2[t] = Eg[t];
    }
    __syncthreads();
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double cb = 1.0, sp = 0.5, G = 1.0, L = 1.0;
    if (lane < cnt) {
#pragma unroll
        for (int i = 0; i < 6; ++i) eps[i] = sbuf[lane * 6 + i];
        cb = cbar[p];
        sp = sphi[p];
        G = shear[p];
        L = lame[p];
    }
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
    __syncthreads();
#pragma unroll
    for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
    __syncthreads();
    {
        double *Sg = S + 6 * p0;
        for (int t = lane; t < cnt * 3; t += W) store2<NT>(Sg + 2 * t, sbuf[2 * t], sbuf[2 * t + 1]);
    }
    if (WITH_DS) {
        const dvec2 *src = reinterpret_cast<const dvec2 *>(sbuf);
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int ch = cnt - h * H < H ? (cnt - h * H > 0 ? cnt - h * H : 0) : H;
            __syncthreads();
            if (lane >= h * H && lane < h * H + H) {
                double *row = sbuf + (lane - h * H) * 36;
#pragma unroll
                for (int j = 0; j < 6; ++j)
#pragma unroll
                    for (int i = 0; i < 6; ++i) row[i + 6 * j] = d[i <= j ? sym6(i, j) : sym6(j, i)];
            }
            __syncthreads();
            double *Dg = DS + 36 * (p0 + h * H);
            const int m = ch * 18;
#pragma unroll
            for (int k = 0; k < H * 18 / W; ++k) {
                const int t = lane + W * k;
                if (t < m) {
                    const dvec2 v = src[t];
                    store2<NT>(Dg + 2 * t, v.x, v.y);
                }
            }
        }
    }
    if (branch && lane < cnt) branch[p] = (int8_t)br;
}

// k_mc3d_x36 with the inputs read straight into registers: every load of
// the wavefront (each lane's 48-byte strain row as three 16-byte loads, and
// the four material values) is issued up front, so a single memory latency
// precedes the arithmetic and no LDS round trip or barrier is spent on E.
// DS is staged in phases of H points.
// MI: the material fields are per material, indexed by the material of the
// point's element (mat_id[p / n_q]) -- the reduced strength of a scalar
// lambda then costs one evaluation per material instead of a pass over the
// points (the same values, so the same bits).
template <bool WITH_DS, bool NT, int H, int MINW = 1, bool SDIRECT = false, bool MI = false>
__global__ __launch_bounds__(64, MINW) void k_mc3d_x36d(int64_t n, const double *__restrict__ E,
                                                  const double *__restrict__ cbar,
                                                  const double *__restrict__ sphi,
                                                  const double *__restrict__ shear,
                                                  const double *__restrict__ lame,
                                                  double *__restrict__ S, double *__restrict__ DS,
                                                  int8_t *__restrict__ branch,
                                                  const int32_t *__restrict__ mat_id = nullptr, int n_q = 1)
{
    constexpr int W = 64;
    static_assert(W % H == 0, "phase size");
    __shared__ __align__(16) double sbuf[WITH_DS ? (H * 36 > W * 6 ? H * 36 : W * 6) : W * 6];
    const int lane = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * W;
    const int cnt = (int)((n - p0) < W ? (n - p0) : W);
    const int64_t p = p0 + lane;
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double cb = 1.0, sp = 0.5, G = 1.0, L = 1.0;
    if (lane < cnt) {
        const double2 *e2 = reinterpret_cast<const double2 *>(E + 6 * p);
        const double2 a = e2[0], b = e2[1], c = e2[2];
        const int64_t m = MI ? (int64_t)mat_id[(uint32_t)p / (uint32_t)n_q] : p;
        cb = cbar[m];
        sp = sphi[m];
        G = shear[m];
        L = lame[m];
        eps[0] = a.x; eps[1] = a.y; eps[2] = b.x; eps[3] = b.y; eps[4] = c.x; eps[5] = c.y;
    }
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
    if (SDIRECT) {
        // each lane's 48-byte stress row as three 16-byte stores
        if (lane < cnt) {
            double *Sg = S + 6 * p;
            store2<NT>(Sg, s[0], s[1]);
            store2<NT>(Sg + 2, s[2], s[3]);
            store2<NT>(Sg + 4, s[4], s[5]);
        }
    } else {
#pragma unroll
        for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
        __syncthreads();
        double *Sg = S + 6 * p0;
        for (int t = lane; t < cnt * 3; t += W) store2<NT>(Sg + 2 * t, sbuf[2 * t], sbuf[2 * t + 1]);
    }
    if (WITH_DS) {
        const dvec2 *src = reinterpret_cast<const dvec2 *>(sbuf);
#pragma unroll
        for (int h = 0; h < W / H; ++h) {
            const int ch = cnt - h * H < H ? (cnt - h * H > 0 ? cnt - h * H : 0) : H;
            __syncthreads();
            if (lane >= h * H && lane < h * H + H) {
                double *row = sbuf + (lane - h * H) * 36;
#pragma unroll
                for (int j = 0; j < 6; ++j)
#pragma unroll
                    for (int i = 0; i < 6; ++i) row[i + 6 * j] = d[i <= j ? sym6(i, j) : sym6(j, i)];
            }
            __syncthreads();
            double *Dg = DS + 36 * (p0 + h * H);
            const int m = ch * 18;
#pragma unroll
            for (int k = 0; k < (H * 18 + W - 1) / W; ++k) {
                const int t = lane + W * k;
                if (t < m) {
                    const dvec2 v = src[t];
                    store2<NT>(Dg + 2 * t, v.x, v.y);
                }
            }
        }
    }
    if (branch && lane < cnt) branch[p] = (int8_t)br;
}

// k_mc3d_x36d with the points of a wavefront spread over the array: lane l
// of wavefront w takes point 16 (q nw + w) + (l mod 16) with q = l / 16 and
// nw wavefronts in the grid.  Each 16-point phase of a wavefront then writes
// its S and DS rows next to those of the neighbouring wavefronts (which the
// dispatcher places on other XCDs), and a wavefront's four 4.6 kB DS spans
// lie a quarter of the array apart, instead of one 18 kB span per wavefront.
// A streaming-write probe of these two patterns measured 6.5 TB/s for 4 kB
// spans per single-wave group against 5.5 TB/s for 18 kB spans.
template <bool WITH_DS, bool NT>
__global__ __launch_bounds__(64) void k_mc3d_x36s(int64_t n, const double *__restrict__ E,
                                                  const double *__restrict__ cbar,
                                                  const double *__restrict__ sphi,
                                                  const double *__restrict__ shear,
                                                  const double *__restrict__ lame,
                                                  double *__restrict__ S, double *__restrict__ DS,
                                                  int8_t *__restrict__ branch)
{
    constexpr int W = 64, H = 16, NPH = W / H;
    __shared__ __align__(16) double sbuf[H * 36 > W * 6 ? H * 36 : W * 6];
    const int lane = threadIdx.x;
    const int64_t nw = gridDim.x;
    const int q = lane / H, j = lane % H;
    const int64_t p = ((int64_t)q * nw + blockIdx.x) * H + j;
    const bool act = p < n;
    double eps[6] = {0, 0, 0, 0, 0, 0};
    double cb = 1.0, sp = 0.5, G = 1.0, L = 1.0;
    if (act) {
        const double2 *e2 = reinterpret_cast<const double2 *>(E + 6 * p);
        const double2 a = e2[0], b = e2[1], c = e2[2];
        cb = cbar[p];
        sp = sphi[p];
        G = shear[p];
        L = lame[p];
        eps[0] = a.x; eps[1] = a.y; eps[2] = b.x; eps[3] = b.y; eps[4] = c.x; eps[5] = c.y;
    }
    double s[6], d[21];
    const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
    // S: each phase's 16 rows are one 768-byte span
#pragma unroll
    for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const int t = lane + W * k;  // double2 index over the 4 phases x 48
        const int qq = t / 48, tt = t - qq * 48;
        const int64_t g0 = ((int64_t)qq * nw + blockIdx.x) * H;
        const int64_t rem = n - g0;
        const int ch = rem < H ? (rem > 0 ? (int)rem : 0) : H;
        if (tt < ch * 3) store2<NT>(S + 6 * g0 + 2 * tt, sbuf[qq * 96 + 2 * tt], sbuf[qq * 96 + 2 * tt + 1]);
    }
    if (WITH_DS) {
        const dvec2 *src = reinterpret_cast<const dvec2 *>(sbuf);
#pragma unroll
        for (int h = 0; h < NPH; ++h) {
            const int64_t g0 = ((int64_t)h * nw + blockIdx.x) * H;
            const int64_t rem = n - g0;
            const int ch = rem < H ? (rem > 0 ? (int)rem : 0) : H;
            __syncthreads();
            if (q == h) {
                double *row = sbuf + j * 36;
#pragma unroll
                for (int jj = 0; jj < 6; ++jj)
#pragma unroll
                    for (int i = 0; i < 6; ++i) row[i + 6 * jj] = d[i <= jj ? sym6(i, jj) : sym6(jj, i)];
            }
            __syncthreads();
            double *Dg = DS + 36 * g0;
            const int m = ch * 18;
#pragma unroll
            for (int k = 0; k < (H * 18 + W - 1) / W; ++k) {
                const int t = lane + W * k;
                if (t < m) {
                    const dvec2 v = src[t];
                    store2<NT>(Dg + 2 * t, v.x, v.y);
                }
            }
        }
    }
    if (branch && act) branch[p] = (int8_t)br;
}

// Persistent, software-pipelined form of k_mc3d_aos_half: each wavefront
// walks 64-point tiles with the grid as stride and issues the loads of its
// next tile (the contiguous strain span and the four material values) before
// the arithmetic of the current one.  The load latency and the draining of
// the previous tile's stores then overlap the return mapping instead of
// waiting for other wavefronts to cover them.
template <bool WITH_DS, bool NT, int MINW>
__global__ __launch_bounds__(64, MINW) void k_mc3d_pipe(int64_t n, int64_t ntiles, const double *__restrict__ E,
                                                        const double *__restrict__ cbar,
                                                        const double *__restrict__ sphi,
                                                        const double *__restrict__ shear,
                                                        const double *__restrict__ lame,
                                                        double *__restrict__ S, double *__restrict__ DS,
                                                        int8_t *__restrict__ branch)
{
    constexpr int W = 64, H = 32;
    __shared__ double sbuf[H * 21];  // >= 3*W double2 of the strain span
    const int lane = threadIdx.x;
    int64_t tile = blockIdx.x;
    if (tile >= ntiles) return;
    double2 pe0, pe1, pe2;
    double pcb, psp, pG, pL;
    auto fetch = [&](int64_t tl) {
        const int64_t q0 = tl * W;
        const int c = (int)((n - q0) < W ? (n - q0) : W);
        const int m = 3 * c;
        const double2 *Eg = reinterpret_cast<const double2 *>(E + 6 * q0);
        const double2 z = make_double2(0.0, 0.0);
        pe0 = lane < m ? Eg[lane] : z;
        pe1 = lane + W < m ? Eg[lane + W] : z;
        pe2 = lane + 2 * W < m ? Eg[lane + 2 * W] : z;
        if (lane < c) {
            const int64_t q = q0 + lane;
            pcb = cbar[q];
            psp = sphi[q];
            pG = shear[q];
            pL = lame[q];
        } else {
            pcb = 1.0; psp = 0.5; pG = 1.0; pL = 1.0;
        }
    };
    fetch(tile);
    for (; tile < ntiles; tile += gridDim.x) {
        const int64_t p0 = tile * W;
        const int cnt = (int)((n - p0) < W ? (n - p0) : W);
        const int64_t p = p0 + lane;
        double2 *s2 = reinterpret_cast<double2 *>(sbuf);
        __syncthreads();
        s2[lane] = pe0;
        s2[lane + W] = pe1;
        s2[lane + 2 * W] = pe2;
        __syncthreads();
        double eps[6];
#pragma unroll
        for (int i = 0; i < 6; ++i) eps[i] = sbuf[lane * 6 + i];
        const double cb = pcb, sp = psp, G = pG, L = pL;
        const int64_t nxt = tile + gridDim.x;
        if (nxt < ntiles) fetch(nxt);
        double s[6], d[21];
        const int br = mc3d_point<WITH_DS>(eps, cb, sp, G, L, s, d);
        __syncthreads();
#pragma unroll
        for (int i = 0; i < 6; ++i) sbuf[lane * 6 + i] = s[i];
        __syncthreads();
        {
            double *Sg = S + 6 * p0;
            for (int t = lane; t < cnt * 3; t += W) store2<NT>(Sg + 2 * t, sbuf[2 * t], sbuf[2 * t + 1]);
        }
        if (WITH_DS) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int ch = cnt - h * H < H ? (cnt - h * H > 0 ? cnt - h * H : 0) : H;
                __syncthreads();
                if (lane >= h * H && lane < h * H + H) {
#pragma unroll
                    for (int k = 0; k < 21; ++k) sbuf[(lane - h * H) * 21 + k] = d[k];
                }
                __syncthreads();
                double *Dg = DS + 36 * (p0 + h * H);
                for (int t = lane; t < ch * 18; t += W) {
                    const int o = 2 * t, q = o / 36, c = o - 36 * q;
                    const int i0 = c % 6, j0 = c / 6, i1 = (c + 1) % 6, j1 = (c + 1) / 6;
                    const double v0 = sbuf[q * 21 + (i0 <= j0 ? sym6(i0, j0) : sym6(j0, i0))];
                    const double v1 = sbuf[q * 21 + (i1 <= j1 ? sym6(i1, j1) : sym6(j1, i1))];
                    store2<NT>(Dg + o, v0, v1);
                }
            }
        }
        if (branch && lane < cnt) branch[p] = (int8_t)br;
    }
}

// ---------------------------------------------------------------------------
// 2D (plane strain)
// ---------------------------------------------------------------------------
// Ordered slot data for the three principal strains.
template <bool WITH_DS>
__host__ __device__ __forceinline__ int mc2d_point(const double eps[3], double cb, double sphi, double G,
                                          double lam, double S[4], double D6[6])
{
    const double x = eps[0], y = eps[1], gm = eps[2];
    double e1, e2, e3;
    int br;
    // slot -> unordered eigenvalue: 0 = l+, 1 = l-, 2 = out-of-plane
    int o1, o2, o3;
    double s1 = 0, s2 = 0, s3 = 0, r = 0, b1 = 0, b2 = 0, b3 = 0, bI = 0;
    double c1 = 0, c2 = 0, c3 = 0, c12 = 0, c23 = 0, cII = 0, bId = 0, sI = 0, sT = 0;
    double h1 = 0, h2 = 0, h3 = 0;  // coefficient of the second-derivative operator of each slot
    double tr, rad, lp, lm;
    {
#pragma clang fp contract(off)
        tr = x + y;
        rad = sqrt((x - y) * (x - y) + gm * gm);
        lp = 0.5 * (tr + rad);
        lm = 0.5 * (tr - rad);
        const double lz = 0.0;
        e1 = lp; e2 = lm; e3 = lz; o1 = 0; o2 = 1; o3 = 2;
        if ((lp >= lz) && (lz > lm)) { e2 = lz; e3 = lm; o2 = 2; o3 = 1; }
        if (lz > lp) { e1 = lz; e2 = lp; e3 = lm; o1 = 2; o2 = 0; o3 = 1; }
        const double trE = e1 + e2 + e3;
        const double f_tr = 2.0 * G * ((1.0 + sphi) * e1 - (1.0 - sphi) * e3) + 2.0 * lam * sphi * trE - cb;
        const double g_sl = (e1 - e2) / (1.0 + sphi);
        const double g_sr = (e2 - e3) / (1.0 - sphi);
        const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sphi);
        const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sphi);
        const double den_s = 4.0 * lam * sphi * sphi + 2.0 * G * (1.0 + sphi) * (1.0 + sphi) +
                             2.0 * G * (1.0 - sphi) * (1.0 - sphi);
        const double den_l = 4.0 * lam * sphi * sphi + G * (1.0 + sphi) * (1.0 + sphi) +
                             2.0 * G * (1.0 - sphi) * (1.0 - sphi);
        const double den_r = 4.0 * lam * sphi * sphi + 2.0 * G * (1.0 + sphi) * (1.0 + sphi) +
                             G * (1.0 - sphi) * (1.0 - sphi);
        const double lam_s = f_tr / den_s;
        const double lam_l = (G * ((1.0 + sphi) * (e1 + e2) - 2.0 * (1.0 - sphi) * e3) + 2.0 * lam * sphi * trE - cb) / den_l;
        const double lam_r = (G * (2.0 * (1.0 + sphi) * e1 - (1.0 - sphi) * (e2 + e3)) + 2.0 * lam * sphi * trE - cb) / den_r;
        if (f_tr <= 0.0) br = 0;
        else if (lam_s <= ((g_sl < g_sr) ? g_sl : g_sr)) br = 1;
        else if ((g_sl < g_sr) && (lam_l >= g_sl) && (lam_l <= g_la)) br = 2;
        else if ((g_sl > g_sr) && (lam_r >= g_sr) && (lam_r <= g_ra)) br = 3;
        else br = 4;

        if (br == 0) {
            sI = lam * tr; sT = 2.0 * G; bId = 2.0 * G; cII = lam;
        } else if (br == 1) {
            s1 = lam * trE + 2.0 * G * e1 - lam_s * (2.0 * lam * sphi + 2.0 * G * (1.0 + sphi));
            s2 = lam * trE + 2.0 * G * e2 - lam_s * (2.0 * lam * sphi);
            s3 = lam * trE + 2.0 * G * e3 - lam_s * (2.0 * lam * sphi - 2.0 * G * (1.0 - sphi));
            h1 = s1; h2 = s2; h3 = s3;
            c1 = c2 = c3 = 2.0 * G; cII = lam;
            b1 = 2.0 * G * (1.0 + sphi); b3 = -2.0 * G * (1.0 - sphi); bI = 2.0 * lam * sphi;
            r = 1.0 / den_s;
        } else if (br == 2) {
            const double sa = lam * trE + G * (e1 + e2) - lam_l * (2.0 * lam * sphi + G * (1.0 + sphi));
            const double sc = lam * trE + 2.0 * G * e3 - lam_l * (2.0 * lam * sphi - 2.0 * G * (1.0 - sphi));
            s1 = sa; s2 = sa; s3 = sc;
            h1 = sa; h2 = sa; h3 = sc;
            c1 = c2 = c12 = G; c3 = 2.0 * G; cII = lam;
            b1 = b2 = G * (1.0 + sphi); b3 = -2.0 * G * (1.0 - sphi); bI = 2.0 * lam * sphi;
            r = 1.0 / den_l;
        } else if (br == 3) {
            const double sa = lam * trE + 2.0 * G * e1 - lam_r * (2.0 * lam * sphi + 2.0 * G * (1.0 + sphi));
            const double sc = lam * trE + G * (e2 + e3) - lam_r * (2.0 * lam * sphi - G * (1.0 - sphi));
            s1 = sa; s2 = sc; s3 = sc;
            h1 = sa; h2 = sc; h3 = sc;
            c1 = 2.0 * G; c2 = c3 = c23 = G; cII = lam;
            b1 = 2.0 * G * (1.0 + sphi); b2 = b3 = -G * (1.0 - sphi); bI = 2.0 * lam * sphi;
            r = 1.0 / den_r;
        } else {
            const double sa = cb / (2.0 * sphi);
            sI = sa;
        }
    }
    // unordered first derivatives of l+ (dp), l- (dm), out-of-plane (dz)
    double dp[3], dm[3];
    if (rad != 0.0) {
        dp[0] = (x - lm) / rad; dp[1] = (y - lm) / rad; dp[2] = 0.5 * gm / rad;
    } else {
        dp[0] = 1.0; dp[1] = 1.0; dp[2] = 0.0;
    }
    dm[0] = 1.0 - dp[0]; dm[1] = 1.0 - dp[1]; dm[2] = -dp[2];
    double v1[3], v2[3], v3[3];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        v1[i] = o1 == 0 ? dp[i] : (o1 == 1 ? dm[i] : 0.0);
        v2[i] = o2 == 0 ? dp[i] : (o2 == 1 ? dm[i] : 0.0);
        v3[i] = o3 == 0 ? dp[i] : (o3 == 1 ? dm[i] : 0.0);
    }
    const double z1 = o1 == 2 ? 1.0 : 0.0, z2 = o2 == 2 ? 1.0 : 0.0, z3 = o3 == 2 ? 1.0 : 0.0;
    // S: (s11, s22, s12, s33); the trial tensor is (x, y, gm/2, 0)
    const double tt[3] = {x, y, 0.5 * gm};
#pragma unroll
    for (int i = 0; i < 3; ++i)
        S[i] = s1 * v1[i] + s2 * v2[i] + s3 * v3[i] + (i < 2 ? sI : 0.0) + sT * tt[i];
    S[3] = s1 * z1 + s2 * z2 + s3 * z3 + sI;
    if (WITH_DS) {
        // second derivative of l+ is Hp, of l- is -Hp, of the out-of-plane value 0
        const double eta1 = o1 == 0 ? 1.0 : (o1 == 1 ? -1.0 : 0.0);
        const double eta2 = o2 == 0 ? 1.0 : (o2 == 1 ? -1.0 : 0.0);
        const double eta3 = o3 == 0 ? 1.0 : (o3 == 1 ? -1.0 : 0.0);
        const double h = h1 * eta1 + h2 * eta2 + h3 * eta3;
        const double irad = rad != 0.0 ? 1.0 / rad : 0.0;
        double w[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) w[i] = b1 * v1[i] + b2 * v2[i] + b3 * v3[i] + (i < 2 ? bI : 0.0);
        int q = 0;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
#pragma unroll
            for (int j = i; j < 3; ++j) {
                const double dg = (i == j) ? (i < 2 ? 1.0 : 0.5) : 0.0;
                const double Hp = (dg - dp[i] * dp[j] - dm[i] * dm[j]) * irad;
                double v = h * Hp + bId * dg;
                v += c1 * v1[i] * v1[j] + c2 * v2[i] * v2[j] + c3 * v3[i] * v3[j];
                v += c12 * (v1[i] * v2[j] + v2[i] * v1[j]) + c23 * (v2[i] * v3[j] + v3[i] * v2[j]);
                v += (i < 2 && j < 2) ? cII : 0.0;
                v -= r * w[i] * w[j];
                D6[q++] = v;
            }
        }
    }
    return br;
}

__host__ __device__ constexpr int sym3(int i, int j) { return i * 3 - (i * (i - 1)) / 2 + (j - i); }

template <bool WITH_DS, bool MI = false>
__global__ __launch_bounds__(256) void k_mc2d_aos(int64_t n, const double *__restrict__ E,
                                                  const double *__restrict__ cbar,
                                                  const double *__restrict__ sphi,
                                                  const double *__restrict__ shear,
                                                  const double *__restrict__ lame,
                                                  double *__restrict__ S, double *__restrict__ DS,
                                                  int8_t *__restrict__ branch,
                                                  const int32_t *__restrict__ mat_id = nullptr, int n_q = 1)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double eps[3] = {E[3 * p], E[3 * p + 1], E[3 * p + 2]};
    double s[4], d[6];
    const int64_t m = MI ? (int64_t)mat_id[(uint32_t)p / (uint32_t)n_q] : p;
    const int br = mc2d_point<WITH_DS>(eps, cbar[m], sphi[m], shear[m], lame[m], s, d);
    double2 *s2 = reinterpret_cast<double2 *>(S + 4 * p);
    s2[0] = make_double2(s[0], s[1]);
    s2[1] = make_double2(s[2], s[3]);
    if (WITH_DS) {
        double *o = DS + 9 * p;
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            const int i = k % 3, j = k / 3;
            o[k] = d[i <= j ? sym3(i, j) : sym3(j, i)];
        }
    }
    if (branch) branch[p] = (int8_t)br;
}

// The same 2D contract with one 64-lane wavefront per block and every global
// access of the point-major arrays coalesced through LDS: the wave reads its
// 64 x 3 strains as one contiguous span of 16-byte loads, and writes S
// (64 x 4) and DS (64 x 9) as contiguous spans of 16-byte non-temporal
// stores (the direct form's lanes store 8 bytes at a 72-byte stride, nine
// times).  The point arithmetic is mc2d_point, as in k_mc2d_aos: the same
// bits.  E, S and DS must be 16-byte aligned (the dispatcher checks).
template <bool WITH_DS, bool MI = false>
__global__ __launch_bounds__(64) void k_mc2d_lds(int64_t n, const double *__restrict__ E,
                                                 const double *__restrict__ cbar, const double *__restrict__ sphi,
                                                 const double *__restrict__ shear, const double *__restrict__ lame,
                                                 double *__restrict__ S, double *__restrict__ DS,
                                                 int8_t *__restrict__ branch, const int32_t *__restrict__ mat_id,
                                                 int n_q)
{
    __shared__ double se[64 * 3];
    __shared__ double ss[64 * 4];
    __shared__ double sd[WITH_DS ? 64 * 9 : 2];
    const int l = threadIdx.x;
    const int64_t p0 = (int64_t)blockIdx.x * 64;
    const int cnt = (int)((n - p0) < 64 ? (n - p0) : 64);
    const double *Eb = E + 3 * p0;
    const int ne = 3 * cnt;
    // the material loads are issued before the strain span is staged
    double mc = 0.0, ms = 0.0, mg = 0.0, ml = 0.0;
    if (l < cnt) {
        const int64_t p = p0 + l;
        const int64_t m = MI ? (int64_t)mat_id[(uint32_t)p / (uint32_t)n_q] : p;
        mc = cbar[m];
        ms = sphi[m];
        mg = shear[m];
        ml = lame[m];
    }
    for (int i = l; 2 * i + 1 < ne; i += 64) {
        const dvec2 v = *reinterpret_cast<const dvec2 *>(Eb + 2 * i);
        se[2 * i] = v.x;
        se[2 * i + 1] = v.y;
    }
    if ((ne & 1) && l == 0) se[ne - 1] = Eb[ne - 1];
    __syncthreads();
    int br = 0;
    if (l < cnt) {
        double eps[3] = {se[3 * l], se[3 * l + 1], se[3 * l + 2]};
        double sv[4], d[6];
        br = mc2d_point<WITH_DS>(eps, mc, ms, mg, ml, sv, d);
#pragma unroll
        for (int k = 0; k < 4; ++k) ss[4 * l + k] = sv[k];
        if (WITH_DS) {
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                const int i = k % 3, j = k / 3;
                sd[9 * l + k] = d[i <= j ? sym3(i, j) : sym3(j, i)];
            }
        }
    }
    __syncthreads();
    double *Sb = S + 4 * p0;
    for (int i = l; i < 2 * cnt; i += 64) store2<true>(Sb + 2 * i, ss[2 * i], ss[2 * i + 1]);
    if (WITH_DS) {
        double *Db = DS + 9 * p0;
        const int nd = 9 * cnt;
        for (int i = l; 2 * i + 1 < nd; i += 64) store2<true>(Db + 2 * i, sd[2 * i], sd[2 * i + 1]);
        if ((nd & 1) && l == 0) Db[nd - 1] = sd[nd - 1];
    }
    if (branch && l < cnt) branch[p0 + l] = (int8_t)br;
}

// Davis strength reduction per point (reduction.m:25-58) for a scalar lambda
// or one lambda per point.  Davis C's vector `if` is resolved on the host.
__global__ void k_reduction(int64_t n, int davis, int c_all, const double *__restrict__ c0,
                            const double *__restrict__ phi, const double *__restrict__ psi,
                            const double *__restrict__ lam, int lam_stride,
                            double *__restrict__ cbar, double *__restrict__ sphi)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
  {
#pragma clang fp contract(off)
    const double l = lam[lam_stride ? p : 0];
    double c_l, phi_l;
    if (davis == 'A') {
        const double beta = cos(phi[p]) * cos(psi[p]) / (1.0 - sin(phi[p]) * sin(psi[p]));
        c_l = beta * c0[p] / l;
        phi_l = atan(beta * tan(phi[p]) / l);
    } else if (davis == 'B') {
        const double c01 = c0[p] / l;
        const double phi1 = atan(tan(phi[p]) / l);
        const double psi1 = atan(tan(psi[p]) / l);
        const double beta = cos(phi1) * cos(psi1) / (1.0 - sin(phi1) * sin(psi1));
        c_l = beta * c01;
        phi_l = atan(beta * tan(phi1));
    } else {
        const double c01 = c0[p] / l;
        const double phi1 = atan(tan(phi[p]) / l);
        const double beta = c_all ? cos(phi1) * cos(psi[p]) / (1.0 - sin(phi1) * sin(psi[p])) : 1.0;
        c_l = beta * c01;
        phi_l = atan(beta * tan(phi1));
    }
    cbar[p] = 2.0 * c_l * cos(phi_l);
    sphi[p] = sin(phi_l);
  }
}

}  // namespace

// ---------------------------------------------------------------------------
// Launchers (device pointers, library stream)
// ---------------------------------------------------------------------------
extern "C" int ssr_dev_constitutive_3D(int64_t n, const double *E, const double *c_bar,
                                       const double *sin_phi, const double *shear,
                                       const double *bulk, const double *lame, double *S,
                                       double *DS, int8_t *branch)
{
    (void)bulk;
    SSR_REQUIRE(n >= 0, "constitutive_3D: n_int must be >= 0");
    if (n == 0) return SSR_OK;
    SSR_REQUIRE(E && c_bar && sin_phi && shear && lame && S, "constitutive_3D: null argument");
    if (ssr_tuning(SSR_TUNE_CONSTITUTIVE_VARIANT) == 0) {
        const int th = 256;
        if (DS)
            hipLaunchKernelGGL(k_mc3d_aos<true>, dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n, E,
                               c_bar, sin_phi, shear, lame, S, DS, branch);
        else
            hipLaunchKernelGGL(k_mc3d_aos<false>, dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n,
                               E, c_bar, sin_phi, shear, lame, S, DS, branch);
    } else if (ssr_tuning(SSR_TUNE_CONSTITUTIVE_VARIANT) == 1) {
        if (DS)
            hipLaunchKernelGGL((k_mc3d_aos_lds<true, false>), dim3(ssr_blocks(n, 64)), dim3(64), 0, ssr_stream(), n,
                               E, c_bar, sin_phi, shear, lame, S, DS, branch);
        else
            hipLaunchKernelGGL((k_mc3d_aos_lds<false, false>), dim3(ssr_blocks(n, 64)), dim3(64), 0, ssr_stream(), n,
                               E, c_bar, sin_phi, shear, lame, S, DS, branch);
    } else {
        const int v = ssr_tuning(SSR_TUNE_CONSTITUTIVE_VARIANT);
        const dim3 grid(ssr_blocks(n, 64)), blk(64);
#define MC3(K_) hipLaunchKernelGGL(K_, grid, blk, 0, ssr_stream(), n, E, c_bar, sin_phi, shear, lame, S, DS, branch)
        const int64_t ntiles = (n + 63) / 64;
        const int wpc = ssr_tuning(SSR_TUNE_PERSIST_WAVES);
        auto pgrid = [&](int minw) {
            const int64_t g = (int64_t)ssr_cu_count() * (wpc > 0 ? wpc : 4 * minw);
            return dim3((unsigned)(g < ntiles ? g : ntiles));
        };
#define MCP(K_, MW_) hipLaunchKernelGGL(K_, pgrid(MW_), blk, 0, ssr_stream(), n, ntiles, E, c_bar, sin_phi, shear, lame, S, DS, branch)
        if (v == 2) {
            if (DS) MC3((k_mc3d_aos_lds<true, true>));
            else MC3((k_mc3d_aos_lds<false, true>));
        } else if (v == 3) {
            if (DS) MC3((k_mc3d_aos_half<true, false>));
            else MC3((k_mc3d_aos_half<false, false>));
        } else if (v == 4) {
            if (DS) MC3((k_mc3d_aos_half<true, true>));
            else MC3((k_mc3d_aos_half<false, true>));
        } else if (v == 5) {
            if (DS) MCP((k_mc3d_pipe<true, true, 1>), 4);
            else MCP((k_mc3d_pipe<false, true, 1>), 4);
        } else if (v == 6) {
            if (DS) MCP((k_mc3d_pipe<true, true, 5>), 5);
            else MCP((k_mc3d_pipe<false, true, 5>), 5);
        } else if (v == 7) {
            if (DS) MC3((k_mc3d_aos_half<true, true, 5>));
            else MC3((k_mc3d_aos_half<false, true, 5>));
        } else if (v == 8) {
            if (DS) MC3((k_mc3d_x36<true, false>));
            else MC3((k_mc3d_x36<false, false>));
        } else if (v == 9) {
            if (DS) MC3((k_mc3d_x36<true, true>));
            else MC3((k_mc3d_x36<false, true>));
        } else if (v == 10) {
            if (DS) MC3((k_mc3d_x36d<true, true, 32>));
            else MC3((k_mc3d_x36d<false, true, 32>));
        } else if (v == 11) {
            if (DS) MC3((k_mc3d_x36d<true, true, 16>));
            else MC3((k_mc3d_x36d<false, true, 16>));
        } else if (v == 12) {
            if (DS) MC3((k_mc3d_x36d<true, true, 16, 5>));
            else MC3((k_mc3d_x36d<false, true, 16, 5>));
        } else if (v == 13) {
            if (DS) MC3((k_mc3d_x36d<true, true, 16, 1, true>));
            else MC3((k_mc3d_x36d<false, true, 16, 1, true>));
        } else if (v == 14) {
            if (DS) MC3((k_mc3d_x36s<true, true>));
            else MC3((k_mc3d_x36s<false, true>));
        } else {
            if (DS) MC3((k_mc3d_x36s<true, false>));
            else MC3((k_mc3d_x36s<false, false>));
        }
#undef MCP
#undef MC3
    }
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// 2D form: the LDS-staged kernel (tuning key 15 = 1, default) for S and DS
// when E, S and DS are 16-byte aligned, else one lane per point with direct
// accesses (S alone: the direct form measured faster, 0.220 against
// 0.235 ms for 13.2M points).
static bool mc2d_staged(const double *E, const double *S, const double *DS)
{
    const uintptr_t a = (uintptr_t)E | (uintptr_t)S | (uintptr_t)DS;
    return ssr_tuning(SSR_TUNE_CONSTITUTIVE_2D) != 0 && DS && (a & 15) == 0;
}

// The same return mappings with per-material fields: point p takes the
// fields of material mat_id[p / n_q] (its element's).  3D runs the default
// kernel form.
extern "C" int ssr_dev_constitutive_mat(int dim, int64_t n, int n_q, const double *E, const int32_t *mat_id,
                                        int64_t n_mat, const double *c_bar, const double *sin_phi,
                                        const double *shear, const double *bulk, const double *lame, double *S,
                                        double *DS, int8_t *branch)
{
    (void)bulk;
    SSR_REQUIRE(dim == 2 || dim == 3, "constitutive_mat: dim must be 2 or 3");
    SSR_REQUIRE(n >= 0 && n_q > 0 && n_mat > 0, "constitutive_mat: bad sizes");
    SSR_REQUIRE(n < ((int64_t)1 << 32), "constitutive_mat: too many points for 32-bit element indices");
    if (n == 0) return SSR_OK;
    SSR_REQUIRE(E && mat_id && c_bar && sin_phi && shear && lame && S, "constitutive_mat: null argument");
    if (dim == 3) {
        const dim3 grid(ssr_blocks(n, 64)), blk(64);
        if (DS)
            hipLaunchKernelGGL((k_mc3d_x36d<true, true, 16, 1, false, true>), grid, blk, 0, ssr_stream(), n, E, c_bar,
                               sin_phi, shear, lame, S, DS, branch, mat_id, n_q);
        else
            hipLaunchKernelGGL((k_mc3d_x36d<false, true, 16, 1, false, true>), grid, blk, 0, ssr_stream(), n, E,
                               c_bar, sin_phi, shear, lame, S, DS, branch, mat_id, n_q);
    } else if (mc2d_staged(E, S, DS)) {
        const dim3 grid(ssr_blocks(n, 64)), blk(64);
        if (DS)
            hipLaunchKernelGGL((k_mc2d_lds<true, true>), grid, blk, 0, ssr_stream(), n, E, c_bar, sin_phi, shear,
                               lame, S, DS, branch, mat_id, n_q);
        else
            hipLaunchKernelGGL((k_mc2d_lds<false, true>), grid, blk, 0, ssr_stream(), n, E, c_bar, sin_phi, shear,
                               lame, S, DS, branch, mat_id, n_q);
    } else {
        const int th = 256;
        if (DS)
            hipLaunchKernelGGL((k_mc2d_aos<true, true>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n, E,
                               c_bar, sin_phi, shear, lame, S, DS, branch, mat_id, n_q);
        else
            hipLaunchKernelGGL((k_mc2d_aos<false, true>), dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n, E,
                               c_bar, sin_phi, shear, lame, S, DS, branch, mat_id, n_q);
    }
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

extern "C" int ssr_dev_constitutive_2D(int64_t n, const double *E, const double *c_bar,
                                       const double *sin_phi, const double *shear,
                                       const double *bulk, const double *lame, double *S,
                                       double *DS, int8_t *branch)
{
    (void)bulk;
    SSR_REQUIRE(n >= 0, "constitutive_2D: n_int must be >= 0");
    if (n == 0) return SSR_OK;
    SSR_REQUIRE(E && c_bar && sin_phi && shear && lame && S, "constitutive_2D: null argument");
    const int th = 256;
    if (mc2d_staged(E, S, DS)) {
        const dim3 grid(ssr_blocks(n, 64)), blk(64);
        if (DS)
            hipLaunchKernelGGL((k_mc2d_lds<true>), grid, blk, 0, ssr_stream(), n, E, c_bar, sin_phi, shear, lame, S,
                               DS, branch, nullptr, 1);
        else
            hipLaunchKernelGGL((k_mc2d_lds<false>), grid, blk, 0, ssr_stream(), n, E, c_bar, sin_phi, shear, lame, S,
                               DS, branch, nullptr, 1);
    } else if (DS)
        hipLaunchKernelGGL(k_mc2d_aos<true>, dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n, E,
                           c_bar, sin_phi, shear, lame, S, DS, branch);
    else
        hipLaunchKernelGGL(k_mc2d_aos<false>, dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n,
                           E, c_bar, sin_phi, shear, lame, S, DS, branch);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

// ---------------------------------------------------------------------------
// Potential Psi(E) of the Mohr-Coulomb model per point (potential_3D.m,
// potential_2D.m): the elastic energy of the trial strain minus
// denom * lambda^2 / 2 of the active return; its integral is the energy
// functional of the direct SSR continuation.  Same branch tests as the
// return mapping, evaluated without FP contraction.
// ---------------------------------------------------------------------------
// tI: trace entering the yield test and the multipliers, tr: trace entering
// the energies (the 3D potential uses the invariant I1 for the former and the
// sum of the eigenvalues for the latter; in 2D both are e1+e2+e3).
__device__ __forceinline__ double psi_branches(double e1, double e2, double e3, double tI, double tr, double cb,
                                               double sp, double G, double K, double lam, double den_s)
{
#pragma clang fp contract(off)
    const double f_tr = 2.0 * G * ((1.0 + sp) * e1 - (1.0 - sp) * e3) + 2.0 * (lam * sp) * tI - cb;
    const double el = 0.5 * lam * tr * tr + G * (e1 * e1 + e2 * e2 + e3 * e3);
    if (f_tr <= 0.0) return el;
    const double g_sl = (e1 - e2) / (1.0 + sp);
    const double g_sr = (e2 - e3) / (1.0 - sp);
    const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sp);
    const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sp);
    const double lam_s = f_tr / den_s;
    if (lam_s <= (g_sl < g_sr ? g_sl : g_sr)) return el - 0.5 * den_s * (lam_s * lam_s);
    const double den_l = 4.0 * lam * sp * sp + G * (1.0 + sp) * (1.0 + sp) + 2.0 * G * (1.0 - sp) * (1.0 - sp);
    const double lam_l = (G * ((1.0 + sp) * (e1 + e2) - 2.0 * (1.0 - sp) * e3) + 2.0 * lam * sp * tI - cb) / den_l;
    if (g_sl < g_sr && lam_l >= g_sl && lam_l <= g_la)
        return 0.5 * lam * tr * tr + G * (e3 * e3 + 0.5 * (e1 + e2) * (e1 + e2)) - 0.5 * den_l * (lam_l * lam_l);
    const double den_r = 4.0 * lam * sp * sp + 2.0 * G * (1.0 + sp) * (1.0 + sp) + G * (1.0 - sp) * (1.0 - sp);
    const double lam_r = (G * (2.0 * (1.0 + sp) * e1 - (1.0 - sp) * (e2 + e3)) + 2.0 * lam * sp * tI - cb) / den_r;
    if (g_sl > g_sr && lam_r >= g_sr && lam_r <= g_ra)
        return 0.5 * lam * tr * tr + G * (e1 * e1 + 0.5 * (e2 + e3) * (e2 + e3)) - 0.5 * den_r * (lam_r * lam_r);
    const double den_a = 4.0 * K * sp * sp;
    const double lam_a = (2.0 * K * sp * tI - cb) / den_a;
    return 0.5 * K * tr * tr - 0.5 * den_a * (lam_a * lam_a);
}

__global__ __launch_bounds__(256) void k_potential3d(int64_t n, const double *__restrict__ E,
                                                     const double *__restrict__ cbar, const double *__restrict__ sphi,
                                                     const double *__restrict__ shear, const double *__restrict__ bulk,
                                                     const double *__restrict__ lame, double *__restrict__ psi)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double *e = E + 6 * p;
    double ps;
    {
#pragma clang fp contract(off)
        const double t0 = e[0], t1 = e[1], t2 = e[2], t3 = 0.5 * e[3], t4 = 0.5 * e[4], t5 = 0.5 * e[5];
        const double I1 = t0 + t1 + t2;
        const double I2 = t0 * t1 + t0 * t2 + t1 * t2 - t3 * t3 - t4 * t4 - t5 * t5;
        const double I3 = t0 * t1 * t2 - t2 * t3 * t3 - t1 * t5 * t5 - t0 * t4 * t4 + 2.0 * t3 * t4 * t5;
        double Q = (1.0 / 9.0) * (I1 * I1 - 3.0 * I2);
        Q = Q > 0.0 ? Q : 0.0;
        const double R = (1.0 / 54.0) * (-2.0 * (I1 * I1 * I1) + 9.0 * (I1 * I2) - 27.0 * I3);
        double th0 = 0.0;
        if (Q > 0.0) th0 = R / sqrt(Q * Q * Q);
        th0 = th0 < -1.0 ? -1.0 : (th0 > 1.0 ? 1.0 : th0);
        const double th = acos(th0) / 3.0;
        const double rq = sqrt(Q);
        const double e1 = -2.0 * rq * cos(th + kTwoPiOver3) + I1 / 3.0;
        const double e2 = -2.0 * rq * cos(th - kTwoPiOver3) + I1 / 3.0;
        const double e3 = -2.0 * rq * cos(th) + I1 / 3.0;
        const double sp = sphi[p], G = shear[p], lam = lame[p];
        const double den_s = 4.0 * lam * sp * sp + 4.0 * G * (1.0 + sp * sp);
        ps = psi_branches(e1, e2, e3, I1, e1 + e2 + e3, cbar[p], sp, G, bulk[p], lam, den_s);
    }
    psi[p] = ps;
}

__global__ __launch_bounds__(256) void k_potential2d(int64_t n, const double *__restrict__ E,
                                                     const double *__restrict__ cbar, const double *__restrict__ sphi,
                                                     const double *__restrict__ shear, const double *__restrict__ bulk,
                                                     const double *__restrict__ lame, double *__restrict__ psi)
{
    const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    const double *e = E + 3 * p;
    double ps;
    {
#pragma clang fp contract(off)
        const double I1 = e[0] + e[1];
        const double I2 = sqrt((e[0] - e[1]) * (e[0] - e[1]) + e[2] * e[2]);
        const double a1 = (I1 + I2) / 2.0, a2 = (I1 - I2) / 2.0, a3 = 0.0;
        double e1 = a1, e2 = a2, e3 = a3;
        if ((a1 >= a3) && (a3 > a2)) { e2 = a3; e3 = a2; }
        if (a3 > a1) { e1 = a3; e2 = a1; e3 = a2; }
        const double tr = e1 + e2 + e3;
        const double sp = sphi[p], G = shear[p], lam = lame[p];
        const double den_s = 4.0 * lam * sp * sp + 2.0 * G * (1.0 + sp) * (1.0 + sp) + 2.0 * G * (1.0 - sp) * (1.0 - sp);
        ps = psi_branches(e1, e2, e3, tr, tr, cbar[p], sp, G, bulk[p], lam, den_s);
    }
    psi[p] = ps;
}

// Psi per point (potential_{2D,3D}.m); E point-major with 3 (2D) or 6 (3D) rows.
extern "C" int ssr_dev_potential(int dim, int64_t n, const double *E, const double *c_bar, const double *sin_phi,
                                 const double *shear, const double *bulk, const double *lame, double *Psi)
{
    SSR_REQUIRE(dim == 2 || dim == 3, "potential: dim must be 2 or 3");
    SSR_REQUIRE(n >= 0, "potential: n_int must be >= 0");
    if (n == 0) return SSR_OK;
    SSR_REQUIRE(E && c_bar && sin_phi && shear && bulk && lame && Psi, "potential: null argument");
    if (dim == 3)
        hipLaunchKernelGGL(k_potential3d, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, E, c_bar, sin_phi,
                           shear, bulk, lame, Psi);
    else
        hipLaunchKernelGGL(k_potential2d, dim3(ssr_blocks(n, 256)), dim3(256), 0, ssr_stream(), n, E, c_bar, sin_phi,
                           shear, bulk, lame, Psi);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}

extern "C" int ssr_dev_reduction(int64_t n, int davis, const double *c0, const double *phi,
                                 const double *psi, const double *lambda, int lambda_per_point,
                                 int davis_c_all, double *c_bar, double *sin_phi)
{
    SSR_REQUIRE(davis == 'A' || davis == 'B' || davis == 'C', "Incorrect choice of the Davis approach.");
    if (n == 0) return SSR_OK;
    const int th = 256;
    hipLaunchKernelGGL(k_reduction, dim3(ssr_blocks(n, th)), dim3(th), 0, ssr_stream(), n, davis,
                       davis_c_all, c0, phi, psi, lambda, lambda_per_point ? 1 : 0, c_bar, sin_phi);
    SSR_HIP_CHECK(hipGetLastError());
    return SSR_OK;
}
