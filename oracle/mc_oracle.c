/*
 * mc_oracle.c -- CPU oracle for the SSR Newton-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * (or as the timed CPU baseline).  The product path never calls it.
 *
 * This file restates, in plain C, the arithmetic of the reference's native
 * per-point and per-element kernels:
 *
 *   orc_reduction      reduction.m:25-58 (Davis A/B/C strength reduction)
 *   orc_mc3d           constitutive_3D_kernel.h:101-410 (Mohr-Coulomb, 3D)
 *   orc_mc2d           constitutive_2D_kernel.h:36-369 (Mohr-Coulomb, plane strain)
 *   orc_assemble3d     assemble_K_tangent_vals.c:85-171 (sum_q B'(w D)B, scatter)
 *   orc_assemble2d     assemble_K_tangent_vals_2D.c:88-160
 *   orc_strain{2,3}d   CONSTITUTIVE.m:218,251 (E = B*U, B as in
 *                      elastic_stiffness_matrix_3D.m:105-132 / _2D.m:90-104)
 *   orc_force{2,3}d    CONSTITUTIVE.m:281 (F = B' * (WEIGHT .* S(1:n_strain,:)))
 *
 * Pinning: orc_mc3d / orc_mc2d are checked against golden vectors written by
 * the reference's own kernels compiled from their sources (tests/golden), and
 * against closed forms (elastic, apex), the yield-surface invariant, symmetry
 * and a central finite difference of S.  orc_reduction is checked against the
 * closed form of Davis B at lambda = 1 and against an independent numpy
 * evaluation.  The assembly, strain and force routines are checked against a
 * dense global B'*D*B / B*U / B'*s built by brute force on small meshes, the
 * elastic identity V_tang(DS_elast) == V_elast and rigid-body null modes.
 *
 * Compiled with -O2 and no -march flags, so on x86-64 no fused multiply-add
 * is generated: the same floating-point contract as the reference MEX build
 * (build_constitutive_3D_mex.sh:33-38).
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846

/* ------------------------------------------------------------------------ */
/* Davis strength reduction (reduction.m:25-58).  lambda may be a scalar     */
/* (lam_stride = 0) or one value per point (lam_stride = 1).                 */
/* Davis C uses an Octave vector `if` (reduction.m:47): the formula branch   */
/* applies to every point only when phi1 > psi holds at every point, else    */
/* beta = 1 everywhere.                                                      */
/* ------------------------------------------------------------------------ */
int orc_reduction(int davis, int64_t n, const double *c0, const double *phi,
                  const double *psi, const double *lam, int lam_stride,
                  double *c_bar, double *sin_phi)
{
    int all_c = 1;
    if (davis == 'C') {
        for (int64_t p = 0; p < n; ++p) {
            double l = lam[lam_stride ? p : 0];
            double phi1 = atan(tan(phi[p]) / l);
            if (!(phi1 > psi[p])) { all_c = 0; break; }
        }
    }
    for (int64_t p = 0; p < n; ++p) {
        double l = lam[lam_stride ? p : 0];
        double c_lam, phi_lam;
        if (davis == 'A') {
            double beta = cos(phi[p]) * cos(psi[p]) / (1.0 - sin(phi[p]) * sin(psi[p]));
            c_lam = beta * c0[p] / l;
            phi_lam = atan(beta * tan(phi[p]) / l);
        } else if (davis == 'B') {
            double c01 = c0[p] / l;
            double phi1 = atan(tan(phi[p]) / l);
            double psi1 = atan(tan(psi[p]) / l);
            double beta = cos(phi1) * cos(psi1) / (1.0 - sin(phi1) * sin(psi1));
            c_lam = beta * c01;
            phi_lam = atan(beta * tan(phi1));
        } else if (davis == 'C') {
            double c01 = c0[p] / l;
            double phi1 = atan(tan(phi[p]) / l);
            double beta = all_c ? cos(phi1) * cos(psi[p]) / (1.0 - sin(phi1) * sin(psi[p])) : 1.0;
            c_lam = beta * c01;
            phi_lam = atan(beta * tan(phi1));
        } else {
            return -1;
        }
        c_bar[p] = 2.0 * c_lam * cos(phi_lam);
        sin_phi[p] = sin(phi_lam);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3D Mohr-Coulomb return mapping                                            */
/* ------------------------------------------------------------------------ */
/* Voigt order (11,22,33,12,23,13).  Tensors carry halved shear ("stress
 * representation"), strains arrive with engineering shear. */
static const double IOTA6[6] = {1, 1, 1, 0, 0, 0};

/* Eigenprojection P_k = (T^2 - (e_m+e_n) T + e_m e_n I) / ((e_k-e_m)(e_k-e_n))
 * (constitutive_3D_kernel.h:33-42). */
static void proj3(const double t2[6], const double t[6], double sum_other,
                  double prod_other, double inv_gap, double P[6])
{
    for (int i = 0; i < 6; ++i)
        P[i] = inv_gap * (t2[i] - sum_other * t[i] + (i < 3 ? prod_other : 0.0));
}

/* d(T^2)/dT as a 6x6 operator in the mixed Voigt convention of
 * constitutive_3D_kernel.h:48-75.  Stored dT2[row][col]. */
static void dsquare3(const double t[6], double M[6][6])
{
    memset(M, 0, sizeof(double) * 36);
    M[0][0] = 2.0 * t[0]; M[3][0] = t[3]; M[5][0] = t[5];
    M[1][1] = 2.0 * t[1]; M[3][1] = t[3]; M[4][1] = t[4];
    M[2][2] = 2.0 * t[2]; M[4][2] = t[4]; M[5][2] = t[5];
    M[0][3] = t[3]; M[1][3] = t[3]; M[3][3] = 0.5 * (t[0] + t[1]);
    M[4][3] = 0.5 * t[5]; M[5][3] = 0.5 * t[4];
    M[1][4] = t[4]; M[2][4] = t[4]; M[3][4] = 0.5 * t[5];
    M[4][4] = 0.5 * (t[1] + t[2]); M[5][4] = 0.5 * t[3];
    M[0][5] = t[5]; M[2][5] = t[5]; M[3][5] = 0.5 * t[4];
    M[4][5] = 0.5 * t[3]; M[5][5] = 0.5 * (t[0] + t[2]);
}

static double identv(int i, int j) { return i == j ? (i < 3 ? 1.0 : 0.5) : 0.0; }

/* Returns the branch code: 0 elastic, 1 smooth, 2 left edge, 3 right edge,
 * 4 apex (constitutive_3D_kernel.h:158,196,261,330,399). */
static int mc3d_point(const double eps[6], double cb, double sphi, double G,
                      double lam, double S[6], double *D /* col-major 36 or NULL */)
{
    const double spsi = sphi; /* associated flow, constitutive_3D_kernel.h:106 */
    double t[6] = {eps[0], eps[1], eps[2], 0.5 * eps[3], 0.5 * eps[4], 0.5 * eps[5]};
    double t2[6];
    t2[0] = t[0] * t[0] + t[3] * t[3] + t[5] * t[5];
    t2[1] = t[1] * t[1] + t[3] * t[3] + t[4] * t[4];
    t2[2] = t[2] * t[2] + t[4] * t[4] + t[5] * t[5];
    t2[3] = t[0] * t[3] + t[1] * t[3] + t[4] * t[5];
    t2[4] = t[3] * t[5] + t[1] * t[4] + t[2] * t[4];
    t2[5] = t[0] * t[5] + t[3] * t[4] + t[2] * t[5];

    /* invariants and trigonometric eigenvalues (kernel.h:122-150) */
    double I1 = t[0] + t[1] + t[2];
    double I2 = t[0] * t[1] + t[0] * t[2] + t[1] * t[2] - t[3] * t[3] - t[4] * t[4] - t[5] * t[5];
    double I3 = t[0] * t[1] * t[2] - t[2] * t[3] * t[3] - t[1] * t[5] * t[5] - t[0] * t[4] * t[4] + 2.0 * t[3] * t[4] * t[5];
    double Qi = (I1 * I1 - 3.0 * I2) / 9.0;
    if (Qi < 0.0) Qi = 0.0;
    double Ri = (-2.0 * I1 * I1 * I1 + 9.0 * I1 * I2 - 27.0 * I3) / 54.0;
    double rq = sqrt(Qi);
    double th = 0.0;
    if (!(Qi < 1e-30)) {
        double c = Ri / (rq * Qi);
        if (c < -1.0) c = -1.0;
        if (c > 1.0) c = 1.0;
        th = acos(c) / 3.0;
    }
    double e1 = -2.0 * rq * cos(th + 2.0 * ORC_PI / 3.0) + I1 / 3.0;
    double e2 = -2.0 * rq * cos(th - 2.0 * ORC_PI / 3.0) + I1 / 3.0;
    double e3 = -2.0 * rq * cos(th) + I1 / 3.0;

    double f_tr = 2.0 * G * ((1.0 + sphi) * e1 - (1.0 - sphi) * e3) + 2.0 * lam * sphi * I1 - cb;

    if (f_tr <= 0.0) { /* elastic, kernel.h:158-185 */
        for (int i = 0; i < 6; ++i) S[i] = (i < 3 ? lam * I1 : 0.0) + 2.0 * G * t[i];
        if (D) {
            for (int j = 0; j < 6; ++j)
                for (int i = 0; i < 6; ++i)
                    D[i + 6 * j] = lam * IOTA6[i] * IOTA6[j] + (i == j ? (i < 3 ? 2.0 * G : G) : 0.0);
        }
        return 0;
    }

    double g_sl = (e1 - e2) / (1.0 + spsi);
    double g_sr = (e2 - e3) / (1.0 - spsi);
    double den_s = 4.0 * lam * sphi * spsi + 4.0 * G * (1.0 + sphi * spsi);
    double lam_s = f_tr / den_s;

    if (lam_s <= g_sl && lam_s <= g_sr) { /* smooth, kernel.h:196-256 */
        double g1 = 1.0 / ((e1 - e2) * (e1 - e3));
        double g2 = 1.0 / ((e2 - e1) * (e2 - e3));
        double g3 = 1.0 / ((e3 - e1) * (e3 - e2));
        double P1[6], P2[6], P3[6];
        proj3(t2, t, e2 + e3, e2 * e3, g1, P1);
        proj3(t2, t, e1 + e3, e1 * e3, g2, P2);
        proj3(t2, t, e1 + e2, e1 * e2, g3, P3);
        double s1 = lam * I1 + 2.0 * G * e1 - lam_s * (2.0 * lam * spsi + 2.0 * G * (1.0 + spsi));
        double s2 = lam * I1 + 2.0 * G * e2 - lam_s * 2.0 * lam * spsi;
        double s3 = lam * I1 + 2.0 * G * e3 - lam_s * (2.0 * lam * spsi - 2.0 * G * (1.0 - spsi));
        for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s2 * P2[i] + s3 * P3[i];
        if (D) {
            double M[6][6];
            dsquare3(t, M);
            double Dp[6], Dq[6];
            for (int i = 0; i < 6; ++i) {
                Dp[i] = 2.0 * G * ((1.0 + sphi) * P1[i] - (1.0 - sphi) * P3[i]) + 2.0 * lam * sphi * IOTA6[i];
                Dq[i] = 2.0 * G * ((1.0 + spsi) * P1[i] - (1.0 - spsi) * P3[i]) + 2.0 * lam * spsi * IOTA6[i];
            }
            double a23 = e2 + e3, a13 = e1 + e3, a12 = e1 + e2;
            double k1a = 2.0 * e1 - e2 - e3, k1b = e2 - e3;
            double k2a = 2.0 * e2 - e1 - e3, k2b = e1 - e3;
            double k3a = 2.0 * e3 - e1 - e2, k3b = e1 - e2;
            double rden = 1.0 / den_s;
            for (int j = 0; j < 6; ++j)
                for (int i = 0; i < 6; ++i) {
                    double Id = identv(i, j);
                    double q11 = P1[i] * P1[j], q22 = P2[i] * P2[j], q33 = P3[i] * P3[j];
                    double d1 = g1 * (M[i][j] - Id * a23 - k1a * q11 - k1b * (q22 - q33));
                    double d2 = g2 * (M[i][j] - Id * a13 - k2a * q22 - k2b * (q11 - q33));
                    double d3 = g3 * (M[i][j] - Id * a12 - k3a * q33 - k3b * (q11 - q22));
                    D[i + 6 * j] = s1 * d1 + s2 * d2 + s3 * d3 + lam * IOTA6[i] * IOTA6[j]
                                   + 2.0 * G * (q11 + q22 + q33) - Dq[i] * Dp[j] * rden;
                }
        }
        return 1;
    }

    if (g_sl < g_sr) { /* left edge, kernel.h:261-325 */
        double den_l = 4.0 * lam * sphi * spsi + G * (1.0 + sphi) * (1.0 + spsi) + 2.0 * G * (1.0 - sphi) * (1.0 - spsi);
        double lam_l = (G * ((1.0 + sphi) * (e1 + e2) - 2.0 * (1.0 - sphi) * e3) + 2.0 * lam * sphi * I1 - cb) / den_l;
        double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - spsi);
        if (lam_l >= g_sl && lam_l <= g_la) {
            double g3 = 1.0 / ((e3 - e1) * (e3 - e2));
            double P3[6], P12[6];
            proj3(t2, t, e1 + e2, e1 * e2, g3, P3);
            for (int i = 0; i < 6; ++i) P12[i] = IOTA6[i] - P3[i];
            double s1 = lam * I1 + G * (e1 + e2) - lam_l * (2.0 * lam * spsi + G * (1.0 + spsi));
            double s3 = lam * I1 + 2.0 * G * e3 - lam_l * (2.0 * lam * spsi - 2.0 * G * (1.0 - spsi));
            for (int i = 0; i < 6; ++i) S[i] = s1 * P12[i] + s3 * P3[i];
            if (D) {
                double M[6][6];
                dsquare3(t, M);
                double a12 = e1 + e2;
                double Dp[6], Dq[6];
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * ((1.0 + sphi) * P12[i] - 2.0 * (1.0 - sphi) * P3[i]) + 2.0 * lam * sphi * IOTA6[i];
                    Dq[i] = G * ((1.0 + spsi) * P12[i] - 2.0 * (1.0 - spsi) * P3[i]) + 2.0 * lam * spsi * IOTA6[i];
                }
                double rden = 1.0 / den_l;
                for (int j = 0; j < 6; ++j)
                    for (int i = 0; i < 6; ++i) {
                        double Id = identv(i, j);
                        double q33 = P3[i] * P3[j];
                        double qaa = P12[i] * P12[j];
                        double qat = P12[i] * t[j], qta = t[i] * P12[j];
                        double qa3 = P12[i] * P3[j], q3a = P3[i] * P12[j];
                        double d3 = g3 * (M[i][j] - Id * a12 - (qat + qta) + a12 * qaa
                                          + (a12 - 2.0 * e3) * q33 + e3 * (qa3 + q3a));
                        D[i + 6 * j] = (s3 - s1) * d3 + lam * IOTA6[i] * IOTA6[j]
                                       + G * (qaa + 2.0 * q33) - Dq[i] * Dp[j] * rden;
                    }
            }
            return 2;
        }
    }

    if (g_sl > g_sr) { /* right edge, kernel.h:330-394 */
        double den_r = 4.0 * lam * sphi * spsi + 2.0 * G * (1.0 + sphi) * (1.0 + spsi) + G * (1.0 - sphi) * (1.0 - spsi);
        double lam_r = (G * (2.0 * (1.0 + sphi) * e1 - (1.0 - sphi) * (e2 + e3)) + 2.0 * lam * sphi * I1 - cb) / den_r;
        double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + spsi);
        if (lam_r >= g_sr && lam_r <= g_ra) {
            double g1 = 1.0 / ((e1 - e2) * (e1 - e3));
            double P1[6], P23[6];
            proj3(t2, t, e2 + e3, e2 * e3, g1, P1);
            for (int i = 0; i < 6; ++i) P23[i] = IOTA6[i] - P1[i];
            double s1 = lam * I1 + 2.0 * G * e1 - lam_r * (2.0 * lam * spsi + 2.0 * G * (1.0 + spsi));
            double s3 = lam * I1 + G * (e2 + e3) - lam_r * (2.0 * lam * spsi - G * (1.0 - spsi));
            for (int i = 0; i < 6; ++i) S[i] = s1 * P1[i] + s3 * P23[i];
            if (D) {
                double M[6][6];
                dsquare3(t, M);
                double a23 = e2 + e3;
                double Dp[6], Dq[6];
                for (int i = 0; i < 6; ++i) {
                    Dp[i] = G * (2.0 * (1.0 + sphi) * P1[i] - (1.0 - sphi) * P23[i]) + 2.0 * lam * sphi * IOTA6[i];
                    Dq[i] = G * (2.0 * (1.0 + spsi) * P1[i] - (1.0 - spsi) * P23[i]) + 2.0 * lam * spsi * IOTA6[i];
                }
                double rden = 1.0 / den_r;
                for (int j = 0; j < 6; ++j)
                    for (int i = 0; i < 6; ++i) {
                        double Id = identv(i, j);
                        double q11 = P1[i] * P1[j];
                        double qbb = P23[i] * P23[j];
                        double qbt = P23[i] * t[j], qtb = t[i] * P23[j];
                        double qb1 = P23[i] * P1[j], q1b = P1[i] * P23[j];
                        double d1 = g1 * (M[i][j] - Id * a23 - (qbt + qtb) + a23 * qbb
                                          + (a23 - 2.0 * e1) * q11 + e1 * (qb1 + q1b));
                        D[i + 6 * j] = (s1 - s3) * d1 + lam * IOTA6[i] * IOTA6[j]
                                       + G * (2.0 * q11 + qbb) - Dq[i] * Dp[j] * rden;
                    }
            }
            return 3;
        }
    }

    /* apex, kernel.h:399-409 */
    double sa = cb / (2.0 * sphi);
    for (int i = 0; i < 6; ++i) S[i] = sa * IOTA6[i];
    if (D) memset(D, 0, sizeof(double) * 36);
    return 4;
}

/* S (6 x n) and optional DS (36 x n), column-major per point, as the MEX
 * outputs (constitutive_problem_3D_SDS_mex.c:49-63).  `branch` (optional)
 * receives the branch code per point. */
void orc_mc3d(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
              const double *shear, const double *bulk, const double *lame,
              double *S, double *DS, int8_t *branch)
{
    (void)bulk;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        int b = mc3d_point(E + 6 * p, c_bar[p], sin_phi[p], shear[p], lame[p],
                           S + 6 * p, DS ? DS + 36 * p : NULL);
        if (branch) branch[p] = (int8_t)b;
    }
}

/* ------------------------------------------------------------------------ */
/* 2D (plane strain) Mohr-Coulomb return mapping                             */
/* ------------------------------------------------------------------------ */
static void outer3(const double a[3], const double b[3], double s, double M[9])
{
    for (int j = 0; j < 3; ++j)
        for (int i = 0; i < 3; ++i) M[i + 3 * j] = s * a[i] * b[j];
}

static int mc2d_point(const double eps[3], double cb, double sphi, double G,
                      double K, double lam, double S[4], double *D /* 9 or NULL */)
{
    /* trial strain, 4-component form with eps33 = 0 (kernel.h:42) */
    const double x = eps[0], y = eps[1], gm = eps[2], z = 0.0;
    const double tr = x + y;
    const double rad = sqrt((x - y) * (x - y) + gm * gm);
    const double lp = 0.5 * (tr + rad), lm = 0.5 * (tr - rad), lz = z;

    /* first derivatives of the unordered eigenvalues (kernel.h:52-66) */
    double dp[4] = {0, 0, 0, 0};
    if (rad != 0.0) {
        dp[0] = (x - lm) / rad;
        dp[1] = (y - lm) / rad;
        dp[2] = 0.5 * gm / rad;
    } else {
        dp[0] = 1.0;
        dp[1] = 1.0;
    }
    const double dm[4] = {1.0 - dp[0], 1.0 - dp[1], -dp[2], 0.0};
    const double dz[4] = {0.0, 0.0, 0.0, 1.0};

    /* second derivatives (kernel.h:69-96) */
    double Hp[9] = {0}, Hm[9] = {0}, Hz[9] = {0};
    if (rad != 0.0) {
        const double ir = 1.0 / rad;
        const double dg[3] = {1.0, 1.0, 0.5};
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i) {
                double v = (i == j ? dg[i] : 0.0) - dp[i] * dp[j] - dm[i] * dm[j];
                /* the reference writes (-a_i a_j - b_i b_j) with operands in
                 * the order of kernel.h:79-87; products commute exactly */
                Hp[i + 3 * j] = v * ir;
            }
        for (int k = 0; k < 9; ++k) Hm[k] = -Hp[k];
    }

    /* order eigenvalues e1 >= e2 >= e3 (kernel.h:98-130) */
    double e1 = lp, e2 = lm, e3 = lz;
    const double *v1 = dp, *v2 = dm, *v3 = dz;
    const double *H1 = Hp, *H2 = Hm, *H3 = Hz;
    if ((lp >= lz) && (lz > lm)) {
        e2 = lz; e3 = lm; v2 = dz; v3 = dm; H2 = Hz; H3 = Hm;
    }
    if (lz > lp) {
        e1 = lz; e2 = lp; e3 = lm; v1 = dz; v2 = dp; v3 = dm; H1 = Hz; H2 = Hp; H3 = Hm;
    }

    const double trE = e1 + e2 + e3;
    const double f_tr = 2.0 * G * ((1.0 + sphi) * e1 - (1.0 - sphi) * e3) + 2.0 * lam * sphi * trE - cb;
    const double g_sl = (e1 - e2) / (1.0 + sphi);
    const double g_sr = (e2 - e3) / (1.0 - sphi);
    const double g_la = (e1 + e2 - 2.0 * e3) / (3.0 - sphi);
    const double g_ra = (2.0 * e1 - e2 - e3) / (3.0 + sphi);
    const double den_s = 4.0 * lam * sphi * sphi + 2.0 * G * (1.0 + sphi) * (1.0 + sphi) + 2.0 * G * (1.0 - sphi) * (1.0 - sphi);
    const double den_l = 4.0 * lam * sphi * sphi + G * (1.0 + sphi) * (1.0 + sphi) + 2.0 * G * (1.0 - sphi) * (1.0 - sphi);
    const double den_r = 4.0 * lam * sphi * sphi + 2.0 * G * (1.0 + sphi) * (1.0 + sphi) + G * (1.0 - sphi) * (1.0 - sphi);
    const double lam_s = f_tr / den_s;
    const double lam_l = (G * ((1.0 + sphi) * (e1 + e2) - 2.0 * (1.0 - sphi) * e3) + 2.0 * lam * sphi * trE - cb) / den_l;
    const double lam_r = (G * (2.0 * (1.0 + sphi) * e1 - (1.0 - sphi) * (e2 + e3)) + 2.0 * lam * sphi * trE - cb) / den_r;
    (void)K;

    /* branch order of kernel.h:162-168 */
    int br;
    if (f_tr <= 0.0) br = 0;
    else if (lam_s <= ((g_sl < g_sr) ? g_sl : g_sr)) br = 1;
    else if ((g_sl < g_sr) && (lam_l >= g_sl) && (lam_l <= g_la)) br = 2;
    else if ((g_sl > g_sr) && (lam_r >= g_sr) && (lam_r <= g_ra)) br = 3;
    else br = 4;

    static const double VOL9[9] = {1, 1, 0, 1, 1, 0, 0, 0, 0};
    if (br == 0) {
        const double vol = x + y + z;
        S[0] = lam * vol + 2.0 * G * x;
        S[1] = lam * vol + 2.0 * G * y;
        S[2] = G * gm;
        S[3] = lam * vol + 2.0 * G * z;
        if (D) {
            memset(D, 0, 9 * sizeof(double));
            D[0] = lam + 2.0 * G; D[1] = lam; D[3] = lam; D[4] = lam + 2.0 * G; D[8] = G;
        }
        return 0;
    }
    if (br == 1) {
        const double s1 = lam * trE + 2.0 * G * e1 - lam_s * (2.0 * lam * sphi + 2.0 * G * (1.0 + sphi));
        const double s2 = lam * trE + 2.0 * G * e2 - lam_s * (2.0 * lam * sphi);
        const double s3 = lam * trE + 2.0 * G * e3 - lam_s * (2.0 * lam * sphi - 2.0 * G * (1.0 - sphi));
        for (int i = 0; i < 4; ++i) S[i] = s1 * v1[i] + s2 * v2[i] + s3 * v3[i];
        if (D) {
            double w[3], M1[9], M2[9], M3[9], M6[9];
            for (int i = 0; i < 3; ++i)
                w[i] = 2.0 * G * (1.0 + sphi) * v1[i] - 2.0 * G * (1.0 - sphi) * v3[i] + (i < 2 ? 2.0 * lam * sphi : 0.0);
            outer3(v1, v1, 2.0 * G, M1);
            outer3(v2, v2, 2.0 * G, M2);
            outer3(v3, v3, 2.0 * G, M3);
            outer3(w, w, 1.0 / den_s, M6);
            for (int k = 0; k < 9; ++k)
                D[k] = s1 * H1[k] + s2 * H2[k] + s3 * H3[k] + lam * VOL9[k] + M1[k] + M2[k] + M3[k] - M6[k];
        }
        return 1;
    }
    if (br == 2) {
        const double s1 = lam * trE + G * (e1 + e2) - lam_l * (2.0 * lam * sphi + G * (1.0 + sphi));
        const double s3 = lam * trE + 2.0 * G * e3 - lam_l * (2.0 * lam * sphi - 2.0 * G * (1.0 - sphi));
        for (int i = 0; i < 4; ++i) S[i] = s1 * (v1[i] + v2[i]) + s3 * v3[i];
        if (D) {
            double v12[3], w[3], H12[9], M1[9], M3[9], M6[9];
            for (int i = 0; i < 3; ++i) v12[i] = v1[i] + v2[i];
            for (int i = 0; i < 3; ++i)
                w[i] = G * (1.0 + sphi) * v12[i] - 2.0 * G * (1.0 - sphi) * v3[i] + (i < 2 ? 2.0 * lam * sphi : 0.0);
            for (int k = 0; k < 9; ++k) H12[k] = H1[k] + H2[k];
            outer3(v12, v12, G, M1);
            outer3(v3, v3, 2.0 * G, M3);
            outer3(w, w, 1.0 / den_l, M6);
            for (int k = 0; k < 9; ++k)
                D[k] = s1 * H12[k] + s3 * H3[k] + lam * VOL9[k] + M1[k] + M3[k] - M6[k];
        }
        return 2;
    }
    if (br == 3) {
        const double s1 = lam * trE + 2.0 * G * e1 - lam_r * (2.0 * lam * sphi + 2.0 * G * (1.0 + sphi));
        const double s3 = lam * trE + G * (e2 + e3) - lam_r * (2.0 * lam * sphi - G * (1.0 - sphi));
        for (int i = 0; i < 4; ++i) S[i] = s1 * v1[i] + s3 * (v2[i] + v3[i]);
        if (D) {
            double v23[3], w[3], H23[9], M1[9], M3[9], M6[9];
            for (int i = 0; i < 3; ++i) v23[i] = v2[i] + v3[i];
            for (int i = 0; i < 3; ++i)
                w[i] = 2.0 * G * (1.0 + sphi) * v1[i] - G * (1.0 - sphi) * v23[i] + (i < 2 ? 2.0 * lam * sphi : 0.0);
            for (int k = 0; k < 9; ++k) H23[k] = H2[k] + H3[k];
            outer3(v1, v1, 2.0 * G, M1);
            outer3(v23, v23, G, M3);
            outer3(w, w, 1.0 / den_r, M6);
            for (int k = 0; k < 9; ++k)
                D[k] = s1 * H1[k] + s3 * H23[k] + lam * VOL9[k] + M1[k] + M3[k] - M6[k];
        }
        return 3;
    }
    {
        const double sa = cb / (2.0 * sphi);
        S[0] = sa; S[1] = sa; S[2] = 0.0; S[3] = sa;
        if (D) memset(D, 0, 9 * sizeof(double));
    }
    return 4;
}

void orc_mc2d(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
              const double *shear, const double *bulk, const double *lame,
              double *S, double *DS, int8_t *branch)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) {
        int b = mc2d_point(E + 3 * p, c_bar[p], sin_phi[p], shear[p], bulk[p], lame[p],
                           S + 4 * p, DS ? DS + 9 * p : NULL);
        if (branch) branch[p] = (int8_t)b;
    }
}

/* ------------------------------------------------------------------------ */
/* Strain-displacement operator, matrix-free from DPhi (per point, n_p       */
/* derivatives contiguous).  DOF numbering is node-major: dof = dim*node+c.  */
/* ------------------------------------------------------------------------ */

/* B_q for one point in 3D: rows (e11,e22,e33,g12,g23,g13), column dof 3*a+c
 * (elastic_stiffness_matrix_3D.m:105-132, assemble_K_tangent_vals.c:110-132). */
static void bmat3(int n_p, const double *d1, const double *d2, const double *d3, double *B /* 6 x 3n_p col-major */)
{
    memset(B, 0, sizeof(double) * 6 * 3 * n_p);
    for (int a = 0; a < n_p; ++a) {
        double *cx = B + 6 * (3 * a), *cy = cx + 6, *cz = cy + 6;
        cx[0] = d1[a]; cx[3] = d2[a]; cx[5] = d3[a];
        cy[1] = d2[a]; cy[3] = d1[a]; cy[4] = d3[a];
        cz[2] = d3[a]; cz[4] = d2[a]; cz[5] = d1[a];
    }
}

/* 2D rows (e11,e22,g12) (elastic_stiffness_matrix_2D.m:90-104). */
static void bmat2(int n_p, const double *d1, const double *d2, double *B /* 3 x 2n_p */)
{
    memset(B, 0, sizeof(double) * 3 * 2 * n_p);
    for (int a = 0; a < n_p; ++a) {
        double *cx = B + 3 * (2 * a), *cy = cx + 3;
        cx[0] = d1[a]; cx[2] = d2[a];
        cy[1] = d2[a]; cy[2] = d1[a];
    }
}

/* Element-level tangent values (assemble_K_tangent_vals.c:85-171):
 *   V[smap(k,e)-1] += (sum_q B_q' (w_q D_q) B_q)_k, smap 1-based, 0 = skip.
 * Sequential element loop (deterministic).  V must be zeroed by the caller. */
static void assemble_generic(int dim, int64_t n_p, int64_t n_e, int n_q,
                             const double *D1, const double *D2, const double *D3,
                             const double *DS, const double *W, const int64_t *smap, double *V)
{
    const int ns = dim == 3 ? 6 : 3;
    const int nl = dim * (int)n_p;
    double B[6 * 3 * 20], DB[6 * 3 * 20], Dq[36], Ke[60 * 60];
    if (nl > 60) return;
    for (int64_t e = 0; e < n_e; ++e) {
        memset(Ke, 0, sizeof(double) * nl * nl);
        for (int q = 0; q < n_q; ++q) {
            const int64_t g = e * n_q + q;
            const double w = W[g];
            for (int k = 0; k < ns * ns; ++k) Dq[k] = w * DS[g * ns * ns + k];
            if (dim == 3) bmat3((int)n_p, D1 + g * n_p, D2 + g * n_p, D3 + g * n_p, B);
            else bmat2((int)n_p, D1 + g * n_p, D2 + g * n_p, B);
            for (int c = 0; c < nl; ++c)
                for (int r = 0; r < ns; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < ns; ++k) s += Dq[r + k * ns] * B[k + c * ns];
                    DB[r + c * ns] = s;
                }
            for (int c = 0; c < nl; ++c)
                for (int r = 0; r < nl; ++r) {
                    double s = 0.0;
                    for (int k = 0; k < ns; ++k) s += B[k + r * ns] * DB[k + c * ns];
                    Ke[r + c * nl] += s;
                }
        }
        const int64_t *m = smap + e * (int64_t)nl * nl;
        for (int k = 0; k < nl * nl; ++k)
            if (m[k] > 0) V[m[k] - 1] += Ke[k];
    }
}

void orc_assemble3d(int64_t n_p, int64_t n_e, int n_q, const double *DPhi1, const double *DPhi2,
                    const double *DPhi3, const double *DS, const double *WEIGHT,
                    const int64_t *smap, int64_t nnz, double *V)
{
    memset(V, 0, sizeof(double) * nnz);
    assemble_generic(3, n_p, n_e, n_q, DPhi1, DPhi2, DPhi3, DS, WEIGHT, smap, V);
}

void orc_assemble2d(int64_t n_p, int64_t n_e, int n_q, const double *DPhi1, const double *DPhi2,
                    const double *DS, const double *WEIGHT, const int64_t *smap, int64_t nnz, double *V)
{
    memset(V, 0, sizeof(double) * nnz);
    assemble_generic(2, n_p, n_e, n_q, DPhi1, DPhi2, NULL, DS, WEIGHT, smap, V);
}

/* E = B*U (CONSTITUTIVE.m:218): elem is n_p x n_e, 0-based node ids,
 * U is dim x n_n column-major, E is ns x (n_e*n_q). */
void orc_strain(int dim, int64_t n_p, int64_t n_e, int n_q, const int64_t *elem,
                const double *D1, const double *D2, const double *D3,
                const double *U, double *E)
{
    const int ns = dim == 3 ? 6 : 3;
#pragma omp parallel for schedule(static)
    for (int64_t e = 0; e < n_e; ++e) {
        double B[6 * 3 * 20];
        for (int q = 0; q < n_q; ++q) {
            const int64_t g = e * n_q + q;
            if (dim == 3) bmat3((int)n_p, D1 + g * n_p, D2 + g * n_p, D3 + g * n_p, B);
            else bmat2((int)n_p, D1 + g * n_p, D2 + g * n_p, B);
            for (int r = 0; r < ns; ++r) {
                double s = 0.0;
                for (int a = 0; a < n_p; ++a)
                    for (int c = 0; c < dim; ++c)
                        s += B[r + (dim * a + c) * ns] * U[dim * elem[e * n_p + a] + c];
                E[g * ns + r] = s;
            }
        }
    }
}

/* F = B' * (WEIGHT .* S(1:ns,:)) (CONSTITUTIVE.m:281).  S has s_rows rows
 * (6 in 3D, 4 in 2D where row 4 = sigma33 is dropped).  F (dim x n_n) is
 * zeroed here; sequential, deterministic. */
void orc_force(int dim, int64_t n_p, int64_t n_e, int n_q, int64_t n_n, const int64_t *elem,
               const double *D1, const double *D2, const double *D3,
               const double *S, int s_rows, const double *W, double *F)
{
    const int ns = dim == 3 ? 6 : 3;
    memset(F, 0, sizeof(double) * dim * n_n);
    double B[6 * 3 * 20];
    for (int64_t e = 0; e < n_e; ++e) {
        for (int q = 0; q < n_q; ++q) {
            const int64_t g = e * n_q + q;
            if (dim == 3) bmat3((int)n_p, D1 + g * n_p, D2 + g * n_p, D3 + g * n_p, B);
            else bmat2((int)n_p, D1 + g * n_p, D2 + g * n_p, B);
            for (int a = 0; a < n_p; ++a)
                for (int c = 0; c < dim; ++c) {
                    double s = 0.0;
                    for (int r = 0; r < ns; ++r) s += B[r + (dim * a + c) * ns] * (W[g] * S[g * s_rows + r]);
                    F[dim * elem[e * n_p + a] + c] += s;
                }
        }
    }
}

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
