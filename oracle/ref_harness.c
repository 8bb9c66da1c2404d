/*
 * ref_harness.c -- loop wrapper around the reference's own per-point kernels.
 *
 * TEST INFRASTRUCTURE ONLY (see mc_oracle.c).  Built by oracle/Makefile in the
 * development container only, from the headers where they lie under the
 * reference checkout (constitutive_3D_kernel.h, constitutive_2D_kernel.h);
 * output goes to oracle/_ref/ and is never committed.  The loops mirror the
 * MEX wrappers (constitutive_problem_3D_SDS_mex.c:56-63,
 * constitutive_problem_2D_SDS_mex.c:56-63): OpenMP static schedule over points.
 * Passing DS == NULL gives the stress-only path (constitutive_problem_*_S_mex.c).
 */
#include <stdint.h>
#include "constitutive_3D_kernel.h"
#include "constitutive_2D_kernel.h"

void ref_mc3d(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
              const double *shear, const double *bulk, const double *lame,
              double *S, double *DS)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p)
        constitutive_3D_point(E + 6 * p, c_bar[p], sin_phi[p], shear[p], bulk[p], lame[p],
                              S + 6 * p, DS ? DS + 36 * p : 0);
}

void ref_mc2d(int64_t n, const double *E, const double *c_bar, const double *sin_phi,
              const double *shear, const double *bulk, const double *lame,
              double *S, double *DS)
{
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p)
        constitutive_2D_point(E + 3 * p, c_bar[p], sin_phi[p], shear[p], bulk[p], lame[p],
                              S + 4 * p, DS ? DS + 9 * p : 0);
}
