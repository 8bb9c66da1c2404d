/*
 * ssr_hip.h -- C ABI of libssr_hip.so, the MI355X (gfx950) implementation of
 * the SSR Newton-step hot path of Beremi/slope_stability_octave.
 *
 * Conventions
 *  - Every function returns 0 on success and a nonzero status otherwise
 *    (1 bad argument, 2 HIP runtime error, 3 allocation failure, 4 bad state);
 *    ssr_last_error() returns the message of the last failure in this thread.
 *    A MEX shim turns a nonzero status into mexErrMsgIdAndTxt with that text.
 *  - Arrays are double precision, Octave column-major.  Per-point blocks are
 *    contiguous ("point-major"): E is n_strain x n_int, S is 6 x n_int (4 in
 *    2D), DS is 36 x n_int (9 in 2D), each block column-major.
 *  - ssr_* functions without "_dev_" take HOST pointers, copy in, compute on
 *    the GPU and copy out -- the MEX boundary semantics.  ssr_dev_* functions
 *    take DEVICE pointers and enqueue on the library stream (ssr_set_stream),
 *    returning before the GPU finishes.
 *  - Sizes are int64_t; a size of 0 is a valid no-op.
 */
#ifndef SSR_HIP_H
#define SSR_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- runtime ------------------------------------------------------------ */
const char *ssr_last_error(void);
int ssr_version(void);
int ssr_device_count(int *count);
int ssr_set_stream(void *hip_stream); /* NULL = default stream */
void *ssr_get_stream(void);
int ssr_dmalloc(void **ptr, size_t bytes);
int ssr_dfree(void *ptr);
int ssr_htod(void *dst, const void *src, size_t bytes); /* synchronous */
int ssr_dtoh(void *dst, const void *src, size_t bytes); /* synchronous */
int ssr_dtod(void *dst, const void *src, size_t bytes); /* enqueued */
int ssr_dzero(void *dst, size_t bytes);                 /* enqueued */
int ssr_synchronize(void);
int ssr_event_create(void **ev);
int ssr_event_destroy(void *ev);
int ssr_event_record(void *ev); /* on the library stream */
int ssr_event_elapsed_ms(void *start, void *stop, float *ms);

/* ---- constitutive problem (Mohr-Coulomb return mapping) ----------------- */

/* [S, DS] = constitutive_problem_3D_SDS_mex(E, c_bar, sin_phi, shear, bulk, lame)
 * replaces slope_stability/+CONSTITUTIVE_PROBLEM/mex/constitutive_problem_3D_SDS_mex.c:29-64
 * (per-point kernel constitutive_3D_kernel.h:101-410). */
int ssr_constitutive_problem_3D_SDS(int64_t n_int, const double *E, const double *c_bar,
                                    const double *sin_phi, const double *shear, const double *bulk,
                                    const double *lame, double *S, double *DS);

/* S = constitutive_problem_3D_S_mex(...)
 * replaces +CONSTITUTIVE_PROBLEM/mex/constitutive_problem_3D_S_mex.c:27-58. */
int ssr_constitutive_problem_3D_S(int64_t n_int, const double *E, const double *c_bar,
                                  const double *sin_phi, const double *shear, const double *bulk,
                                  const double *lame, double *S);

/* [S, DS] = constitutive_problem_2D_SDS_mex(E, ...), E 3 x n_int, S 4 x n_int, DS 9 x n_int
 * replaces +CONSTITUTIVE_PROBLEM/mex/constitutive_problem_2D_SDS_mex.c:29-64
 * (constitutive_2D_kernel.h:36-369).  rows_E must be 3 (check of :45). */
int ssr_constitutive_problem_2D_SDS(int64_t n_int, int64_t rows_E, const double *E,
                                    const double *c_bar, const double *sin_phi, const double *shear,
                                    const double *bulk, const double *lame, double *S, double *DS);

/* S = constitutive_problem_2D_S_mex(E, ...)
 * replaces +CONSTITUTIVE_PROBLEM/mex/constitutive_problem_2D_S_mex.c:27-59 (check of :43). */
int ssr_constitutive_problem_2D_S(int64_t n_int, int64_t rows_E, const double *E,
                                  const double *c_bar, const double *sin_phi, const double *shear,
                                  const double *bulk, const double *lame, double *S);

/* Device-resident forms of the four calls above (same layouts; DS == NULL
 * gives the stress-only variant; branch, if not NULL, receives per point
 * 0 elastic, 1 smooth, 2 left edge, 3 right edge, 4 apex). */
int ssr_dev_constitutive_3D(int64_t n_int, const double *E, const double *c_bar,
                            const double *sin_phi, const double *shear, const double *bulk,
                            const double *lame, double *S, double *DS, int8_t *branch);
int ssr_dev_constitutive_2D(int64_t n_int, const double *E, const double *c_bar,
                            const double *sin_phi, const double *shear, const double *bulk,
                            const double *lame, double *S, double *DS, int8_t *branch);

/* [c_bar, sin_phi] = reduction(c0, phi, psi, lambda, Davis_type) per point
 * replaces +CONSTITUTIVE_PROBLEM/reduction.m:25-58 (called by CONSTITUTIVE.m:197).
 * davis is 'A', 'B' or 'C'; lambda is a scalar (lambda_per_point = 0) or one
 * value per point.  davis_c_all carries Octave's vector `if` of reduction.m:47
 * (1 when phi1 > psi at every point). */
int ssr_dev_reduction(int64_t n_int, int davis, const double *c0, const double *phi,
                      const double *psi, const double *lambda, int lambda_per_point,
                      int davis_c_all, double *c_bar, double *sin_phi);

#ifdef __cplusplus
}
#endif

#endif /* SSR_HIP_H */
