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
/* Per host thread: launch this thread's work on its own stream (NULL: back to
 * the library stream).  Independent problems driven from several threads then
 * run concurrently on the device (no reference counterpart: the reference is
 * single-threaded Octave). */
int ssr_set_thread_stream(void *hip_stream);
int ssr_stream_create(void **out); /* default priority, non-blocking */
int ssr_stream_destroy(void *hip_stream);
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
/* The same with the five material fields per material (n_mat entries each): point p takes
 * those of material mat_id[p / n_q], its element's.  With c_bar and sin_phi reduced per
 * material (ssr_dev_reduction on the material table) this is CONSTITUTIVE.build_F_all's
 * reduction (CONSTITUTIVE.m:318,332,472) folded into the return mapping's material loads
 * for a scalar lambda: the same values at every point, so the same S and DS. */
int ssr_dev_constitutive_mat(int dim, int64_t n_int, int n_q, const double *E, const int32_t *mat_id,
                             int64_t n_mat, const double *c_bar, const double *sin_phi, const double *shear,
                             const double *bulk, const double *lame, double *S, double *DS, int8_t *branch);

/* [c_bar, sin_phi] = reduction(c0, phi, psi, lambda, Davis_type) per point
 * replaces +CONSTITUTIVE_PROBLEM/reduction.m:25-58 (called by CONSTITUTIVE.m:197).
 * davis is 'A', 'B' or 'C'; lambda is a scalar (lambda_per_point = 0) or one
 * value per point.  davis_c_all carries Octave's vector `if` of reduction.m:47
 * (1 when phi1 > psi at every point). */
int ssr_dev_reduction(int64_t n_int, int davis, const double *c0, const double *phi,
                      const double *psi, const double *lambda, int lambda_per_point,
                      int davis_c_all, double *c_bar, double *sin_phi);
/* Psi per point of the Mohr-Coulomb model (potential_3D.m / potential_2D.m),
 * E point-major with 6 (3D) or 3 (2D) rows; the integral WEIGHT*Psi' is
 * CONSTITUTIVE.potential_energy (CONSTITUTIVE.m:363-387). */
int ssr_dev_potential(int dim, int64_t n_int, const double *E, const double *c_bar, const double *sin_phi,
                      const double *shear, const double *bulk, const double *lame, double *Psi);

/* ---- tuning ------------------------------------------------------------- */
/* Kernel forms and launch shapes, all parity-identical (no reference
 * counterpart; autotune.py picks them on the caller's data):
 *  key 0: 3D constitutive kernel form (11 = default: register inputs, LDS-staged
 *         S and DS, non-temporal stores; the other forms are listed in
 *         constitutive.hip);
 *  key 1: lanes per block row of the node-block SpMV (0 = by dimension,
 *         4, 8, 16, 32; 64 = one wavefront per row);
 *  key 2: wavefronts per compute unit of persistent grids (0 = automatic);
 *  key 3: 0 only (eager V-cycle launches); replay from a captured HIP graph is
 *         disabled and any other value is rejected;
 *  key 4: 1 = node-block Jacobi scaling in the finest-level Chebyshev smoother
 *         (default), 0 = scalar;
 *  key 5: assembly and internal-force form (0 = element contributions, then
 *         a deterministic per-block / per-node gather, default; 1 = FP64 atomics);
 *  key 6: lanes per block row of the multi-vector SpMV (0 = by dimension);
 *  key 7: blocks per lane per pass of the node-block SpMV (0 = by dimension:
 *         2 in 3D with 16 lanes per row, 1 in 2D; or 1, 2);
 *  key 8: element tangent kernel (0 = two pairs of one column node per lane,
 *         sharing D*B_b, default (2D as 2);
 *         1 = D*B shared through LDS per point; 2 = per-lane products from
 *         LDS-staged point data);
 *  key 9: coarse AMG operators, prolongators and restrictions as block CSR
 *         over whole nodes / aggregates (1, default) or scalar CSR (0);
 *         applies to hierarchies built after the call;
 *  key 10: AMG hierarchy built on the device (1, default; the host build is
 *         used when an aggregate loses a near-null vector) or on the host (0);
 *  key 11: node-block SpMV with its block values staged through LDS by
 *         16-byte loads (1; for the default lane shapes) or read by each
 *         lane (0, default);
 *  key 12: lane-group sums of the multi-vector SpMV and of the block-CSR
 *         SpMV (AMG coarse levels and transfers) by halving exchanges (1,
 *         default) or by a full butterfly per value (0): the same bits;
 *  key 13: finest-level Chebyshev block step with one thread per unknown (1,
 *         default) or per node (0): the same bits;
 *  key 14: node-block SpMV of a solver's symmetric operator (and of the AMG
 *         finest level it hands over) reading only the blocks on and above
 *         the block diagonal, the others as transposed mirrors (1) or every
 *         block (0, default): the same bits for an operator whose mirrored
 *         blocks are transposes bit for bit, as K_r is (slower on the
 *         structured slope meshes, whose mirrors lie far apart);
 *  key 15: 2D constitutive kernel form with DS (1 = one wavefront per block,
 *         strains read and S, DS written through LDS as contiguous 16-byte
 *         spans, default when E, S and DS are 16-byte aligned; 0 = one lane
 *         per point with direct accesses, also the form for S alone): the
 *         same bits;
 *  key 16: the native Newton loops (ssr_newton_ind_SSR, ssr_newton,
 *         ssr_newton_ind_LL) replay their element phases -- strain, stress
 *         and tangent, internal force of each build_F, and the K_r
 *         assembly -- from HIP graphs captured on their second use within
 *         one Newton solve (1, default), or launch them eagerly (0): the
 *         same bits;
 *  key 17: the Krylov multi-dot V'u issuing each pass's loads (clamped
 *         indices) before its products, and the double update u += V h
 *         reading h directly instead of through LDS (1, default), or the
 *         guarded, staged forms (0): the same bits;
 *  key 18: the float block-CSR products of the AMG transfers (3 x 6 and
 *         6 x 3 blocks) loading a lane's next block with its current one (1,
 *         default) or in turn (0): the same bits. */
int ssr_set_tuning(int key, int value);
int ssr_get_tuning(int key, int *value);

/* ---- tangent assembly ---------------------------------------------------- */

/* V_tang = assemble_K_tangent_vals(DPhi1, DPhi2, DPhi3, DS, WEIGHT, scatter_map, n_q, nnz_out)
 * replaces slope_stability/+ASSEMBLY/mex/assemble_K_tangent_vals.c:38-172 (dim = 3,
 * P2 tetrahedra, 11 points) and assemble_K_tangent_vals_2D.c:31-165 (dim = 2,
 * DPhi3 = NULL, P2 triangles, 7 points).  DPhi* are n_p x n_int, DS is ns^2 x n_int,
 * scatter_map is int64 (dim n_p)^2 x n_e, 1-based, 0 = skip.  V is zero-filled. */
int ssr_assemble_K_tangent_vals(int dim, int64_t n_p, int64_t n_int, int64_t n_e, int n_q,
                                const double *DPhi1, const double *DPhi2, const double *DPhi3,
                                const double *DS, const double *WEIGHT, const int64_t *scatter_map,
                                int64_t nnz_out, double *V);
int ssr_dev_assemble_K_tangent_vals(int dim, int64_t n_p, int64_t n_int, int64_t n_e, int n_q,
                                    const double *DPhi1, const double *DPhi2, const double *DPhi3,
                                    const double *DS, const double *WEIGHT, const int64_t *smap,
                                    int64_t nnz, double *V);

/* Fixed sparsity pattern: replaces CONSTITUTIVE.init_K_r_pattern (CONSTITUTIVE.m:393-464)
 * and build_scatter_map (:582-635).  Node-block CSR over all nodes with dim x dim
 * blocks; Dirichlet DOFs (Q == 0) masked.  elem: n_p x n_e (0-based), Q: dim x n_n. */
typedef struct ssr_pattern ssr_pattern;
int ssr_pattern_create(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const int64_t *elem,
                       const uint8_t *Q, int upload_device, ssr_pattern **out);
int ssr_pattern_destroy(ssr_pattern *p);
/* info[8] = dim, n_n, n_p, n_e, n_blocks, n_free, nnz_free, on_device */
int ssr_pattern_info(const ssr_pattern *p, int64_t *info);
/* host tables: 0 row_ptr, 1 col, 2 element block map, 3 transpose block, 4 diagonal
 * block, 5 reference-order positions (int64), 6 free mask (uint8) */
int ssr_pattern_get(const ssr_pattern *p, int which, void *dst);
/* device tables: 0 row_ptr, 1 col, 2 emap, 3 tblk, 4 diag, 5 elem, 6 ref_pos, 7 mask, 8 brow,
 * 9 upper blocks, 10/11 slot-list bounds and element pair slots, 12/13 per-block contribution
 * lists, 14/15 per-node element incidences (NULL when the table was not built) */
int ssr_pattern_dev_ptr(const ssr_pattern *p, int which, void **out);
/* The same pattern built on the device (CONSTITUTIVE.init_K_r_pattern, CONSTITUTIVE.m:393-464,
 * and build_scatter_map, :582-635): every table equal to ssr_pattern_create's.  DEVICE
 * pointers: elem int32 n_p x n_e (0-based), Q uint8 dim x n_n (nonzero = free). */
int ssr_pattern_create_dev(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const int32_t *elem,
                           const uint8_t *Q, ssr_pattern **out);

/* ---- FE setup on the device ---------------------------------------------- */
/* DPhi1..3 (n_p x n_int point-major) and WEIGHT (1 x n_int) of P2 elements from
 * coord (dim x n_n node-major) and elem (int32 n_p x n_e): elastic_stiffness_matrix_3D.m:60-96,148
 * (2D: elastic_stiffness_matrix_2D.m:50-80,118) with the P2 basis and quadrature of
 * local_basis_volume_{3D,2D}.m and quadrature_volume_{3D,2D}.m.  DPhi3 unused in 2D. */
int ssr_dev_fe_gradients(int dim, int64_t n_n, int64_t n_p, int64_t n_e, const double *coord,
                         const int32_t *elem, double *DPhi1, double *DPhi2, double *DPhi3, double *WEIGHT);
/* f_V (dim x n_n node-major) = vector_volume_{3D,2D}.m: f_int is f_rows x n_int point-major with
 * f_rows = dim, or f_rows = 1 for a density acting along -y (f_V_int = [0; -gamma(; 0)]). */
int ssr_dev_volume_force(const ssr_pattern *P, int n_q, const double *WEIGHT, const double *f_int,
                         int f_rows, double *f);
/* Per-point fields from a per-element material id (heterogenous_materials.m:35-87):
 * out[k][p] = table[mat_id[p / n_q] * n_fields + k]; out is a HOST array of n_fields DEVICE
 * pointers (NULL entries skipped). */
int ssr_dev_materials(int64_t n_e, int n_q, int n_mat, int n_fields, const double *table,
                      const int32_t *mat_id, double *const *out);

/* sum_e sum_q B'(w D)B into the block values (zeroed first); ds_layout 0 = MEX
 * layout (ns^2 per point), 1 = packed symmetric (ns(ns+1)/2 per point).
 * Replaces CONSTITUTIVE.build_K_tangent_QQ_vals (CONSTITUTIVE.m:487-501). */
int ssr_dev_assemble_bsr(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                         const double *DPhi3, const double *DS, int ds_layout, const double *WEIGHT,
                         double *Kvals);
/* E = B*U (CONSTITUTIVE.m:218,251); U node-major dim x n_n, E ns x n_int point-major. */
int ssr_dev_strain(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                   const double *DPhi3, const double *U, double *E);
/* F = B'*(WEIGHT.*S(1:ns,:)) (CONSTITUTIVE.m:268-285); S has s_rows rows per point. */
int ssr_dev_force(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                  const double *DPhi3, const double *S, int s_rows, const double *WEIGHT, double *F);
/* ssr_dev_strain and ssr_dev_force with each integration point's gradients
 * computed from the node coordinates (dim x n_n node-major, device) instead of
 * read from DPhi: the arithmetic of ssr_dev_fe_gradients, so with DPhi from
 * ssr_dev_fe_gradients on the same coordinates the results are the same bits,
 * and the 3 x n_p doubles of gradients per point are not read from memory.
 * The same products as CONSTITUTIVE.m:218,251 and :268-285 (no counterpart
 * entry in the reference, which stores B). */
int ssr_dev_strain_geo(const ssr_pattern *P, int n_q, const double *coord, const double *U, double *E);
int ssr_dev_force_geo(const ssr_pattern *P, int n_q, const double *coord, const double *S, int s_rows,
                      const double *WEIGHT, double *F);
/* K_r = r K_el + (1-r) K_t, (K_r + K_r')/2, Dirichlet masked (CONSTITUTIVE.m:515-532). */
int ssr_dev_regularize(const ssr_pattern *P, double r, const double *Kel, const double *Kt, double *Kr);
/* K_r straight from DS: the tangent assembly and the regularisation of
 * CONSTITUTIVE.build_K_r_QQ_vals (CONSTITUTIVE.m:487-532) in one call, K_t never
 * stored; equal bit for bit to ssr_dev_assemble_bsr + ssr_dev_regularize.
 * K_elast is the regularised elastic operator (as ssr_dev_regularize with r = 1
 * leaves it), K_r must not alias it. */
int ssr_dev_assemble_K_r(const ssr_pattern *P, int n_q, const double *DPhi1, const double *DPhi2,
                         const double *DPhi3, const double *DS, int ds_layout, const double *WEIGHT, double r,
                         const double *K_elast, double *K_r);
/* block values -> V in the reference's (ref_I, ref_J) order (CONSTITUTIVE.m:445). */
int ssr_dev_extract_ref(const ssr_pattern *P, const double *Kvals, double *Vref);
/* DS_elast = 2 DEV shear + VOL bulk per point (CONSTITUTIVE.m:413-422), MEX layout. */
int ssr_dev_ds_elastic(int dim, int64_t n_int, const double *shear, const double *bulk, double *DS);

/* ---- sparse / dense kernels of the Krylov path --------------------------- */
/* y = alpha K x + beta y on the node-block matrix (the librsb SpMV of newton_ind_SSR.m:91) */
int ssr_dev_spmv_bsr(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col,
                     const double *val, const double *x, double *y, double alpha, double beta);
/* The same product on float block values and vectors (the operator copy of
 * the float Krylov iterations; sums in float), and the double-to-float copy
 * that produces such an operator: no reference counterpart (the reference
 * runs in double only). */
int ssr_dev_spmv_bsr_f32(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col,
                         const float *val, const float *x, float *y, double alpha, double beta);
int ssr_dev_convert_f64_f32(int64_t n, const double *x, float *y);
/* The same products for a symmetric operator (librsb's symmetric SpMV behind
 * the same call sites): diag = the diagonal block of each row, tblk = the
 * transposed block of each block (pattern tables 4 and 3).  With tuning key
 * 14 set, blocks left of the diagonal are read as the transposes of their
 * mirrors, so only the blocks on and above the diagonal stream from memory;
 * for values whose mirrored blocks are transposes bit for bit the result is
 * the same bits as ssr_dev_spmv_bsr. */
int ssr_dev_spmv_bsr_sym(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col,
                         const int32_t *diag, const int32_t *tblk, const double *val, const double *x,
                         double *y, double alpha, double beta);
int ssr_dev_spmv_bsr_sym_f32(int dim, int64_t n_rows, const int32_t *row_ptr, const int32_t *col,
                             const int32_t *diag, const int32_t *tblk, const float *val, const float *x,
                             float *y, double alpha, double beta);
int ssr_dev_dot(int64_t n, const double *a, const double *b, const double *mask, double *result);
int ssr_dot(int64_t n, const double *a, const double *b, const double *mask, double *host_result);
/* max over nodes of the Euclidean norm of U's node vector (dim x n_n node-major), to the
 * host: the U_max history of SSR_indirect_continuation.m:92-93,182 (max(sqrt(sum(U.^2, 1)))). */
int ssr_node_norm_max(int dim, int64_t n_n, const double *U, double *host_result);
int ssr_dev_multidot(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h);
/* ssr_dev_multidot without any allocation: no counterpart in the reference.
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  The partial sums live in a static device
 * array of 8192 x 128 doubles (8 MiB, held per device until the process
 * ends); k above 128 is rejected.  Calls that overlap on one device share the
 * array and race. */
int ssr_dev_multidot_noalloc(int64_t n, int k, const double *V, int64_t ld, const double *u, double *h);
int ssr_dev_gemv_update(int64_t n, int k, const double *V, int64_t ld, const double *h, double alpha,
                        double beta, double *u);
/* float basis and vector, double coefficients and sums: the forms the precision-32 Krylov
 * iterations run (same call sites). */
int ssr_dev_multidot_f32(int64_t n, int k, const float *V, int64_t ld, const float *u, double *h);
int ssr_dev_gemv_update_f32(int64_t n, int k, const float *V, int64_t ld, const double *h, double alpha,
                            double beta, float *u);
int ssr_dev_axpby(int64_t n, double a, const double *x, double b, double *y);
int ssr_dev_lincomb(int64_t n, double a, const double *x, double b, const double *y, double *z);
int ssr_dev_mul(int64_t n, const double *m, double *x);
/* x = v (write-only fill; clears buffers without a memset node in captured graphs) */
int ssr_dev_fill(int64_t n, double v, double *x);
int ssr_dev_block_jacobi_setup(int dim, int64_t n_n, const int32_t *diag, const double *val, double *inv);
int ssr_dev_block_jacobi_apply(int dim, int64_t n_n, const double *inv, const double *x,
                               const double *mask, double *y);
int ssr_dev_csr_spmv(int64_t n_rows, const int64_t *rp, const int32_t *ci, const double *v,
                     const double *x, double *y, double alpha, double beta);
int ssr_dev_ewmul(int64_t n, const double *a, const double *b, double *z);
int ssr_dev_cheb_step(int64_t n, double c2, double c3, const double *dinv, const double *r, double *d,
                      double *x);
int ssr_dev_dense_gemv(int n, const double *M, const double *x, double *y);

/* ---- deflated Krylov solvers ---------------------------------------------- */
/* Replaces the DFGMRES class (+LINEAR_SOLVERS/DFGMRES.m:77-221) with dfgmres_solver.m,
 * dcg_solver.m and A_orthogonalize.m.  precond_type 0 = node-block Jacobi;
 * ssr_solver_set_amg installs smoothed-aggregation AMG (the BoomerAMG role of
 * hypre_boomeramg_mex.cpp: setup_ijv / update_values / apply). */
typedef struct ssr_solver ssr_solver;
int ssr_solver_create(const ssr_pattern *P, int maxit, double tol, double defl_tol, int precond_type,
                      ssr_solver **out);
int ssr_solver_destroy(ssr_solver *s);
/* operator values (device, borrowed); setup_precond = 1 rebuilds/refreshes M */
int ssr_solver_set_matrix(ssr_solver *s, const double *Kvals, int setup_precond);
/* Arithmetic of the Krylov iterations: 64 = double throughout (the reference's
 * FP64 DFGMRES); 32 = float Arnoldi, preconditioner and deflation projector on
 * float copies of K and W, with x0, the true residual and x += Z y in double
 * and repeated float solves on the double residual until the relative
 * residual meets the tolerance.  No reference counterpart (the reference runs
 * FP64 only); call before set_matrix. */
int ssr_solver_set_precision(ssr_solver *s, int bits);
/* The operator is symmetric (as K_r is, CONSTITUTIVE.m:515-532): the
 * deflation projector x - W W'A x then takes W'A x as (A W)' x with the A W
 * left by ssr_solver_A_orthogonalize, one SpMV less per Krylov iteration.
 * No reference counterpart (dfgmres_solver.m:59-64 applies A each time). */
int ssr_solver_set_symmetric(ssr_solver *s, int symmetric);
int ssr_solver_get_precision(const ssr_solver *s, int *bits);
/* near_null: n_null columns over all dim*n_n DOFs (host, column-major);
 * opts: NULL or double[8] {theta, max_levels, max_coarse, cheb_degree, cheb_lower,
 * cheb_upper, refresh_every, rebuild_every} */
int ssr_solver_set_amg(ssr_solver *s, const double *near_null, int n_null, const double *opts);
int ssr_solver_deflation_size(const ssr_solver *s, int *k);
int ssr_solver_clear_deflation(ssr_solver *s);
int ssr_solver_expand_deflation(ssr_solver *s, const double *v);       /* DFGMRES.m:163-170 */
int ssr_solver_copy_deflation(ssr_solver *dst, const ssr_solver *src); /* DFGMRES.m:204-221 */
/* DFGMRES.copy (DFGMRES.m:204-221): a new solver with src's settings
 * (tolerances, precision, symmetry) and a copy of its deflation basis; with
 * share_precond the copy applies src's preconditioner object (one AMG
 * hierarchy for all copies), else it has a node-block Jacobi of its own. */
int ssr_solver_clone(const ssr_solver *src, int share_precond, ssr_solver **out);
/* dst uses src's preconditioner object (shared; freed with the last user).
 * No reference counterpart: DFGMRES.m:204-221 gives a copy its own
 * preconditioner, set up again on first use.  Sharing keeps one AMG hierarchy
 * across the per-step copies of the continuation (the finest level always
 * applies the current operator).  Calls on dst and src must not overlap. */
int ssr_solver_share_preconditioner(ssr_solver *dst, const ssr_solver *src);
/* The next setup_precond of s rebuilds its preconditioner in full (AMG: new
 * aggregates, prolongators and Galerkin levels).  No reference counterpart:
 * the reference rebuilds on every update (hypre_boomeramg_mex.cpp:1151-1211). */
int ssr_solver_invalidate_preconditioner(ssr_solver *s);
int ssr_solver_get_deflation(const ssr_solver *s, int c, double *out);
int ssr_solver_A_orthogonalize(ssr_solver *s);                         /* A_orthogonalize.m */
int ssr_solver_dfgmres(ssr_solver *s, const double *b, double *x, int *iters, double *relres);
int ssr_solver_dcg(ssr_solver *s, const double *b, double *x, int *iters, double *relres);
int ssr_solver_precond_apply(ssr_solver *s, const double *x, double *y);
/* out[4] = total iterations, solves, preconditioner applications, SpMVs */
int ssr_solver_stats(const ssr_solver *s, int64_t *out);

/* ---- AMG instances from COO triplets (the BoomerAMG binding's commands) ---- */
/* info = hypre_boomeramg_mex('setup_ijv', I, J, V, n, Z, opts, instance_id)
 * replaces +LINEAR_SOLVERS/mex/hypre_boomeramg_mex.cpp:1013-1149 (called by
 * hypre_boomeramg_setup_ijv.m:23) with this library's smoothed-aggregation AMG.
 * I, J: 1-based indices as doubles, V values, nnz triplets of an n x n
 * symmetric matrix (host); Z: n x n_null near-null space, column-major
 * (NULL with n_null = 0: one constant per function); dofs_per_node: 2 or 3,
 * consecutive unknowns of one node (num_functions; an n that is not a
 * multiple is padded with decoupled unknowns); opts: NULL or double[8]
 * as ssr_solver_set_amg.  An existing id with the same n and nnz takes the
 * values-only path (:1054-1104) unless force_rebuild.  info (NULL or 5
 * doubles): n, nnz, num_functions, near-null columns, setup seconds. */
int ssr_amg_setup_ijv(const char *id, const double *I, const double *J, const double *V, int64_t nnz, int64_t n,
                      const double *Z, int n_null, int dofs_per_node, const double *opts, int force_rebuild,
                      double *info);
/* t = hypre_boomeramg_mex('update_values', V, id): values in the triplet order
 * of setup_ijv, hierarchy rebuilt; replaces hypre_boomeramg_mex.cpp:1151-1211. */
int ssr_amg_update_values(const char *id, const double *V, int64_t nnz, double *setup_seconds);
/* y = hypre_boomeramg_mex('apply', r, id): one V-cycle from x = 0 on host
 * vectors of length n; replaces hypre_boomeramg_mex.cpp:751-796.  The _dev
 * form takes device vectors. */
int ssr_amg_apply(const char *id, const double *r, int64_t n, double *y);
int ssr_amg_apply_dev(const char *id, const double *r, double *y);
/* hypre_boomeramg_mex('clear'[, id]): id NULL clears every instance. */
int ssr_amg_clear(const char *id);
/* hypre_boomeramg_mex('info'): number of instances and their total size. */
int ssr_amg_info(int64_t *count, int64_t *n_total);

/* ---- Newton step of the indirect SSR continuation ---------------------------- */
/* The device state of the CONSTITUTIVE handle class (CONSTITUTIVE.m) that the
 * Newton loop touches.  All pointers are device memory owned by the caller
 * except phi_host/psi_host (host, read only for Davis C): n_friction entries
 * each -- the (phi, psi) of the distinct materials -- or n_int entries when
 * n_friction is 0. */
typedef struct ssr_constitutive {
    int dim;     /* 2 or 3 */
    int n_q;     /* integration points per element; n_int = n_e * n_q */
    int davis;   /* 'A', 'B' or 'C' (reduction.m) */
    int s_rows;  /* rows of S per point: 6 in 3D, 4 in 2D */
    int64_t n_int;
    const ssr_pattern *pattern;               /* with the Dirichlet mask of Q */
    const double *DPhi1, *DPhi2, *DPhi3;      /* n_p x n_int; DPhi3 NULL in 2D */
    const double *WEIGHT;                     /* n_int */
    const double *c0, *phi, *psi;             /* n_int, unreduced material */
    const double *shear, *bulk, *lame;        /* n_int */
    const double *phi_host, *psi_host;        /* host copies of phi, psi */
    double *lambda;                           /* 1 double: lambda of the reduction */
    double *c_bar, *sin_phi;                  /* n_int, reduced material */
    double *E, *S, *DS;                       /* n_strain, s_rows, n_strain^2 per point */
    const double *K_elast;                    /* block values of the elastic operator */
    double *K_tang, *K_r;                     /* block values, overwritten (K_tang: unused by the loops) */
    const double *mask;                       /* dim*n_n: 1 at free DOFs, 0 at Dirichlet DOFs */
    const double *f;                          /* dim*n_n load vector */
    int64_t n_friction;                       /* entries of phi_host, psi_host (0: n_int) */
    /* optional per-material fields (mat_id NULL: the per-point arrays above are used).  With
     * them the reduction runs once per material and the return mappings read the material
     * of each point's element; c_bar and sin_phi above are then not written. */
    const int32_t *mat_id;                    /* n_int / n_q: material of each element */
    int64_t n_mat;
    const double *c0_m, *phi_m, *psi_m;       /* n_mat, unreduced */
    const double *shear_m, *bulk_m, *lame_m;  /* n_mat */
    double *c_bar_m, *sin_phi_m;              /* n_mat, reduced */
    /* optional node coordinates (dim x n_n node-major; NULL: not used).  Set only when
     * DPhi1..3 came from ssr_dev_fe_gradients on them: the strain and the internal force
     * then compute each point's gradients from the coordinates (ssr_dev_strain_geo,
     * ssr_dev_force_geo, the same bits) instead of reading DPhi. */
    const double *coord;
} ssr_constitutive;

#define SSR_NEWTON_UNKNOWN 0
#define SSR_NEWTON_CONVERGED 1
#define SSR_NEWTON_STALLED 2 /* 'stalled_regularization' */
#define SSR_NEWTON_MAXIT 3
#define SSR_NEWTON_NAN 4

/* [U_it, lambda_it, flag_N, it, history] = newton_ind_SSR(U_ini, omega, lambda_ini,
 *     it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, constitutive_matrix_builder,
 *     linear_solver)
 * replaces slope_stability/+NEWTON/newton_ind_SSR.m:1-206 with damping_ALG5.m:1-60;
 * the whole Newton loop runs in the library.  method 0 = DFGMRES, 1 = DCG on
 * `solver` (whose deflation basis it expands, DFGMRES.m:163-170).  U_ini and
 * U_it are device dim*n_n fields (they may be the same array); work is device
 * scratch of ssr_newton_work_len doubles, disjoint from every other argument.  hist (host, 6 x it_newt_max, or
 * NULL) receives per iteration: relative residual, r, alpha, Krylov
 * iterations for W and for V, deflation basis size; NaN where an iteration
 * stopped before that point.  times (host, 4 doubles, or NULL) accumulates host
 * seconds: Krylov solves, A-orthogonalisation, preconditioner setup, the rest.
 * A NaN residual fails the solve (flag_N = 1). */
int ssr_newton_work_len(const ssr_constitutive *C, int64_t *len);
int ssr_newton_ind_SSR(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini,
                       double omega, double lambda_ini, int it_newt_max, int it_damp_max, double tol,
                       double r_min, double *work, int64_t work_len, double *U_it, double *lambda_out,
                       int *flag_N, int *it, int *status, double *hist, double *times);
/* [U_it, flag_N, it, history] = newton(U_ini, tol, it_newt_max, it_damp_max, r_min,
 *     K_elast, Q, f, constitutive_matrix_builder, linear_system_solver)
 * replaces slope_stability/+NEWTON/newton.m:1-192 with damping.m:1-56: Newton at
 * the material already reduced on C (c_bar, sin_phi), as the initial phase of
 * the continuation calls it.  Arguments as for ssr_newton_ind_SSR; hist rows
 * (5 x it_newt_max): relative residual, r, alpha, Krylov iterations,
 * deflation basis size. */
int ssr_newton(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini, int it_newt_max,
               int it_damp_max, double tol, double r_min, double *work, int64_t work_len, double *U_it,
               int *flag_N, int *it, int *status, double *hist, double *times);
/* [U_it, t_it, flag_N, it] = newton_ind_LL(U_ini, t_ini, omega, it_newt_max, it_damp_max,
 *     tol, r_min, K_elast, Q, f, constitutive_matrix_builder, linear_solver)
 * replaces slope_stability/+NEWTON/newton_ind_LL.m:1-198 (limit-load method): bordered
 * Newton for F(U) = t f, f'U = omega at the material on C, damped by
 * damping.m against a zero field; t takes the full increment and reaching
 * it_newt_max is a failure.  Arguments and hist rows as for ssr_newton_ind_SSR
 * (relative residual, r, alpha, Krylov iterations for W and V, deflation size). */
int ssr_newton_ind_LL(const ssr_constitutive *C, ssr_solver *solver, int method, const double *U_ini,
                      double t_ini, double omega, int it_newt_max, int it_damp_max, double tol, double r_min,
                      double *work, int64_t work_len, double *U_it, double *t_out, int *flag_N, int *it,
                      int *status, double *hist, double *times);

/* [U, lambda_hist, omega_hist, Umax_hist] = SSR_indirect_continuation(lambda_init,
 *     d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min, step_max, omega_max_stop,
 *     it_newt_max, it_damp_max, tol, r_min, K_elast, Q, f, constitutive_matrix_builder,
 *     linear_system_solver)
 * replaces slope_stability/+CONTINUATION/SSR_indirect_continuation.m:1-237 with
 * init_phase_SSR_indirect_continuation.m:1-143 (secant predictor, :119): the whole
 * continuation in the library, over ssr_newton (initial phase) and
 * ssr_newton_ind_SSR, each Newton solve on a copy of `solver`
 * (ssr_solver_clone with share_precond) whose basis the accepted fields expand.
 * params: lambda_init, d_lambda_init, d_lambda_min, d_lambda_diff_scaled_min,
 * step_max, omega_max_stop, it_newt_max, it_damp_max, tol, r_min.  U_out (device,
 * dim*n_n) receives the last accepted field; the histories (host, hist_cap each)
 * the accepted (lambda, omega, max nodal |U|) from the two initial points on;
 * attempt_its / init_its (host) the Newton iterations of every attempt of the
 * main loop and of the initial phase; krylov_its their Krylov iterations.  stop:
 * 1 minimal increase of lambda, 2 maximal omega, 3 omega cannot increase,
 * 4 maximal number of steps; on the reference's errors (:initial choice of lambda
 * too large, FoS equal to lambda_init, FoS above 10, omega_max_stop too small)
 * stop is -1..-4, the call returns SSR_ERR_STATE and ssr_last_error holds the
 * reference's message. */
int ssr_SSR_indirect_continuation(const ssr_constitutive *C, ssr_solver *solver, int method, int share_precond,
                                  const double *params, double *U_out, double *lambda_hist, double *omega_hist,
                                  double *Umax_hist, int hist_cap, int *n_hist, int *attempt_its, int attempt_cap,
                                  int *n_attempts, int *init_its, int init_cap, int *n_init, int64_t *krylov_its,
                                  int *stop);

/* ---- learned continuation predictor (small MLP) --------------------------- */
/* No reference counterpart: the reference continuation has the secant
 * predictor only (SSR_indirect_continuation.m:119), which remains the
 * fallback.  Widths {n_in, n_hidden x (n_layers-1), n_out}; tanh hidden
 * layers; bf16 weights (round-to-nearest-even of the float master) with float
 * accumulation in the forward pass; parameters per layer W (out x in,
 * row-major) then b.  X, Y, dY, grad and the loss data are device pointers. */
typedef struct ssr_mlp ssr_mlp;
int ssr_mlp_create(int n_in, int n_hidden, int n_out, int n_layers, uint64_t seed, double out_scale,
                   ssr_mlp **out);
int ssr_mlp_destroy(ssr_mlp *m);
/* info[0] = n_params, info[1] = n_layers, info[2..] = widths */
int ssr_mlp_info(const ssr_mlp *m, int64_t *info, int cap);
int ssr_mlp_get_params(const ssr_mlp *m, float *host);
int ssr_mlp_set_params(ssr_mlp *m, const float *host);
int ssr_mlp_forward(ssr_mlp *m, int B, const float *X, float *Y);
/* grad of sum_s dY[s]'Y[s] w.r.t. the parameters at X */
int ssr_mlp_backward(ssr_mlp *m, int B, const float *X, const float *dY, float *grad);
/* per-sample loss delta'G delta - 2 delta'g + h + w_lam (y_nc - lam_t)^2 with
 * delta = Y[:, 0:nc], and d(mean loss)/dY (dY may be NULL) */
int ssr_mlp_predictor_loss(int B, int nc, const float *Y, const float *G, const float *g, const float *h,
                           const float *lam_t, float w_lam, float *dY, float *loss);
/* one AdamW step on the mean predictor loss; mean loss before the step to *loss_host */
int ssr_mlp_train_step(ssr_mlp *m, int B, const float *X, const float *G, const float *g, const float *h,
                       const float *lam_t, float w_lam, float lr, float wd, float *loss_host);

#ifdef __cplusplus
}
#endif

#endif /* SSR_HIP_H */
