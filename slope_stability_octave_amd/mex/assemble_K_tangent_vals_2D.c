/* assemble_K_tangent_vals_2D.c -- Octave MEX drop-in for
 * slope_stability/+ASSEMBLY/mex/assemble_K_tangent_vals_2D.c, forwarding to libssr_hip.so
 * (ssr_assemble_K_tangent_vals, include/ssr_hip.h).  Same inputs, outputs and error
 * messages; the computation runs on the MI355X.  Build in the package
 * directory of the reference (the file name is what CONSTITUTIVE.m and
 * is_hypre_mex_available.m look for):
 *
 *   mkoctfile --mex -O2 -I<repo>/include assemble_K_tangent_vals_2D.c \
 *       -L<repo>/slope_stability_octave_amd -lssr_hip -Wl,-rpath,<repo>/slope_stability_octave_amd
 */
#include <stdint.h>

#include "mex.h"
#include "ssr_hip.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 7)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:nrhs",
                          "Seven inputs required: DPhi1, DPhi2, DS, WEIGHT, scatter_map, n_q, nnz_out");
    if (nlhs > 1)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:nlhs", "At most one output.");
    if (!mxIsInt64(prhs[4]))
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:type", "scatter_map must be int64.");
    const int n_q = (int)mxGetScalar(prhs[5]);
    const int nnz_out = (int)mxGetScalar(prhs[6]);
    const mwSize n_p = mxGetM(prhs[0]), n_int = mxGetN(prhs[0]), n_e = mxGetN(prhs[4]);
    if (mxGetM(prhs[1]) != n_p || mxGetN(prhs[1]) != n_int)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:size", "DPhi1 and DPhi2 must have the same size.");
    if (mxGetM(prhs[2]) != 9 || mxGetN(prhs[2]) != n_int)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:size", "DS must have size (9, n_int) for 2D.");
    if (mxGetM(prhs[4]) != (mwSize)(4 * n_p * n_p))
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:size", "scatter_map row count must be n_local_dof^2.");
    plhs[0] = mxCreateDoubleMatrix((mwSize)nnz_out, 1, mxREAL);
    if (ssr_assemble_K_tangent_vals(2, (int64_t)n_p, (int64_t)n_int, (int64_t)n_e, n_q, mxGetPr(prhs[0]),
                                    mxGetPr(prhs[1]), NULL, mxGetPr(prhs[2]), mxGetPr(prhs[3]),
                                    (const int64_t *)mxGetData(prhs[4]), (int64_t)nnz_out, mxGetPr(plhs[0])) != 0)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals_2D:hip", "%s", ssr_last_error());
}
