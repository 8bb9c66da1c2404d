/* assemble_K_tangent_vals.c -- Octave MEX drop-in for
 * slope_stability/+ASSEMBLY/mex/assemble_K_tangent_vals.c, forwarding to libssr_hip.so
 * (ssr_assemble_K_tangent_vals, include/ssr_hip.h).  Same inputs, outputs and error
 * messages; the computation runs on the MI355X.  Build in the package
 * directory of the reference (the file name is what CONSTITUTIVE.m and
 * is_hypre_mex_available.m look for):
 *
 *   mkoctfile --mex -O2 -I<repo>/include assemble_K_tangent_vals.c \
 *       -L<repo>/slope_stability_octave_amd -lssr_hip -Wl,-rpath,<repo>/slope_stability_octave_amd
 */
#include <stdint.h>

#include "mex.h"
#include "ssr_hip.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 8)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals:nrhs",
                          "Eight inputs required: DPhi1, DPhi2, DPhi3, DS, WEIGHT, scatter_map, n_q, nnz_out");
    if (nlhs > 1)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals:nlhs", "At most one output.");
    if (!mxIsInt64(prhs[5]))
        mexErrMsgIdAndTxt("assemble_K_tangent_vals:type", "scatter_map must be int64.");
    const int n_q = (int)mxGetScalar(prhs[6]);
    const int nnz_out = (int)mxGetScalar(prhs[7]);
    const mwSize n_p = mxGetM(prhs[0]), n_int = mxGetN(prhs[0]), n_e = mxGetN(prhs[5]);
    plhs[0] = mxCreateDoubleMatrix((mwSize)nnz_out, 1, mxREAL); /* zero-filled */
    if (ssr_assemble_K_tangent_vals(3, (int64_t)n_p, (int64_t)n_int, (int64_t)n_e, n_q, mxGetPr(prhs[0]),
                                    mxGetPr(prhs[1]), mxGetPr(prhs[2]), mxGetPr(prhs[3]), mxGetPr(prhs[4]),
                                    (const int64_t *)mxGetData(prhs[5]), (int64_t)nnz_out, mxGetPr(plhs[0])) != 0)
        mexErrMsgIdAndTxt("assemble_K_tangent_vals:hip", "%s", ssr_last_error());
}
