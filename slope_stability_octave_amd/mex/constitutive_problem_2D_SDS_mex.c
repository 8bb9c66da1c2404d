/* constitutive_problem_2D_SDS_mex.c -- Octave MEX drop-in for
 * slope_stability/+CONSTITUTIVE_PROBLEM/mex/constitutive_problem_2D_SDS_mex.c, forwarding to libssr_hip.so
 * (ssr_constitutive_problem_2D_SDS, include/ssr_hip.h).  Same inputs, outputs and error
 * messages; the computation runs on the MI355X.  Build in the package
 * directory of the reference (the file name is what CONSTITUTIVE.m and
 * is_hypre_mex_available.m look for):
 *
 *   mkoctfile --mex -O2 -I<repo>/include constitutive_problem_2D_SDS_mex.c \
 *       -L<repo>/slope_stability_octave_amd -lssr_hip -Wl,-rpath,<repo>/slope_stability_octave_amd
 */
#include <stdint.h>

#include "mex.h"
#include "ssr_hip.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 6)
        mexErrMsgTxt("constitutive_problem_2D_SDS_mex: "
                     "6 inputs required (E, c_bar, sin_phi, shear, bulk, lame)");
    if (nlhs > 2)
        mexErrMsgTxt("constitutive_problem_2D_SDS_mex: up to 2 outputs (S, DS)");
    if (mxGetM(prhs[0]) != 3)
        mexErrMsgTxt("constitutive_problem_2D_SDS_mex: E must have 3 rows.");
    const mwSize n_int = mxGetN(prhs[0]);
    plhs[0] = mxCreateDoubleMatrix(4, n_int, mxREAL);
    plhs[1] = mxCreateDoubleMatrix(9, n_int, mxREAL);
    if (ssr_constitutive_problem_2D_SDS((int64_t)n_int, (int64_t)mxGetM(prhs[0]), mxGetPr(prhs[0]), mxGetPr(prhs[1]), mxGetPr(prhs[2]),
        mxGetPr(prhs[3]), mxGetPr(prhs[4]), mxGetPr(prhs[5]), mxGetPr(plhs[0]), mxGetPr(plhs[1])) != 0)
        mexErrMsgIdAndTxt("constitutive_problem_2D_SDS_mex:hip", "%s", ssr_last_error());
}
