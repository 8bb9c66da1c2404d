/* hypre_boomeramg_mex.c -- Octave MEX drop-in for
 * slope_stability/+LINEAR_SOLVERS/mex/hypre_boomeramg_mex.cpp: the same
 * command surface ('setup', 'setup_ijv', 'update_values', 'apply', 'clear',
 * 'info') and error ids, over the smoothed-aggregation AMG instances of
 * libssr_hip.so (ssr_amg_*, include/ssr_hip.h) instead of BoomerAMG.
 * Options read from the opts struct: num_functions (DOFs per node, default
 * 3), force_rebuild, and this AMG's own theta, max_levels, max_coarse,
 * cheb_degree; BoomerAMG-only fields are accepted and ignored.  Build:
 *
 *   mkoctfile --mex -O2 -I<repo>/include hypre_boomeramg_mex.c \
 *       -L<repo>/slope_stability_octave_amd -lssr_hip -Wl,-rpath,<repo>/slope_stability_octave_amd
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mex.h"
#include "ssr_hip.h"

static void fail(const char *id, const char *msg) { mexErrMsgIdAndTxt(id, "%s", msg); }

static void check(int rc)
{
    if (rc == 0) return;
    const char *msg = ssr_last_error();
    if (strstr(msg, "unknown instance")) fail("hypre:noInstance", msg);
    if (strstr(msg, "does not match")) fail("hypre:badRhs", msg);
    if (strstr(msg, "values")) fail("hypre:badV", msg);
    fail("hypre:exception", msg);
}

static char *id_string(const mxArray *a)
{
    if (mxIsChar(a)) {
        char *s = mxArrayToString(a);
        if (!s) fail("hypre:badId", "Failed to convert char instance_id.");
        return s;
    }
    if (mxIsNumeric(a) && mxGetNumberOfElements(a) == 1) {
        char *s = (char *)mxMalloc(32);
        snprintf(s, 32, "%lld", (long long)mxGetScalar(a));
        return s;
    }
    fail("hypre:badId", "instance_id must be a string or a scalar.");
    return NULL;
}

static double opt_field(const mxArray *opts, const char *name, double dflt)
{
    if (!opts || !mxIsStruct(opts)) return dflt;
    const mxArray *f = mxGetField(opts, 0, name);
    if (!f || mxIsEmpty(f)) return dflt;
    if (mxIsLogical(f)) return mxIsLogicalScalarTrue(f) ? 1.0 : 0.0;
    return mxGetScalar(f);
}

/* {theta, max_levels, max_coarse, cheb_degree, cheb_lower, cheb_upper, refresh_every, rebuild_every} */
static void amg_opts(const mxArray *opts, double o[8])
{
    o[0] = opt_field(opts, "theta", 0.0);
    o[1] = opt_field(opts, "max_levels", 12);
    o[2] = opt_field(opts, "max_coarse", 600);
    o[3] = opt_field(opts, "cheb_degree", 3);
    o[4] = 1.0 / 30.0;
    o[5] = 1.1;
    o[6] = 0;
    o[7] = 0;
}

static mxArray *info_struct(const char *id, const double info[5])
{
    const char *fields[] = {"instance_id", "n", "nnz", "num_functions", "num_interp_vectors", "setup_time_seconds"};
    mxArray *out = mxCreateStructMatrix(1, 1, 6, fields);
    mxSetField(out, 0, "instance_id", mxCreateString(id));
    for (int k = 0; k < 5; ++k) mxSetField(out, 0, fields[k + 1], mxCreateDoubleScalar(info[k]));
    return out;
}

static void setup_ijv(int nlhs, mxArray *plhs[], const double *I, const double *J, const double *V, int64_t nnz,
                      int64_t n, const mxArray *Z, const mxArray *opts, const mxArray *idarr)
{
    if (Z && !mxIsEmpty(Z) && (!mxIsDouble(Z) || (int64_t)mxGetM(Z) != n))
        fail("hypre:badZ", "Z must be a real double matrix with n rows.");
    char *id = id_string(idarr);
    double o[8], info[5];
    amg_opts(opts, o);
    const int nz = (Z && !mxIsEmpty(Z)) ? (int)mxGetN(Z) : 0;
    check(ssr_amg_setup_ijv(id, I, J, V, nnz, n, nz ? mxGetPr(Z) : NULL, nz,
                            (int)opt_field(opts, "num_functions", 3), o, (int)opt_field(opts, "force_rebuild", 0),
                            info));
    if (nlhs > 0) plhs[0] = info_struct(id, info);
    mxFree(id);
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs < 1 || !mxIsChar(prhs[0]))
        fail("hypre:usage", "Usage: hypre_boomeramg_mex('setup'| 'apply' | 'clear' | 'info', ...)");
    char *cmd = mxArrayToString(prhs[0]);
    if (!cmd) fail("hypre:usage", "Failed to parse command.");
    if (!strcmp(cmd, "setup_ijv")) {
        /* hypre_boomeramg_mex('setup_ijv', I, J, V, n, Z, opts, id) */
        if (nrhs != 8)
            fail("hypre:usage", "setup_ijv usage: hypre_boomeramg_mex('setup_ijv', I, J, V, n, Z, opts, id)");
        if (!mxIsDouble(prhs[1]) || !mxIsDouble(prhs[2]) || !mxIsDouble(prhs[3]))
            fail("hypre:badIJV", "I, J, V must be real double vectors.");
        const int64_t nnz = (int64_t)mxGetNumberOfElements(prhs[3]);
        if ((int64_t)mxGetNumberOfElements(prhs[1]) != nnz || (int64_t)mxGetNumberOfElements(prhs[2]) != nnz)
            fail("hypre:badIJV", "I, J, V must have equal length.");
        setup_ijv(nlhs, plhs, mxGetPr(prhs[1]), mxGetPr(prhs[2]), mxGetPr(prhs[3]), nnz,
                  (int64_t)mxGetScalar(prhs[4]), prhs[5], prhs[6], prhs[7]);
    } else if (!strcmp(cmd, "setup")) {
        /* hypre_boomeramg_mex('setup', A, Z, opts, id): A sparse, converted to 1-based COO */
        if (nrhs != 5) fail("hypre:usage", "setup usage: hypre_boomeramg_mex('setup', A, Z, opts, id)");
        const mxArray *A = prhs[1];
        if (!mxIsSparse(A) || !mxIsDouble(A) || mxGetM(A) != mxGetN(A))
            fail("hypre:badA", "A must be a square real sparse matrix.");
        const mwIndex *ir = mxGetIr(A), *jc = mxGetJc(A);
        const double *pr = mxGetPr(A);
        const int64_t n = (int64_t)mxGetN(A), nnz = (int64_t)jc[n];
        double *I = (double *)mxMalloc(sizeof(double) * (nnz ? nnz : 1));
        double *J = (double *)mxMalloc(sizeof(double) * (nnz ? nnz : 1));
        for (int64_t c = 0; c < n; ++c)
            for (mwIndex t = jc[c]; t < jc[c + 1]; ++t) {
                I[t] = (double)ir[t] + 1.0;
                J[t] = (double)c + 1.0;
            }
        setup_ijv(nlhs, plhs, I, J, pr, nnz, n, prhs[2], prhs[3], prhs[4]);
        mxFree(I);
        mxFree(J);
    } else if (!strcmp(cmd, "update_values")) {
        /* t = hypre_boomeramg_mex('update_values', V, id) */
        if (nrhs != 3) fail("hypre:usage", "update_values usage: hypre_boomeramg_mex('update_values', V, id)");
        if (!mxIsDouble(prhs[1]) || mxIsComplex(prhs[1])) fail("hypre:badV", "V must be a real double vector.");
        char *id = id_string(prhs[2]);
        double t = 0.0;
        check(ssr_amg_update_values(id, mxGetPr(prhs[1]), (int64_t)mxGetNumberOfElements(prhs[1]), &t));
        if (nlhs > 0) plhs[0] = mxCreateDoubleScalar(t);
        mxFree(id);
    } else if (!strcmp(cmd, "apply")) {
        /* y = hypre_boomeramg_mex('apply', r, id) */
        if (nrhs != 3) fail("hypre:usage", "apply usage: hypre_boomeramg_mex('apply', rhs, instance_id)");
        if (nlhs > 1) fail("hypre:usage", "apply returns one output vector.");
        if (!mxIsDouble(prhs[1]) || mxIsComplex(prhs[1])) fail("hypre:badRhs", "rhs must be a real double vector.");
        char *id = id_string(prhs[2]);
        const int64_t n = (int64_t)mxGetNumberOfElements(prhs[1]);
        plhs[0] = mxCreateDoubleMatrix((mwSize)n, 1, mxREAL);
        check(ssr_amg_apply(id, mxGetPr(prhs[1]), n, mxGetPr(plhs[0])));
        mxFree(id);
    } else if (!strcmp(cmd, "clear")) {
        if (nrhs == 1) {
            check(ssr_amg_clear(NULL));
        } else {
            char *id = id_string(prhs[1]);
            check(ssr_amg_clear(id));
            mxFree(id);
        }
    } else if (!strcmp(cmd, "info")) {
        int64_t count = 0, total = 0;
        check(ssr_amg_info(&count, &total));
        if (nlhs > 0) {
            const char *fields[] = {"num_instances", "total_n"};
            plhs[0] = mxCreateStructMatrix(1, 1, 2, fields);
            mxSetField(plhs[0], 0, "num_instances", mxCreateDoubleScalar((double)count));
            mxSetField(plhs[0], 0, "total_n", mxCreateDoubleScalar((double)total));
        } else {
            mexPrintf("hypre_boomeramg_mex: %lld instance(s), %lld unknowns in all\n", (long long)count,
                      (long long)total);
        }
    } else {
        fail("hypre:usage", "Unknown command.");
    }
    mxFree(cmd);
}
