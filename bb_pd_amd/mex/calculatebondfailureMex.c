/* calculatebondfailureMex -- gateway of the reference's
 * solvers/functions/mex/calculatebondfailureMex.c:47-71 on libbbpd_hip:
 * calculatebondfailureMex(fail, failureFunctionality, BONDTYPE, stretch, criticalStretchConcrete,
 *                         criticalStretchSteel); writes argument 1. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 6) mexErrMsgTxt("Six input arguments required \n");
    const int nBonds = (int)mxGetM(prhs[0]);
    const double ff = mxGetScalar(prhs[1]), csc = mxGetScalar(prhs[4]), css = mxGetScalar(prhs[5]);
    gw_stage g = {{0}, 0};
    double *fail = gw_up(&g, prhs[0]), *bt = gw_up(&g, prhs[2]), *s = gw_up(&g, prhs[3]);
    gw_check(&g, bbpd_calculatebondfailure(fail, ff, bt, s, csc, css, nBonds, NULL));
    gw_down(&g, prhs[0], fail);
    gw_free(&g);
}
