/* calculatebsfexponentialMex -- gateway of the reference's
 * solvers/functions/mex/calculatebsfexponentialMex.c:69-96 on libbbpd_hip:
 * calculatebsfexponentialMex(stretch, s0, sc, k, alpha, bsf, bondSofteningFactor, BONDTYPE,
 *                            flagBondSoftening); writes arguments 6, 7 and 9. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 9) mexErrMsgTxt("Nine input arguments required \n");
    const int nBonds = (int)mxGetM(prhs[0]);
    const double k = mxGetScalar(prhs[3]), alpha = mxGetScalar(prhs[4]);
    gw_stage g = {{0}, 0};
    double *s = gw_up(&g, prhs[0]), *s0 = gw_up(&g, prhs[1]), *sc = gw_up(&g, prhs[2]);
    double *bsf = gw_up(&g, prhs[5]), *H = gw_up(&g, prhs[6]), *bt = gw_up(&g, prhs[7]), *fl = gw_up(&g, prhs[8]);
    gw_check(&g, bbpd_calculatebsfexponential(s, s0, sc, k, alpha, bsf, H, bt, fl, nBonds, NULL));
    gw_down(&g, prhs[5], bsf);
    gw_down(&g, prhs[6], H);
    gw_down(&g, prhs[8], fl);
    gw_free(&g);
}
