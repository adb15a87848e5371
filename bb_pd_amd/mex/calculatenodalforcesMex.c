/* calculatenodalforcesMex -- gateway of the reference's serial
 * solvers/functions/mex/calculatenodalforcesMex.c:49-73 on libbbpd_hip:
 * calculatenodalforcesMex(BONDLIST, nodalForce, bForceX, bForceY, bForceZ); writes argument 2 in
 * the serial sum order. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 5) mexErrMsgTxt("Five input arguments required \n");
    const int nNodes = (int)mxGetM(prhs[1]);
    const int nBonds = (int)mxGetM(prhs[0]);
    gw_stage g = {{0}, 0};
    double *bl = gw_up(&g, prhs[0]), *F = gw_up(&g, prhs[1]);
    double *fx = gw_up(&g, prhs[2]), *fy = gw_up(&g, prhs[3]), *fz = gw_up(&g, prhs[4]);
    gw_check(&g, bbpd_calculatenodalforces(nBonds, nNodes, bl, F, fx, fy, fz, NULL, NULL, NULL, 1, NULL));
    gw_down(&g, prhs[1], F);
    gw_free(&g);
}
