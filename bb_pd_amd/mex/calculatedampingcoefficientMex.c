/* calculatedampingcoefficientMex -- new MEX name for the MATLAB-only ADR damping coefficient
 * solvers/functions/calculatedampingcoefficient.m:1 (its .m argument list):
 * cn = calculatedampingcoefficientMex(nodalForce, massVector, nodalForcePrevious, DT,
 *                                     nodalVelocityPreviousHalf, nodalDisplacement) */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 6) mexErrMsgTxt("Six input arguments required \n");
    const int nNodes = (int)mxGetM(prhs[0]);
    const double DT = mxGetScalar(prhs[3]);
    gw_stage g = {{0}, 0};
    double *F = gw_up(&g, prhs[0]), *m = gw_up(&g, prhs[1]), *Fp = gw_up(&g, prhs[2]);
    double *vp = gw_up(&g, prhs[4]), *u = gw_up(&g, prhs[5]);
    double cn = 0;
    gw_check(&g, bbpd_calculatedampingcoefficient(F, m, Fp, DT, vp, u, nNodes, &cn, NULL));
    gw_free(&g);
    plhs[0] = mxCreateDoubleScalar(cn);
}
