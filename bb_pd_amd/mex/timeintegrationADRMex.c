/* timeintegrationADRMex -- new MEX name for the MATLAB-only ADR update
 * solvers/functions/timeintegrationADR.m:1 (its .m argument list):
 * timeintegrationADRMex(iTimeStep, DT, nodalVelocityForwardHalf, nodalForce, massVector, cn,
 *                       nodalVelocityPreviousHalf, nodalDisplacement, CONSTRAINTFLAG); the two
 * returned arrays are written in place (arguments 8 and 3).  The two velocities may be the same
 * array. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 9) mexErrMsgTxt("Nine input arguments required \n");
    const int it = (int)mxGetScalar(prhs[0]);
    const double DT = mxGetScalar(prhs[1]), cn = mxGetScalar(prhs[5]);
    const int nNodes = (int)mxGetM(prhs[3]);
    gw_stage g = {{0}, 0};
    double *vf = gw_up(&g, prhs[2]), *F = gw_up(&g, prhs[3]), *m = gw_up(&g, prhs[4]);
    double *vp = prhs[6] == prhs[2] ? vf : gw_up(&g, prhs[6]);
    double *u = gw_up(&g, prhs[7]), *cons = gw_up(&g, prhs[8]);
    gw_check(&g, bbpd_timeintegrationADR(it, DT, vf, F, m, cn, vp, u, cons, nNodes, NULL));
    gw_down(&g, prhs[2], vf);
    gw_down(&g, prhs[7], u);
    gw_free(&g);
}
