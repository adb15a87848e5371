/* timeintegrationeulercromerMex -- gateway of the reference's
 * solvers/functions/mex/timeintegrationeulercromerMex.c:56-85 on libbbpd_hip:
 * timeintegrationeulercromerMex(nodalAcceleration, nodalForce, DAMPING, nodalVelocity, DENSITY,
 *                               CONSTRAINTFLAG, DT, nodalDisplacement, deformedCoordinates,
 *                               undeformedCoordinates); writes arguments 1, 4, 8 and 9. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 10) mexErrMsgTxt("Ten input arguments required \n");
    const int nNodes = (int)mxGetM(prhs[0]);
    const int NOD = (int)mxGetN(prhs[0]);
    const double D = mxGetScalar(prhs[2]), DT = mxGetScalar(prhs[6]);
    gw_stage g = {{0}, 0};
    double *a = gw_up(&g, prhs[0]), *F = gw_up(&g, prhs[1]), *v = gw_up(&g, prhs[3]), *rho = gw_up(&g, prhs[4]);
    double *cons = gw_up(&g, prhs[5]), *u = gw_up(&g, prhs[7]), *x = gw_up(&g, prhs[8]), *X = gw_up(&g, prhs[9]);
    gw_check(&g, bbpd_timeintegrationeulercromer(a, F, D, v, rho, cons, DT, u, x, X, nNodes, NOD, NULL));
    gw_down(&g, prhs[0], a);
    gw_down(&g, prhs[3], v);
    gw_down(&g, prhs[7], u);
    gw_down(&g, prhs[8], x);
    gw_free(&g);
}
