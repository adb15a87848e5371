/* calculatenodalforcesopenmpMex -- gateway of the reference's
 * solvers/functions/mex/calculatenodalforcesopenmpMex.c:88-116 on libbbpd_hip:
 * calculatenodalforcesopenmpMex(BONDLIST, nodalForce, bForceX, bForceY, bForceZ, nodalForceX,
 *                               nodalForceY, nodalForceZ); writes arguments 2 and 6-8.  The sum
 * order is the reference's with the 8 threads it forces (:39), which fill columns 1-8 of the
 * nNodes x nThreads scratch arrays; scratch with fewer than 8 columns, which the reference would
 * write past, is refused. */
#include "bbpd_gateway.h"

#define GW_THREADS 8

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 8) mexErrMsgTxt("Eight input arguments required \n");
    const int nNodes = (int)mxGetM(prhs[1]);
    const int nBonds = (int)mxGetM(prhs[0]);
    for (int k = 5; k < 8; k++)
        if (mxGetM(prhs[k]) != (size_t)nNodes || mxGetN(prhs[k]) < GW_THREADS)
            mexErrMsgTxt("nodalForceX/Y/Z must be nNodes x nThreads with nThreads >= 8 \n");
    gw_stage g = {{0}, 0};
    double *bl = gw_up(&g, prhs[0]), *F = gw_up(&g, prhs[1]);
    double *fx = gw_up(&g, prhs[2]), *fy = gw_up(&g, prhs[3]), *fz = gw_up(&g, prhs[4]);
    double *sx = gw_up(&g, prhs[5]), *sy = gw_up(&g, prhs[6]), *sz = gw_up(&g, prhs[7]);
    gw_check(&g, bbpd_calculatenodalforces(nBonds, nNodes, bl, F, fx, fy, fz, sx, sy, sz, GW_THREADS, NULL));
    gw_down(&g, prhs[1], F);
    gw_down(&g, prhs[5], sx);
    gw_down(&g, prhs[6], sy);
    gw_down(&g, prhs[7], sz);
    gw_free(&g);
}
