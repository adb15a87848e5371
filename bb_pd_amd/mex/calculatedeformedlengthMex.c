/* calculatedeformedlengthMex -- gateway of the reference's
 * solvers/functions/mex/calculatedeformedlengthMex.c:50-79 on libbbpd_hip:
 * calculatedeformedlengthMex(deformedCoordinates, BONDLIST, UNDEFORMEDLENGTH, deformedLength,
 *                            deformedX, deformedY, deformedZ, stretch); writes arguments 4-8. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 8) mexErrMsgTxt("Eight input arguments required \n");
    const int nNodes = (int)mxGetM(prhs[0]);
    const int nBonds = (int)mxGetM(prhs[1]);
    gw_stage g = {{0}, 0};
    double *d[8];
    for (int k = 0; k < 8; k++) d[k] = gw_up(&g, prhs[k]);
    gw_check(&g, bbpd_calculatedeformedlength(d[0], d[1], d[2], nBonds, nNodes, d[3], d[4], d[5], d[6], d[7], NULL));
    for (int k = 3; k < 8; k++) gw_down(&g, prhs[k], d[k]);
    gw_free(&g);
}
