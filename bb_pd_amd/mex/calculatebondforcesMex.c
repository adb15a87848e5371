/* calculatebondforcesMex -- gateway of the reference's
 * solvers/functions/mex/calculatebondforcesMex.c:48-80 on libbbpd_hip:
 * calculatebondforcesMex(bForceX, bForceY, bForceZ, fail, deformedX, deformedY, deformedZ,
 *                        deformedLength, stretch, stretchPlastic, BONDSTIFFNESS, cellVolume,
 *                        VOLUMECORRECTIONFACTORS, bondSofteningFactor); writes arguments 1-3. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 14) mexErrMsgTxt("Fourteen input arguments required \n");
    const int nBonds = (int)mxGetM(prhs[0]);
    const double V = mxGetScalar(prhs[11]);
    gw_stage g = {{0}, 0};
    double *d[14] = {0};
    for (int k = 0; k < 14; k++)
        if (k != 11) d[k] = gw_up(&g, prhs[k]);
    gw_check(&g, bbpd_calculatebondforces(d[0], d[1], d[2], d[3], d[4], d[5], d[6], d[7], d[8], d[9], d[10], V,
                                          d[12], d[13], nBonds, NULL));
    for (int k = 0; k < 3; k++) gw_down(&g, prhs[k], d[k]);
    gw_free(&g);
}
