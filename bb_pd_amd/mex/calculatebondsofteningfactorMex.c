/* calculatebondsofteningfactorMex -- new MEX name for the MATLAB-only bilinear law
 * solvers/functions/calculatebondsofteningfactor.m:1 (its .m argument list):
 * calculatebondsofteningfactorMex(stretch, linearElasticLimit, criticalStretchConcrete,
 *                                 flagBondSoftening, bondSofteningFactor, BONDTYPE); the two
 * returned arrays of the .m are written in place (arguments 4 and 5).  linearElasticLimit and
 * criticalStretchConcrete are per-bond arrays or scalars, as MATLAB broadcasts them. */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 6) mexErrMsgTxt("Six input arguments required \n");
    const int nBonds = (int)mxGetM(prhs[0]);
    const int s0s = mxGetNumberOfElements(prhs[1]) == 1 ? 0 : 1;
    const int scs = mxGetNumberOfElements(prhs[2]) == 1 ? 0 : 1;
    gw_stage g = {{0}, 0};
    double *s = gw_up(&g, prhs[0]), *s0 = gw_up(&g, prhs[1]), *sc = gw_up(&g, prhs[2]);
    double *fl = gw_up(&g, prhs[3]), *H = gw_up(&g, prhs[4]), *bt = gw_up(&g, prhs[5]);
    gw_check(&g, bbpd_calculatebondsofteningfactor(s, s0, s0s, sc, scs, fl, H, bt, nBonds, NULL));
    gw_down(&g, prhs[3], fl);
    gw_down(&g, prhs[4], H);
    gw_free(&g);
}
