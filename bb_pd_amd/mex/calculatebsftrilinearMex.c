/* calculatebsftrilinearMex -- new MEX name for the MATLAB-only trilinear law
 * solvers/functions/calculatebsftrilinear.m:1 (its .m argument list):
 * calculatebsftrilinearMex(stretch, s0, s1, sc, bondSofteningFactor, BONDTYPE, flagBondSoftening)
 * with scalar s0, s1, sc; the two returned arrays are written in place (arguments 5 and 7). */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    (void)plhs;
    if (nrhs != 7) mexErrMsgTxt("Seven input arguments required \n");
    const int nBonds = (int)mxGetM(prhs[0]);
    const double s0 = mxGetScalar(prhs[1]), s1 = mxGetScalar(prhs[2]), sc = mxGetScalar(prhs[3]);
    gw_stage g = {{0}, 0};
    double *s = gw_up(&g, prhs[0]), *H = gw_up(&g, prhs[4]), *bt = gw_up(&g, prhs[5]), *fl = gw_up(&g, prhs[6]);
    gw_check(&g, bbpd_calculatebsftrilinear(s, s0, s1, sc, H, bt, fl, nBonds, NULL));
    gw_down(&g, prhs[4], H);
    gw_down(&g, prhs[6], fl);
    gw_free(&g);
}
