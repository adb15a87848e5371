/* calculatecontactforceMex -- new MEX name for the MATLAB-only rigid-cylinder contact
 * solvers/functions/calculatecontactforce.m:1 (its .m argument list):
 * penetratorfz = calculatecontactforceMex(penetrator, displacementIncrement, undeformedCoordinates,
 *                                         deformedCoordinates, nodalDisplacement, nodalVelocity, DT,
 *                                         cellVolume, DENSITY)
 * `penetrator` is the struct of input/functions/buildpenetrator.m:5-24 (family, centre, radius);
 * the returned deformedCoordinates, nodalDisplacement and nodalVelocity are written in place
 * (arguments 4-6). */
#include "bbpd_gateway.h"

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    (void)nlhs;
    if (nrhs != 9) mexErrMsgTxt("Nine input arguments required \n");
    const mxArray *pen = prhs[0];
    const mxArray *fam = mxGetField(pen, 0, "family"), *cen = mxGetField(pen, 0, "centre"),
                  *rad = mxGetField(pen, 0, "radius");
    if (!fam || !cen || !rad || mxGetNumberOfElements(cen) != 2)
        mexErrMsgTxt("penetrator must be a struct with family, centre (1 x 2) and radius \n");
    const double cx = mxGetPr(cen)[0], cz = mxGetPr(cen)[1], R = mxGetScalar(rad);
    const double disp = mxGetScalar(prhs[1]), DT = mxGetScalar(prhs[6]), V = mxGetScalar(prhs[7]);
    const int nNodes = (int)mxGetM(prhs[2]);
    const int nFam = (int)mxGetNumberOfElements(fam);
    gw_stage g = {{0}, 0};
    double *f = gw_up(&g, fam), *X = gw_up(&g, prhs[2]), *x = gw_up(&g, prhs[3]), *u = gw_up(&g, prhs[4]);
    double *v = gw_up(&g, prhs[5]), *rho = gw_up(&g, prhs[8]);
    double fz = 0;
    gw_check(&g, bbpd_calculatecontactforce(f, nFam, cx, cz, R, disp, X, x, u, v, DT, V, rho, nNodes, &fz, NULL));
    gw_down(&g, prhs[3], x);
    gw_down(&g, prhs[4], u);
    gw_down(&g, prhs[5], v);
    gw_free(&g);
    plhs[0] = mxCreateDoubleScalar(fz);
}
