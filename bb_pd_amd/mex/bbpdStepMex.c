/* bbpdStepMex -- the fused-step gateway: one MEX for the device-resident solver
 * (bbpd_solver_* in include/bbpd.h).  No counterpart in the reference: it replaces the loop body
 * of solvers/scripts/dynamic/dynamicsolverdisplacementcontrolled.m:71-121 (its five MEX calls and
 * the MATLAB-side law, contact and observables) with whole steps that keep every array on the
 * device.  Sub-commands:
 *
 *   h = bbpdStepMex('create', problem, precision)      problem: struct with the reference's
 *       variable names as fields (undeformedCoordinates, BONDLIST, UNDEFORMEDLENGTH,
 *       VOLUMECORRECTIONFACTORS, BONDSTIFFNESS, BONDTYPE, fail, DENSITY, CONSTRAINTFLAG, massVector,
 *       cellVolume, horizon, RADIJ, law, s0, s1, sc, k, alpha, sc_steel, failureFunctionality,
 *       integrator, DT, DAMPING, bodies = {penetrator, support1, ...} as structs of
 *       buildpenetrator.m with `moves`, schedule, nTimeSteps, appliedDisplacement, increment,
 *       itersPerIncrement, itersPerIncrementMin, eqTol, referenceNode, cmodNodes); missing
 *       optional fields take 0 / NULL; precision 0 = FP64, 1 = FP32-mixed
 *   bbpdStepMex('step', h, n)                          n steps (asynchronous)
 *   bbpdStepMex('graph', h, n, reps)                   n steps captured as a HIP graph, replayed reps times
 *   rows = bbpdStepMex('history', h)                   steps x BBPD_HISTORY_COLUMNS since the last call
 *   o = bbpdStepMex('observables', h)                  [step reaction reaction_static deflection cmod cn
 *                                                       n_softening n_broken n_failed]
 *   [u, v, F, x] = bbpdStepMex('nodal', h)
 *   [H, flag, fail] = bbpdStepMex('bonds', h)
 *   [step, u, v, x, F, Fp, H, flag, fail] = bbpdStepMex('checkpoint', h)
 *   bbpdStepMex('restore', h, step, u, v, x, F, Fp, H, flag, fail)
 *   [inc, iters] = bbpdStepMex('loadstate', h)  /  bbpdStepMex('setloadstate', h, inc, iters)
 *   variant = bbpdStepMex('autotune', h, reps)
 *   bbpdStepMex('destroy', h)
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bbpd.h"
#include "mex.h"

typedef struct {
    bbpd_solver *s;
    int nN, nB;
} handle_t;

static void fail_rc(int rc)
{
    if (rc != BBPD_OK) mexErrMsgTxt(bbpd_last_error());
}

static handle_t *get_handle(const mxArray *a)
{
    if (!a || mxGetClassID(a) != mxUINT64_CLASS || mxGetNumberOfElements(a) != 1)
        mexErrMsgTxt("the handle must be the uint64 scalar 'create' returned \n");
    handle_t *h = (handle_t *)(uintptr_t)((uint64_t *)mxGetData(a))[0];
    if (!h || !h->s) mexErrMsgTxt("invalid handle \n");
    return h;
}

static const double *field_pr(const mxArray *p, const char *name, size_t want, int required)
{
    const mxArray *f = mxGetField(p, 0, name);
    if (!f || mxGetNumberOfElements(f) == 0) {
        if (required) mexErrMsgTxt("problem is missing a required array field \n");
        return NULL;
    }
    if (!mxIsDouble(f) || (want && mxGetNumberOfElements(f) != want))
        mexErrMsgTxt("a problem array field has the wrong size or class \n");
    return mxGetPr(f);
}

static double field_scalar(const mxArray *p, const char *name, double dflt)
{
    const mxArray *f = mxGetField(p, 0, name);
    return f && mxGetNumberOfElements(f) > 0 ? mxGetScalar(f) : dflt;
}

static mxArray *matrix_from(const double *src, size_t m, size_t n)
{
    mxArray *a = mxCreateDoubleMatrix(m, n, mxREAL);
    if (m > 0 && n > 0) memcpy(mxGetPr(a), src, sizeof(double) * m * n);
    return a;
}

static void create(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs != 3 || !mxIsStruct(prhs[1])) mexErrMsgTxt("usage: h = bbpdStepMex('create', problem, precision) \n");
    const mxArray *p = prhs[1];
    const mxArray *X = mxGetField(p, 0, "undeformedCoordinates"), *BL = mxGetField(p, 0, "BONDLIST");
    if (!X || !BL || mxGetN(X) != 3 || mxGetN(BL) != 2)
        mexErrMsgTxt("problem needs undeformedCoordinates (nNodes x 3) and BONDLIST (nBonds x 2) \n");
    bbpd_problem pr;
    memset(&pr, 0, sizeof(pr));
    pr.nNodes = (int)mxGetM(X);
    pr.nBonds = (int)mxGetM(BL);
    const size_t nN = (size_t)pr.nNodes, nB = (size_t)pr.nBonds;
    pr.undeformedCoordinates = mxGetPr(X);
    pr.BONDLIST = mxGetPr(BL);
    pr.UNDEFORMEDLENGTH = field_pr(p, "UNDEFORMEDLENGTH", nB, 1);
    pr.VOLUMECORRECTIONFACTORS = field_pr(p, "VOLUMECORRECTIONFACTORS", nB, 1);
    pr.BONDSTIFFNESS = field_pr(p, "BONDSTIFFNESS", nB, 1);
    pr.BONDTYPE = field_pr(p, "BONDTYPE", nB, 0);
    pr.fail = field_pr(p, "fail", nB, 0);
    pr.DENSITY = field_pr(p, "DENSITY", nN, 1);
    pr.CONSTRAINTFLAG = field_pr(p, "CONSTRAINTFLAG", 3 * nN, 0);
    pr.massVector = field_pr(p, "massVector", nN, 0);
    pr.cellVolume = field_scalar(p, "cellVolume", 0);
    pr.horizon = field_scalar(p, "horizon", 0);
    pr.RADIJ = field_scalar(p, "RADIJ", 0);
    pr.law = (int)field_scalar(p, "law", BBPD_LAW_BILINEAR);
    /* s0 and sc: a per-bond array or a scalar */
    const mxArray *s0 = mxGetField(p, 0, "s0"), *sc = mxGetField(p, 0, "sc");
    if (s0 && mxGetNumberOfElements(s0) == nB && nB > 1) pr.s0 = mxGetPr(s0);
    else pr.s0_scalar = s0 ? mxGetScalar(s0) : 0;
    if (sc && mxGetNumberOfElements(sc) == nB && nB > 1) pr.sc = mxGetPr(sc);
    else pr.sc_scalar = sc ? mxGetScalar(sc) : 0;
    pr.s1 = field_scalar(p, "s1", 0);
    pr.k = field_scalar(p, "k", 25);
    pr.alpha = field_scalar(p, "alpha", 0.25);
    pr.sc_steel = field_scalar(p, "sc_steel", 1);
    pr.failureFunctionality = field_scalar(p, "failureFunctionality", 0);
    pr.integrator = (int)field_scalar(p, "integrator", BBPD_INTEGRATOR_EULER_CROMER);
    pr.DT = field_scalar(p, "DT", 0);
    pr.DAMPING = field_scalar(p, "DAMPING", 0);
    const mxArray *bodies = mxGetField(p, 0, "bodies");
    if (bodies) {
        if (!mxIsCell(bodies) || mxGetNumberOfElements(bodies) > BBPD_MAX_BODIES)
            mexErrMsgTxt("bodies must be a cell array of at most 8 penetrator structs \n");
        pr.nBodies = (int)mxGetNumberOfElements(bodies);
        for (int b = 0; b < pr.nBodies; b++) {
            const mxArray *B = mxGetCell(bodies, (mwIndex)b);
            const mxArray *fam = mxGetField(B, 0, "family"), *cen = mxGetField(B, 0, "centre");
            if (!fam || !cen || mxGetNumberOfElements(cen) != 2)
                mexErrMsgTxt("a body needs family and centre (1 x 2) \n");
            pr.family[b] = mxGetPr(fam);
            pr.nFamily[b] = (int)mxGetNumberOfElements(fam);
            pr.centreX[b] = mxGetPr(cen)[0];
            pr.centreZ[b] = mxGetPr(cen)[1];
            pr.radius[b] = field_scalar(B, "radius", 0);
            pr.moves[b] = (int)field_scalar(B, "moves", 0);
        }
    }
    pr.schedule = (int)field_scalar(p, "schedule", 0);
    pr.nTimeSteps = (int)field_scalar(p, "nTimeSteps", 0);
    pr.appliedDisplacement = field_scalar(p, "appliedDisplacement", 0);
    pr.increment = field_scalar(p, "increment", 0);
    pr.itersPerIncrement = (int)field_scalar(p, "itersPerIncrement", 0);
    pr.itersPerIncrementMin = (int)field_scalar(p, "itersPerIncrementMin", 0);
    pr.eqTol = field_scalar(p, "eqTol", 0);
    pr.referenceNode = (int)field_scalar(p, "referenceNode", 0);
    const mxArray *cm = mxGetField(p, 0, "cmodNodes");
    if (cm && mxGetNumberOfElements(cm) == 2) {
        pr.cmodNode1 = (int)mxGetPr(cm)[0];
        pr.cmodNode2 = (int)mxGetPr(cm)[1];
    }
    bbpd_solver *s = NULL;
    fail_rc(bbpd_solver_create(&pr, (int)mxGetScalar(prhs[2]), 0, &s));
    handle_t *h = (handle_t *)malloc(sizeof(handle_t));
    h->s = s;
    h->nN = pr.nNodes;
    h->nB = pr.nBonds;
    (void)nlhs;
    plhs[0] = mxCreateNumericMatrix(1, 1, mxUINT64_CLASS, mxREAL);
    ((uint64_t *)mxGetData(plhs[0]))[0] = (uint64_t)(uintptr_t)h;
}

void mexFunction(int nlhs, mxArray *plhs[], int nrhs, const mxArray *prhs[])
{
    if (nrhs < 1 || !mxIsChar(prhs[0])) mexErrMsgTxt("the first argument must be a sub-command \n");
    char *cmd = mxArrayToString(prhs[0]);
    char c[32];
    strncpy(c, cmd, sizeof(c) - 1);
    c[sizeof(c) - 1] = 0;
    mxFree(cmd);
    if (!strcmp(c, "create")) {
        create(nlhs, plhs, nrhs, prhs);
        return;
    }
    if (nrhs < 2) mexErrMsgTxt("the sub-command needs the handle \n");
    handle_t *h = get_handle(prhs[1]);
    const size_t nN = (size_t)h->nN, nB = (size_t)h->nB;
    if (!strcmp(c, "step")) {
        if (nrhs != 3) mexErrMsgTxt("usage: bbpdStepMex('step', h, n) \n");
        fail_rc(bbpd_solver_step(h->s, (int)mxGetScalar(prhs[2])));
    } else if (!strcmp(c, "graph")) {
        if (nrhs != 4) mexErrMsgTxt("usage: bbpdStepMex('graph', h, n, reps) \n");
        fail_rc(bbpd_solver_run_graph(h->s, (int)mxGetScalar(prhs[2]), (int)mxGetScalar(prhs[3])));
    } else if (!strcmp(c, "history")) {
        int n = 0;
        const int cap = 1 << 20;
        double *rows = (double *)malloc(sizeof(double) * BBPD_HISTORY_COLUMNS * (size_t)cap);
        int rc = bbpd_solver_history(h->s, rows, cap, &n);
        if (rc != BBPD_OK) {
            free(rows);
            fail_rc(rc);
        }
        plhs[0] = mxCreateDoubleMatrix((size_t)n, BBPD_HISTORY_COLUMNS, mxREAL);
        double *o = mxGetPr(plhs[0]);
        for (int r = 0; r < n; r++)
            for (int k = 0; k < BBPD_HISTORY_COLUMNS; k++)
                o[r + (size_t)n * k] = rows[(size_t)r * BBPD_HISTORY_COLUMNS + k];
        free(rows);
    } else if (!strcmp(c, "observables")) {
        bbpd_observables ob;
        fail_rc(bbpd_solver_observables(h->s, &ob));
        const double v[9] = {(double)ob.step, ob.reaction, ob.reaction_static, ob.deflection, ob.cmod, ob.cn,
                             (double)ob.n_softening, (double)ob.n_broken, (double)ob.n_failed};
        plhs[0] = matrix_from(v, 1, 9);
    } else if (!strcmp(c, "nodal")) {
        for (int k = 0; k < 4; k++) plhs[k] = mxCreateDoubleMatrix(nN, 3, mxREAL);
        fail_rc(bbpd_solver_get_nodal(h->s, mxGetPr(plhs[0]), mxGetPr(plhs[1]), mxGetPr(plhs[2]), mxGetPr(plhs[3])));
    } else if (!strcmp(c, "bonds")) {
        for (int k = 0; k < 3; k++) plhs[k] = mxCreateDoubleMatrix(nB, 1, mxREAL);
        fail_rc(bbpd_solver_get_bonds(h->s, mxGetPr(plhs[0]), mxGetPr(plhs[1]), mxGetPr(plhs[2])));
    } else if (!strcmp(c, "checkpoint")) {
        int64_t step = 0;
        for (int k = 1; k < 6; k++) plhs[k] = mxCreateDoubleMatrix(nN, 3, mxREAL);
        for (int k = 6; k < 9; k++) plhs[k] = mxCreateDoubleMatrix(nB, 1, mxREAL);
        fail_rc(bbpd_solver_checkpoint(h->s, &step, mxGetPr(plhs[1]), mxGetPr(plhs[2]), mxGetPr(plhs[3]),
                                       mxGetPr(plhs[4]), mxGetPr(plhs[5]), mxGetPr(plhs[6]), mxGetPr(plhs[7]),
                                       mxGetPr(plhs[8])));
        plhs[0] = mxCreateDoubleScalar((double)step);
    } else if (!strcmp(c, "restore")) {
        if (nrhs != 11) mexErrMsgTxt("usage: bbpdStepMex('restore', h, step, u, v, x, F, Fp, H, flag, fail) \n");
        const double *a[8];
        for (int k = 0; k < 8; k++) a[k] = mxGetNumberOfElements(prhs[3 + k]) ? mxGetPr(prhs[3 + k]) : NULL;
        fail_rc(bbpd_solver_restore(h->s, (int64_t)mxGetScalar(prhs[2]), a[0], a[1], a[2], a[3], a[4], a[5], a[6],
                                    a[7]));
    } else if (!strcmp(c, "loadstate")) {
        int64_t inc = 0, it = 0;
        fail_rc(bbpd_solver_get_load_state(h->s, &inc, &it));
        plhs[0] = mxCreateDoubleScalar((double)inc);
        plhs[1] = mxCreateDoubleScalar((double)it);
    } else if (!strcmp(c, "setloadstate")) {
        if (nrhs != 4) mexErrMsgTxt("usage: bbpdStepMex('setloadstate', h, inc, iters) \n");
        fail_rc(bbpd_solver_set_load_state(h->s, (int64_t)mxGetScalar(prhs[2]), (int64_t)mxGetScalar(prhs[3])));
    } else if (!strcmp(c, "autotune")) {
        int choice = 0;
        fail_rc(bbpd_solver_autotune(h->s, nrhs > 2 ? (int)mxGetScalar(prhs[2]) : 5, NULL, &choice));
        plhs[0] = mxCreateDoubleScalar((double)choice);
    } else if (!strcmp(c, "destroy")) {
        fail_rc(bbpd_solver_destroy(h->s));
        h->s = NULL;
        free(h);
        ((uint64_t *)mxGetData(prhs[1]))[0] = 0;
    } else {
        mexErrMsgTxt("unknown sub-command \n");
    }
}
