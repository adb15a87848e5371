/*
 * bbpd_gateway.h -- host staging shared by the MEX gateways of libbbpd_hip.
 *
 * Each gateway in this directory is the file a MATLAB driver builds in place of the reference's MEX
 * file of the same name (solvers/functions/mex/makefile.m:15-28): the same mexFunction argument
 * list and error texts, the same in-place results.  MATLAB's arrays live in host memory, so a
 * gateway copies the arguments a routine reads to the device, calls the library's entry point
 * (device pointers, include/bbpd.h) and copies the arguments the reference writes back into
 * MATLAB's own buffers.  Error paths free what was staged before mexErrMsgTxt, which does not
 * return.
 */
#ifndef BBPD_GATEWAY_H
#define BBPD_GATEWAY_H

#include <hip/hip_runtime_api.h>
#include <stddef.h>

#include "bbpd.h"
#include "mex.h"

#define GW_MAX 16

typedef struct {
    double *d[GW_MAX];
    int n;
} gw_stage;

static inline void gw_free(gw_stage *g)
{
    for (int k = 0; k < g->n; k++)
        if (g->d[k]) (void)hipFree(g->d[k]);
    g->n = 0;
}

static inline void gw_fail(gw_stage *g, const char *msg)
{
    gw_free(g);
    mexErrMsgTxt(msg);
}

/* a device buffer of n doubles */
static inline double *gw_alloc(gw_stage *g, size_t n)
{
    double *d = NULL;
    if (g->n >= GW_MAX) gw_fail(g, "too many staged arrays \n");
    if (hipMalloc((void **)&d, (n ? n : 1) * sizeof(double)) != hipSuccess) gw_fail(g, "hipMalloc failed \n");
    g->d[g->n++] = d;
    return d;
}

/* a device copy of one MATLAB double array */
static inline double *gw_up(gw_stage *g, const mxArray *a)
{
    if (!mxIsDouble(a)) gw_fail(g, "arguments must be double arrays \n");
    size_t n = mxGetNumberOfElements(a);
    double *d = gw_alloc(g, n);
    if (n && hipMemcpy(d, mxGetPr(a), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        gw_fail(g, "hipMemcpy to the device failed \n");
    return d;
}

/* the device result back into MATLAB's buffer of the argument (in place, as the reference writes) */
static inline void gw_down(gw_stage *g, const mxArray *a, const double *d)
{
    size_t n = mxGetNumberOfElements(a);
    if (n && hipMemcpy(mxGetPr(a), d, n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        gw_fail(g, "hipMemcpy to the host failed \n");
}

/* a library error: free, then report the library's message */
static inline void gw_check(gw_stage *g, int rc)
{
    if (rc != BBPD_OK) gw_fail(g, bbpd_last_error());
}

#endif
