/* mexshim.c -- the implementation of this project's minimal mex.h stand-in (see mex.h). */
#include "mex.h"

#include <setjmp.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

struct mxArray_tag {
    mxClassID cls;
    size_t m, n;
    void *data;      /* double / uint64 elements, or the char string */
    int owns;        /* data allocated here */
    int nfields;     /* struct: field names and values; cell: values */
    char **names;
    mxArray **values;
};

static char *copy_str(const char *s)
{
    size_t n = strlen(s);
    char *d = (char *)malloc(n + 1);
    memcpy(d, s, n + 1);
    return d;
}

static mxArray *make(mxClassID c, size_t m, size_t n)
{
    mxArray *a = (mxArray *)calloc(1, sizeof(mxArray));
    a->cls = c;
    a->m = m;
    a->n = n;
    return a;
}

double *mxGetPr(const mxArray *a) { return a && a->cls == mxDOUBLE_CLASS ? (double *)a->data : NULL; }
void *mxGetData(const mxArray *a) { return a ? a->data : NULL; }
size_t mxGetM(const mxArray *a) { return a ? a->m : 0; }
size_t mxGetN(const mxArray *a) { return a ? a->n : 0; }
size_t mxGetNumberOfElements(const mxArray *a) { return a ? a->m * a->n : 0; }
int mxIsChar(const mxArray *a) { return a && a->cls == mxCHAR_CLASS; }
int mxIsStruct(const mxArray *a) { return a && a->cls == mxSTRUCT_CLASS; }
int mxIsCell(const mxArray *a) { return a && a->cls == mxCELL_CLASS; }
int mxIsDouble(const mxArray *a) { return a && a->cls == mxDOUBLE_CLASS; }
mxClassID mxGetClassID(const mxArray *a) { return a ? a->cls : mxUNKNOWN_CLASS; }

double mxGetScalar(const mxArray *a)
{
    if (!a || a->m * a->n == 0) return 0.0;
    if (a->cls == mxDOUBLE_CLASS) return ((double *)a->data)[0];
    if (a->cls == mxUINT64_CLASS) return (double)((uint64_t *)a->data)[0];
    return 0.0;
}

mxArray *mxGetField(const mxArray *a, mwIndex i, const char *name)
{
    if (!mxIsStruct(a) || i != 0) return NULL;
    for (int f = 0; f < a->nfields; f++)
        if (strcmp(a->names[f], name) == 0) return a->values[f];
    return NULL;
}

mxArray *mxGetCell(const mxArray *a, mwIndex i)
{
    if (!mxIsCell(a) || (int)i >= a->nfields) return NULL;
    return a->values[i];
}

char *mxArrayToString(const mxArray *a)
{
    if (!mxIsChar(a)) return NULL;
    size_t n = strlen((const char *)a->data);
    char *s = (char *)malloc(n + 1);
    memcpy(s, a->data, n + 1);
    return s;
}

mxArray *mxCreateDoubleMatrix(mwSize m, mwSize n, mxComplexity c)
{
    (void)c;
    mxArray *a = make(mxDOUBLE_CLASS, m, n);
    a->data = calloc((m * n) > 0 ? m * n : 1, sizeof(double));
    a->owns = 1;
    return a;
}

mxArray *mxCreateDoubleScalar(double v)
{
    mxArray *a = mxCreateDoubleMatrix(1, 1, mxREAL);
    ((double *)a->data)[0] = v;
    return a;
}

mxArray *mxCreateNumericMatrix(mwSize m, mwSize n, mxClassID c, mxComplexity cx)
{
    (void)cx;
    if (c == mxDOUBLE_CLASS) return mxCreateDoubleMatrix(m, n, mxREAL);
    if (c != mxUINT64_CLASS) return NULL;
    mxArray *a = make(mxUINT64_CLASS, m, n);
    a->data = calloc((m * n) > 0 ? m * n : 1, sizeof(uint64_t));
    a->owns = 1;
    return a;
}

void mxDestroyArray(mxArray *a)
{
    if (!a) return;
    if (a->owns) free(a->data);
    for (int f = 0; f < a->nfields; f++) {
        if (a->names) free(a->names[f]);
        mxDestroyArray(a->values[f]);
    }
    free(a->names);
    free(a->values);
    free(a);
}

void mxFree(void *p) { free(p); }

int mexPrintf(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    int r = vprintf(fmt, ap);
    va_end(ap);
    return r;
}

/* ---- harness ------------------------------------------------------------------------------ */
mxArray *mexshim_double(size_t m, size_t n, double *pr)
{
    mxArray *a = make(mxDOUBLE_CLASS, m, n);
    a->data = pr;
    return a;
}

mxArray *mexshim_string(const char *s)
{
    size_t n = strlen(s);
    mxArray *a = make(mxCHAR_CLASS, 1, n);
    a->data = malloc(n + 1);
    memcpy(a->data, s, n + 1);
    a->owns = 1;
    return a;
}

mxArray *mexshim_struct(int nfields, const char **names, mxArray **values)
{
    mxArray *a = make(mxSTRUCT_CLASS, 1, 1);
    a->nfields = nfields;
    a->names = (char **)calloc(nfields ? nfields : 1, sizeof(char *));
    a->values = (mxArray **)calloc(nfields ? nfields : 1, sizeof(mxArray *));
    for (int f = 0; f < nfields; f++) {
        a->names[f] = copy_str(names[f]);
        a->values[f] = values[f];
    }
    return a;
}

mxArray *mexshim_cell(int n, mxArray **values)
{
    mxArray *a = make(mxCELL_CLASS, 1, (size_t)n);
    a->nfields = n;
    a->values = (mxArray **)calloc(n ? n : 1, sizeof(mxArray *));
    for (int f = 0; f < n; f++) a->values[f] = values[f];
    return a;
}

static __thread jmp_buf *g_env;
static __thread char g_msg[1024];

void mexErrMsgTxt(const char *msg)
{
    snprintf(g_msg, sizeof(g_msg), "%s", msg ? msg : "");
    if (g_env) longjmp(*g_env, 1);
    fprintf(stderr, "mexErrMsgTxt outside a gateway call: %s\n", g_msg);
    abort();
}

int mexshim_call(mexshim_gateway fn, int nlhs, mxArray **plhs, int nrhs, const mxArray **prhs, char *err,
                 int errlen)
{
    jmp_buf env;
    jmp_buf *prev = g_env;
    g_env = &env;
    if (setjmp(env)) {
        g_env = prev;
        if (err && errlen > 0) snprintf(err, (size_t)errlen, "%s", g_msg);
        return 1;
    }
    fn(nlhs, plhs, nrhs, prhs);
    g_env = prev;
    return 0;
}
