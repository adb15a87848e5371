/*
 * bbpd_oracle.c -- CPU restatement of the reference bond-based peridynamics time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity oracle: only tests/, the smoke check and the
 * CPU-baseline leg of bench.py load it, and only as the checker or the timed CPU baseline.  The
 * product (bb_pd_amd/) never links or calls it.
 *
 * Every routine restates one function of mark-hobbs/BB_PD operation for operation, in the
 * reference data model: FP64, column-major nodal arrays (x block, then y at +nNodes, then z at
 * +2*nNodes), BONDLIST as nBonds x 2 doubles holding 1-based node ids, flags as doubles.  The
 * citation above each routine names the reference file and lines it follows.  Build with
 * -ffp-contract=off so that no multiply-add is fused: the reference's gcc build on x86-64 does not
 * fuse either, and the bit patterns of this file are compared with the GPU's.
 *
 * Parity pins (what checks each routine without running the reference, which needs MATLAB's
 * mex.h and so cannot be built in this image) are listed in the header of oracle/oracle.py.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------------------------------
 * A  calculatedeformedlength  -- solvers/functions/mex/calculatedeformedlengthMex.c:27-47
 * ---------------------------------------------------------------------------------------------- */
void or_calculatedeformedlength(const double *x, const double *BONDLIST, const double *L0, int nBonds,
                                int nNodes, double *l, double *dX, double *dY, double *dZ, double *s)
{
    for (int k = 0; k < nBonds; k++) {
        int i = (int)BONDLIST[k];
        int j = (int)BONDLIST[nBonds + k];
        dX[k] = x[j - 1] - x[i - 1];
        dY[k] = x[j + nNodes - 1] - x[i + nNodes - 1];
        dZ[k] = x[j + 2 * nNodes - 1] - x[i + 2 * nNodes - 1];
        l[k] = sqrt(dX[k] * dX[k] + dY[k] * dY[k] + dZ[k] * dZ[k]);
        s[k] = (l[k] - L0[k]) / L0[k];
    }
}

/* ------------------------------------------------------------------------------------------------
 * B4 exponential softening -- solvers/functions/mex/calculatebsfexponentialMex.c:26-66
 *    bsf is the per-step scratch the caller zeroes (dynamicsolverdisplacementcontrolled.m:86).
 * ---------------------------------------------------------------------------------------------- */
void or_calculatebsfexponential(const double *s, const double *s0, const double *sc, double k, double alpha,
                                double *bsf, double *H, const double *BONDTYPE, double *flag, int nBonds)
{
    for (int b = 0; b < nBonds; b++) {
        if (BONDTYPE[b] == 0) {
            if (s0[b] < s[b] && s[b] <= sc[b]) {
                flag[b] = 1;
                double numerator = 1 - exp(-k * (s[b] - s0[b]) / (sc[b] - s0[b]));
                double denominator = 1 - exp(-k);
                double residual = alpha * (1 - (s[b] - s0[b]) / (sc[b] - s0[b]));
                bsf[b] = 1 - (s0[b] / s[b]) * ((1 - (numerator / denominator)) + residual) / (1 + alpha);
            } else if (s[b] > sc[b]) {
                bsf[b] = 1;
            }
            if (bsf[b] > H[b]) H[b] = bsf[b];
            if (H[b] > 1) H[b] = 1;
            if (H[b] != H[b]) H[b] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------------------------------
 * B2 bilinear softening -- solvers/functions/calculatebondsofteningfactor.m:42-48
 *    s0/sc are per-bond arrays when s0_stride/sc_stride is 1, scalars when 0 (MATLAB broadcasting).
 *    MATLAB's max(a,b) returns the non-NaN operand, so a NaN candidate leaves H unchanged.
 * ---------------------------------------------------------------------------------------------- */
void or_calculatebondsofteningfactor(const double *s, const double *s0, int s0_stride, const double *sc,
                                     int sc_stride, double *flag, double *H, const double *BONDTYPE, int nBonds)
{
    for (int b = 0; b < nBonds; b++) {
        double a = s0[(size_t)b * s0_stride];
        double c = sc[(size_t)b * sc_stride];
        if (BONDTYPE[b] == 0 && s[b] > a) flag[b] = 1;                   /* :42 */
        double cand = (((s[b] - a) / s[b]) * (c / (c - a))) * flag[b];    /* :44, :46 (bsf .* flag) */
        if (cand == cand && cand > H[b]) H[b] = cand;                     /* :46 max ignores NaN */
        if (H[b] > 1) H[b] = 1;                                           /* :47 */
        if (H[b] != H[b]) H[b] = 0;                                       /* :48 */
    }
}

/* ------------------------------------------------------------------------------------------------
 * B3 trilinear softening -- solvers/functions/calculatebsftrilinear.m:40-112 (scalar s0, s1, sc)
 * ---------------------------------------------------------------------------------------------- */
void or_calculatebsftrilinear(const double *s, double s0, double s1, double sc, double *H,
                              const double *BONDTYPE, double *flag, int nBonds)
{
    const double beta = 0.25;                                             /* :41 */
    const double eta = s1 / s0;                                           /* :42 */
    for (int b = 0; b < nBonds; b++) {
        double bsf = 0;                                                   /* :44 */
        if (BONDTYPE[b] == 0) {
            if ((s[b] > s0) && (s[b] <= s1)) {                            /* :84-87 */
                bsf = (1 - ((eta - beta) / (eta - 1) * (s0 / s[b]))) + ((1 - beta) / (eta - 1));
                flag[b] = 1;
            } else if ((s[b] > s1) && (s[b] <= sc)) {                     /* :89-91 */
                bsf = 1 - ((s0 * beta) / s[b]) * ((sc - s[b]) / (sc - s1));
            } else if (s[b] > sc) {                                       /* :93-96 */
                bsf = 1;
            }
            if (bsf > H[b]) H[b] = bsf;                                   /* :100-104 */
        }
        if (H[b] != H[b]) H[b] = 0;                                       /* :112 */
    }
}

/* ------------------------------------------------------------------------------------------------
 * B1 linear (brittle) failure latch -- solvers/functions/mex/calculatebondfailureMex.c:26-44
 * ---------------------------------------------------------------------------------------------- */
void or_calculatebondfailure(double *fail, double ff, const double *BONDTYPE, const double *s, double csc,
                             double css, int nBonds)
{
    for (int b = 0; b < nBonds; b++) {
        if (fail[b] == 1 && BONDTYPE[b] == 0 && s[b] > csc) fail[b] = ff;
        if (fail[b] == 1 && BONDTYPE[b] == 1 && s[b] > 3 * csc) fail[b] = ff;
        if (fail[b] == 1 && BONDTYPE[b] == 2 && s[b] > css) fail[b] = ff;
    }
}

/* ------------------------------------------------------------------------------------------------
 * C  calculatebondforces -- solvers/functions/mex/calculatebondforcesMex.c:26-45
 *    left-to-right product, NaN -> 0 per component.
 * ---------------------------------------------------------------------------------------------- */
void or_calculatebondforces(double *fX, double *fY, double *fZ, const double *fail, const double *dX,
                            const double *dY, const double *dZ, const double *l, const double *s,
                            const double *sp, const double *c, double V, const double *vcf, const double *H,
                            int nBonds)
{
    for (int b = 0; b < nBonds; b++) {
        fX[b] = fail[b] * c[b] * (1 - H[b]) * (s[b] - sp[b]) * V * vcf[b] * (dX[b] / l[b]);
        fY[b] = fail[b] * c[b] * (1 - H[b]) * (s[b] - sp[b]) * V * vcf[b] * (dY[b] / l[b]);
        fZ[b] = fail[b] * c[b] * (1 - H[b]) * (s[b] - sp[b]) * V * vcf[b] * (dZ[b] / l[b]);
        if (fX[b] != fX[b]) fX[b] = 0.0;
        if (fY[b] != fY[b]) fY[b] = 0.0;
        if (fZ[b] != fZ[b]) fZ[b] = 0.0;
    }
}

/* ------------------------------------------------------------------------------------------------
 * D  nodal force reduction, serial order -- solvers/functions/mex/calculatenodalforcesMex.c:19-46
 *    F must be zeroed by the caller (dynamicsolverdisplacementcontrolled.m:97).
 * ---------------------------------------------------------------------------------------------- */
void or_calculatenodalforces_serial(int nBonds, int nNodes, const double *BONDLIST, double *F, const double *fX,
                                    const double *fY, const double *fZ)
{
    for (int k = 0; k < nBonds; k++) {
        int i = (int)BONDLIST[k];
        int j = (int)BONDLIST[nBonds + k];
        F[i - 1] += fX[k];
        F[j - 1] -= fX[k];
        F[i + nNodes - 1] += fY[k];
        F[j + nNodes - 1] -= fY[k];
        F[i + 2 * nNodes - 1] += fZ[k];
        F[j + 2 * nNodes - 1] -= fZ[k];
    }
}

/* Static-schedule block of thread t for n iterations over T threads (libgomp's schedule(static)
 * without a chunk: the first n%T threads take one extra iteration). */
static void static_block(int n, int T, int t, int *lo, int *hi)
{
    int q = n / T, r = n % T;
    *lo = t * q + (t < r ? t : r);
    *hi = *lo + q + (t < r ? 1 : 0);
}

/* ------------------------------------------------------------------------------------------------
 * D' nodal force reduction, OpenMP order -- solvers/functions/mex/calculatenodalforcesopenmpMex.c:29-85
 *    Emulated sequentially: thread t scatters its static block of bonds into column t of the
 *    nNodes x T scratch (:50-66), then every node adds the columns in thread order (:69-81).  The
 *    result is what the reference returns with T threads (it forces T = 8, :39).  Scratch arrays
 *    must be zeroed by the caller (dynamicsolverdisplacementcontrolled.m:98-100).
 * ---------------------------------------------------------------------------------------------- */
void or_calculatenodalforces_openmp(int nBonds, int nNodes, const double *BONDLIST, double *F, const double *fX,
                                    const double *fY, const double *fZ, double *sX, double *sY, double *sZ,
                                    int T)
{
    for (int t = 0; t < T; t++) {
        int lo, hi;
        static_block(nBonds, T, t, &lo, &hi);
        for (int k = lo; k < hi; k++) {
            int i = (int)BONDLIST[k];
            int j = (int)BONDLIST[nBonds + k];
            sX[(i - 1) + (size_t)nNodes * t] += fX[k];
            sX[(j - 1) + (size_t)nNodes * t] -= fX[k];
            sY[(i - 1) + (size_t)nNodes * t] += fY[k];
            sY[(j - 1) + (size_t)nNodes * t] -= fY[k];
            sZ[(i - 1) + (size_t)nNodes * t] += fZ[k];
            sZ[(j - 1) + (size_t)nNodes * t] -= fZ[k];
        }
    }
    for (int n = 0; n < nNodes; n++)
        for (int t = 0; t < T; t++) {
            F[n] += sX[n + (size_t)nNodes * t];
            F[n + nNodes] += sY[n + (size_t)nNodes * t];
            F[n + 2 * nNodes] += sZ[n + (size_t)nNodes * t];
        }
}

/* The same reduction executed by a real OpenMP team of T threads, with the loop structure of
 * calculatenodalforcesopenmpMex.c:41-82 (parallel region, `omp for` over bonds into the thread's
 * scratch column, `omp for` over nodes adding the columns).  Used to check that the sequential
 * emulation above reproduces libgomp's static schedule. */
void or_calculatenodalforces_omp_real(int nBonds, int nNodes, const double *BONDLIST, double *F, const double *fX,
                                      const double *fY, const double *fZ, double *sX, double *sY, double *sZ, int T)
{
#ifdef _OPENMP
    omp_set_num_threads(T);
#pragma omp parallel
    {
        int nT = omp_get_num_threads();
        int t = omp_get_thread_num();
#pragma omp for
        for (int k = 0; k < nBonds; k++) {
            int i = (int)BONDLIST[k];
            int j = (int)BONDLIST[nBonds + k];
            sX[(i - 1) + (size_t)nNodes * t] += fX[k];
            sX[(j - 1) + (size_t)nNodes * t] -= fX[k];
            sY[(i - 1) + (size_t)nNodes * t] += fY[k];
            sY[(j - 1) + (size_t)nNodes * t] -= fY[k];
            sZ[(i - 1) + (size_t)nNodes * t] += fZ[k];
            sZ[(j - 1) + (size_t)nNodes * t] -= fZ[k];
        }
#pragma omp for
        for (int n = 0; n < nNodes; n++)
            for (int tt = 0; tt < nT; tt++) {
                F[n] += sX[n + (size_t)nNodes * tt];
                F[n + nNodes] += sY[n + (size_t)nNodes * tt];
                F[n + 2 * nNodes] += sZ[n + (size_t)nNodes * tt];
            }
    }
#else
    or_calculatenodalforces_openmp(nBonds, nNodes, BONDLIST, F, fX, fY, fZ, sX, sY, sZ, T);
#endif
}

/* ------------------------------------------------------------------------------------------------
 * E  Euler-Cromer update -- solvers/functions/mex/timeintegrationeulercromerMex.c:26-53
 * ---------------------------------------------------------------------------------------------- */
void or_timeintegrationeulercromer(double *a, const double *F, double D, double *v, const double *rho,
                                   const double *cons, double DT, double *u, double *x, const double *X,
                                   int nNodes, int NOD)
{
    for (int n = 0; n < nNodes; n++)
        for (int d = 0; d < NOD; d++) {
            size_t q = n + (size_t)nNodes * d;
            a[q] = (F[q] - D * v[q]) / rho[n];
            if (cons[q] == 1) a[q] = 0;
            v[q] = v[q] + (a[q] * DT);
            u[q] = u[q] + (v[q] * DT);
            x[q] = X[q] + u[q];
        }
}

/* ------------------------------------------------------------------------------------------------
 * ADR damping coefficient -- solvers/functions/calculatedampingcoefficient.m:59-112 (NOD == 3)
 *    Sequential node-major sum, massVector(i,1) for every DOF (:65,:71,:77).
 *    `skip` (NULL = none) leaves nodes out of both sums: the stepped solver passes the members of
 *    the contact families, whose DOFs the rigid-body projection drives.
 * ---------------------------------------------------------------------------------------------- */
double or_calculatedampingcoefficient_masked(const double *F, const double *m, const double *Fp, double DT,
                                             const double *vp, const double *u, int nNodes,
                                             const unsigned char *skip)
{
    double cn = 0, cn1 = 0, cn2 = 0;
    for (int i = 0; i < nNodes; i++) {
        if (skip && skip[i]) continue;
        for (int d = 0; d < 3; d++) {
            size_t q = i + (size_t)nNodes * d;
            if (vp[q] != 0)
                cn1 = cn1 - (u[q] * u[q]) * ((F[q] / m[i] - Fp[q] / m[i]) / (DT * vp[q]));
        }
        for (int d = 0; d < 3; d++) {
            size_t q = i + (size_t)nNodes * d;
            cn2 = cn2 + u[q] * u[q];
        }
    }
    if (cn2 != 0) {
        if ((cn1 / cn2) > 0) cn = 2 * sqrt(cn1 / cn2);
        else cn = 0;
    } else {
        cn = 0;
    }
    if (cn > 2) cn = 1.9;
    return cn;
}

double or_calculatedampingcoefficient(const double *F, const double *m, const double *Fp, double DT,
                                      const double *vp, const double *u, int nNodes)
{
    return or_calculatedampingcoefficient_masked(F, m, Fp, DT, vp, u, nNodes, NULL);
}

/* ------------------------------------------------------------------------------------------------
 * The device's summation orders.  Not reference routines: the reference sums sequentially
 * (calculatedampingcoefficient.m:61-85, MATLAB's sum in calculatereactionforceFast.m:3-5).  They let
 * the oracle take the same sums in the order the GPU's fixed-shape reductions add their terms, so
 * that trajectories whose damping coefficient, static reaction or equilibrium ratio feeds back into
 * the state (ADR, the equilibrium staircase) can be compared bit for bit.
 *   or_tree_sum   n values (n a power of two) by the halving tree v[t] = v[t] + v[t + w] for
 *                 w = n/2, n/4, .., 1 and t < w; returns v[0] (v is overwritten)
 *   or_slot_sum   one value per node slot (nSlots a multiple of 256): each block of 256 consecutive
 *                 slots by the halving tree; then lane t of one 256-lane block adds the block sums
 *                 t, t + 256, t + 512, .. in that order starting from 0, and the 256 lanes are
 *                 summed by the halving tree
 * ---------------------------------------------------------------------------------------------- */
double or_tree_sum(double *v, int n)
{
    for (int w = n / 2; w > 0; w >>= 1)
        for (int t = 0; t < w; t++) v[t] = v[t] + v[t + w];
    return v[0];
}

double or_slot_sum(const double *by_slot, long nSlots)
{
    long nblk = nSlots / 256;
    double *part = (double *)malloc(sizeof(double) * (size_t)(nblk > 0 ? nblk : 1));
    double buf[256];
    for (long b = 0; b < nblk; b++) {
        memcpy(buf, by_slot + 256 * b, sizeof(buf));
        part[b] = or_tree_sum(buf, 256);
    }
    for (int t = 0; t < 256; t++) {
        double c = 0;
        for (long q = t; q < nblk; q += 256) c = c + part[q];
        buf[t] = c;
    }
    free(part);
    return or_tree_sum(buf, 256);
}

/* The ADR sums of calculatedampingcoefficient.m:61-85 per node (massVector(i,1) for every DOF, the
 * terms of a node added d = 1..3 from 0), over the nodes with skip == 0 (NULL = all), placed at the
 * node's slot and summed by or_slot_sum: out[0] = cn1, out[1] = cn2.  out[2] is the squared
 * out-of-balance force sum F_d^2 over the DOFs that neither CONSTRAINTFLAG (cons, nNodes x 3, may be
 * NULL) nor skip holds, the numerator of the equilibrium ratio (see or_equilibrium_ratio). */
void or_dampingsums_device(const double *F, const double *m, const double *Fp, double DT, const double *vp,
                           const double *u, int nNodes, const unsigned char *skip, const double *cons,
                           const int *slot_of, long nSlots, double *out)
{
    double *v[3];
    for (int k = 0; k < 3; k++) v[k] = (double *)calloc((size_t)nSlots, sizeof(double));
    for (int i = 0; i < nNodes; i++) {
        if (skip && skip[i]) continue;
        double c1 = 0, c2 = 0, c3 = 0;
        for (int d = 0; d < 3; d++) {
            size_t q = i + (size_t)nNodes * d;
            if (vp[q] != 0) c1 = c1 - (u[q] * u[q]) * ((F[q] / m[i] - Fp[q] / m[i]) / (DT * vp[q]));
            c2 = c2 + u[q] * u[q];
            if (!(cons && cons[q] == 1)) c3 = c3 + F[q] * F[q];
        }
        v[0][slot_of[i]] = c1;
        v[1][slot_of[i]] = c2;
        v[2][slot_of[i]] = c3;
    }
    for (int k = 0; k < 3; k++) {
        out[k] = or_slot_sum(v[k], nSlots);
        free(v[k]);
    }
}

/* cn from the two sums (calculatedampingcoefficient.m:90-112) */
double or_damping_from_sums(double cn1, double cn2)
{
    double cn = 0;
    if (cn2 != 0) {
        if ((cn1 / cn2) > 0) cn = 2 * sqrt(cn1 / cn2);
        else cn = 0;
    }
    if (cn > 2) cn = 1.9;
    return cn;
}

/* Static reaction of calculatereactionforceFast.m:3-5 (F_z * cellVolume over a body's family) in the
 * device's order: lane t of a 1024-lane block adds members t, t + 1024, .. in family order starting
 * from 0, then the halving tree over the lanes. */
double or_static_reaction(const double *F, const double *family, int nFam, double V, int nNodes)
{
    double lanes[1024];
    for (int t = 0; t < 1024; t++) {
        double c = 0;
        for (int f = t; f < nFam; f += 1024) c = c + F[(int)family[f] - 1 + 2 * (size_t)nNodes] * V;
        lanes[t] = c;
    }
    return or_tree_sum(lanes, 1024);
}

/* Equilibrium ratio of a displacement-controlled step, after calculateequilibriumstate.m:6-8 (the
 * out-of-balance force over the applied load): sqrt(sum of F_d^2 over the free DOFs) * cellVolume
 * divided by |static reaction|; +infinity while the reaction is 0. */
double or_equilibrium_ratio(double res2, double V, double reaction)
{
    double a = fabs(reaction);
    return a > 0 ? (sqrt(res2) * V) / a : INFINITY;
}

/* ------------------------------------------------------------------------------------------------
 * ADR update -- solvers/functions/timeintegrationADR.m:51-80 (NOD == 3); m^-1 is 1/m.
 * ---------------------------------------------------------------------------------------------- */
void or_timeintegrationADR(int iTimeStep, double DT, double *vf, const double *F, const double *m, double cn,
                           const double *vp, double *u, const double *cons, int nNodes)
{
    for (int n = 0; n < nNodes; n++) {
        double minv = 1 / m[n];
        for (int d = 0; d < 3; d++) {
            size_t q = n + (size_t)nNodes * d;
            if (iTimeStep == 1)
                vf[q] = 0.5 * DT * minv * F[q];
            else
                vf[q] = (((2 - cn * DT) * vp[q]) + (2 * DT * minv * F[q])) / (2 + cn * DT);
            u[q] = u[q] + (DT * vf[q]);
        }
    }
    for (size_t q = 0; q < (size_t)nNodes * 3; q++)
        if (cons[q] == 1) {
            u[q] = 0;
            vf[q] = 0;
        }
}

/* ------------------------------------------------------------------------------------------------
 * Fictitious ADR mass -- solvers/functions/buildmassmatrix.m:32-40
 * ---------------------------------------------------------------------------------------------- */
void or_buildmassmatrix(int nNodes, double DX, double c0, double DT, const double *nFAM, double *m)
{
    for (int n = 0; n < nNodes; n++) {
        double Kij = (nFAM[n] * (DX * DX)) * (c0 / DX);
        m[n] = 0.25 * (DT * DT) * Kij;
    }
}

/* ------------------------------------------------------------------------------------------------
 * Rigid-cylinder contact -- solvers/functions/calculatecontactforce.m:62-121
 *    family holds 1-based node ids.  The centre moves by `disp` for this call only (:70; the
 *    struct is passed by value).  Returns penetratorfz (:121), summed in family order.
 * ---------------------------------------------------------------------------------------------- */
double or_calculatecontactforce(const double *family, int nFam, double cx, double cz, double R, double disp,
                                const double *X, double *x, double *u, double *v, double DT, double V,
                                const double *rho, int nNodes)
{
    double czm = cz + disp;
    double fz = 0;
    for (int f = 0; f < nFam; f++) {
        int i = (int)family[f] - 1;
        double ddx = x[i] - cx;
        double ddz = x[i + 2 * (size_t)nNodes] - czm;
        double dist = sqrt((ddx * ddx) + (ddz * ddz));
        if (dist < R) {
            double ux = (ddx / dist) * R, uz = (ddz / dist) * R;
            x[i] = cx + ux;
            x[i + 2 * (size_t)nNodes] = czm + uz;
            double dfz = 0;
            for (int d = 0; d < 3; d++) {
                size_t q = i + (size_t)nNodes * d;
                double uo = u[q], vo = v[q];
                u[q] = x[q] - X[q];
                v[q] = (u[q] - uo) / DT;
                double dF = -1 * rho[i] * (v[q] - vo) / DT * V;
                if (d == 2) dfz = dF;
            }
            fz = fz + dfz;
        }
    }
    return fz;
}

/* smoothstepdata -- solvers/functions/smoothstepdata.m:40-42 */
double or_smoothstepdata(double cur, double start, double fin, double a0, double a1)
{
    double xi = (cur - start) / (fin - start);
    return a0 + (a1 - a0) * (xi * xi * xi) * (10 - 15 * xi + 6 * (xi * xi));
}

/* damage field -- solvers/functions/calculatedamage.m:34-45 (broken = fail == 0 or H == 1 here) */
void or_calculatedamage(const double *BONDLIST, const double *unbroken, const double *nFAM, int nBonds, int nNodes,
                        double *damage)
{
    double *cnt = (double *)calloc((size_t)nNodes, sizeof(double));
    for (int k = 0; k < nBonds; k++) {
        int i = (int)BONDLIST[k], j = (int)BONDLIST[nBonds + k];
        cnt[i - 1] += unbroken[k];
        cnt[j - 1] += unbroken[k];
    }
    for (int n = 0; n < nNodes; n++) damage[n] = 1 - (cnt[n] / nFAM[n]);
    free(cnt);
}

/* strain energy -- solvers/functions/calculatestrainenergy.m:10-27.  The micropotential of :16 left
 * to right (BFMULTIPLIER NULL = 1, buildbonddata.m:56-57), added to both nodes of each bond in
 * BONDLIST order (:19-24), halved (:24), then sum(density) * nNodes * cellVolume (:26). */
double or_calculatestrainenergy(int nNodes, const double *fail, const double *c, const double *BF, const double *s,
                                const double *L0, const double *vcf, const double *BONDLIST, int nBonds, double V,
                                double *density)
{
    for (int n = 0; n < nNodes; n++) density[n] = 0;
    for (int k = 0; k < nBonds; k++) {
        double mu = 0.5 * fail[k] * (BF ? BF[k] : 1.0) * c[k] * (s[k] * s[k]) * L0[k] * V * vcf[k];
        int i = (int)BONDLIST[k], j = (int)BONDLIST[nBonds + k];
        density[i - 1] = density[i - 1] + mu;
        density[j - 1] = density[j - 1] + mu;
    }
    double t = 0;
    for (int n = 0; n < nNodes; n++) {
        density[n] = 0.5 * density[n];
        t = t + density[n];
    }
    return t * nNodes * V;
}

/* dissipated damage energy -- solvers/functions/calculatedamageenergy.m:35-62 */
void or_calculatedamageenergy(const double *c, const double *s0, const double *sc, const double *L0,
                              const double *s, double k, double alpha, double V, double *de_hist, int nBonds)
{
    double denominator = 2 * k * (exp(k) - 1) * (1 + alpha);
    double shape = (2 * k) - (2 * exp(k)) + (alpha * k) - (alpha * k * exp(k)) + 2;
    for (int b = 0; b < nBonds; b++) {
        double de = 0;
        if ((s0[b] < s[b]) && (s[b] <= sc[b])) {
            double numerator = c[b] * s0[b] * L0[b] * (s0[b] - s[b]) * shape;
            de = (numerator / denominator) * (V * V);
        } else if (s[b] > sc[b]) {
            double numerator = c[b] * s0[b] * L0[b] * (s0[b] - sc[b]) * shape;
            de = (numerator / denominator) * (V * V);
        }
        if (de > de_hist[b]) de_hist[b] = de;
    }
}

/* Load-point displacement of step i under a staircase of increments: increment * ceil(i / ipi), the
 * penetrator moved by `increment` every ipi ADR iterations (the "5e-6 m per increment" protocol;
 * the legacy prescribed displacement of solvers/scripts/dynamic/timeintegration.m:40 applies one
 * such increment). */
double or_staircase(long i, long ipi, double increment)
{
    return increment * (double)((i + ipi - 1) / ipi);
}

/* ================================================================================================
 * One explicit step in the reference driver's order
 * (solvers/scripts/dynamic/dynamicsolverdisplacementcontrolled.m:71-121), ADR variant from
 * calculatedampingcoefficient.m + timeintegrationADR.m with the commented hand-over lines of
 * dynamicsolverdisplacementcontrolledADR.m:153-155.  All arrays in reference layout.
 * ============================================================================================== */
typedef struct or_state {
    int nNodes, nBonds;
    /* inputs */
    const double *X, *BONDLIST, *L0, *vcf, *c, *BONDTYPE, *s0, *sc, *rho, *cons, *mass;
    int s0_stride, sc_stride; /* 0 = scalar in s0[0]/sc[0] */
    double V, DT, DAMPING, s1, k, alpha, sc_steel, ff;
    int law;        /* 0 linear, 1 bilinear, 2 trilinear, 3 exponential */
    int integrator; /* 0 Euler-Cromer, 1 ADR */
    int nodal_threads; /* 0 serial reduction order, T>0 OpenMP order with T threads */
    /* contact bodies */
    int nBodies;
    const double *const *families;
    const int *nFams;
    const double *centres; /* nBodies x 2 (cx, cz) */
    const double *radii;
    const int *moves; /* 1 if the body follows the load schedule */
    /* state (in/out) */
    double *x, *u, *v, *a, *F, *Fp, *H, *flag, *fail;
    /* scratch */
    double *l, *dX, *dY, *dZ, *s, *bsf, *fX, *fY, *fZ, *sp, *sXs, *sYs, *sZs;
    double *s0_full, *sc_full; /* expanded per-bond copies for the exponential routine */
    /* load schedule, the load-point displacement of step i:
     *   0  smoothstepdata(i, 1, nTimeSteps, 0, applied)   (dynamicsolverdisplacementcontrolled.m:76)
     *   1  or_staircase(i, ipi, increment)
     *   2  equilibrium staircase: increment * inc, where inc (from 1) advances after the step at
     *      whose end the equilibrium ratio (or_equilibrium_ratio of this step's forces and static
     *      reaction) is below eq_tol after at least ipi_min iterations at inc, or after ipi
     *      iterations at inc (ADR only) */
    int schedule, nTimeSteps, ipi, ipi_min;
    double applied, increment, eq_tol;
    /* order of the ADR sums and of the static reaction: slot_of NULL = the reference's sequential
     * order; else node i's terms sit at slot slot_of[i] of nSlots and are summed by or_slot_sum */
    const int *slot_of;
    long nSlots;
    /* load state (in/out): the current increment and the iterations done at it (schedule 2) */
    long inc, inc_iters;
    /* records of the last step (out): displacement, static reaction of body 0, equilibrium ratio,
     * damping coefficient, 1 when the step ended a load increment (schedules 1 and 2) */
    double last_disp, last_static, last_ratio, last_cn;
    int last_inc_end;
} or_state;

/* returns the contact reaction of body 0 (calculatecontactforce.m:121) */
double or_step(or_state *S, int iTimeStep, double *fz_bodies)
{
    int nN = S->nNodes, nB = S->nBonds;
    or_calculatedeformedlength(S->x, S->BONDLIST, S->L0, nB, nN, S->l, S->dX, S->dY, S->dZ, S->s);
    switch (S->law) {
    case 0:
        or_calculatebondfailure(S->fail, S->ff, S->BONDTYPE, S->s, S->sc[0], S->sc_steel, nB);
        break;
    case 1:
        or_calculatebondsofteningfactor(S->s, S->s0, S->s0_stride, S->sc, S->sc_stride, S->flag, S->H, S->BONDTYPE, nB);
        break;
    case 2:
        or_calculatebsftrilinear(S->s, S->s0[0], S->s1, S->sc[0], S->H, S->BONDTYPE, S->flag, nB);
        break;
    default:
        memset(S->bsf, 0, sizeof(double) * nB);
        or_calculatebsfexponential(S->s, S->s0_full, S->sc_full, S->k, S->alpha, S->bsf, S->H, S->BONDTYPE, S->flag, nB);
        break;
    }
    or_calculatebondforces(S->fX, S->fY, S->fZ, S->fail, S->dX, S->dY, S->dZ, S->l, S->s, S->sp, S->c, S->V, S->vcf,
                           S->H, nB);
    memset(S->F, 0, sizeof(double) * 3 * nN);
    if (S->nodal_threads <= 0) {
        or_calculatenodalforces_serial(nB, nN, S->BONDLIST, S->F, S->fX, S->fY, S->fZ);
    } else {
        size_t sz = sizeof(double) * (size_t)nN * S->nodal_threads;
        memset(S->sXs, 0, sz);
        memset(S->sYs, 0, sz);
        memset(S->sZs, 0, sz);
        or_calculatenodalforces_openmp(nB, nN, S->BONDLIST, S->F, S->fX, S->fY, S->fZ, S->sXs, S->sYs, S->sZs,
                                       S->nodal_threads);
    }
    double res2 = 0;
    S->last_cn = 0;
    if (S->integrator == 0) {
        or_timeintegrationeulercromer(S->a, S->F, S->DAMPING, S->v, S->rho, S->cons, S->DT, S->u, S->x, S->X, nN, 3);
    } else {
        /* The Rayleigh quotient is taken over the nodes no contact body drives.  A family member's
         * velocity after a projection is the projection's correction, not a mode of the structure,
         * and its term (u^2 dF / (m DT v)) outweighs the rest by many orders of magnitude, which pins
         * cn to 0 or the 1.9 clamp.  Membership is fixed, so cn stays a continuous function of the
         * state. */
        unsigned char *driven = (unsigned char *)calloc((size_t)nN, 1);
        for (int b = 0; b < S->nBodies; b++)
            for (int f = 0; f < S->nFams[b]; f++) driven[(int)S->families[b][f] - 1] = 1;
        double cn;
        if (S->slot_of) {
            double sums[3];
            or_dampingsums_device(S->F, S->mass, S->Fp, S->DT, S->v, S->u, nN, driven, S->cons, S->slot_of,
                                  S->nSlots, sums);
            cn = or_damping_from_sums(sums[0], sums[1]);
            res2 = sums[2];
        } else {
            cn = or_calculatedampingcoefficient_masked(S->F, S->mass, S->Fp, S->DT, S->v, S->u, nN, driven);
            for (int i = 0; i < nN; i++) {
                if (driven[i]) continue;
                for (int d = 0; d < 3; d++) {
                    size_t q = i + (size_t)nN * d;
                    if (S->cons[q] != 1) res2 = res2 + S->F[q] * S->F[q];
                }
            }
        }
        free(driven);
        S->last_cn = cn;
        /* v holds the previous half-step velocity on entry and the new one on exit */
        or_timeintegrationADR(iTimeStep, S->DT, S->v, S->F, S->mass, cn, S->v, S->u, S->cons, nN);
        for (size_t q = 0; q < (size_t)nN * 3; q++) S->x[q] = S->X[q] + S->u[q];
    }
    double disp;
    if (S->schedule == 0) disp = or_smoothstepdata((double)iTimeStep, 1, (double)S->nTimeSteps, 0, S->applied);
    else if (S->schedule == 1) disp = or_staircase(iTimeStep, S->ipi, S->increment);
    else disp = S->increment * (double)S->inc;
    S->last_disp = disp;
    double r0 = 0;
    for (int b = 0; b < S->nBodies; b++) {
        double d = S->moves[b] ? disp : 0;
        double fz = or_calculatecontactforce(S->families[b], S->nFams[b], S->centres[2 * b], S->centres[2 * b + 1],
                                             S->radii[b], d, S->X, S->x, S->u, S->v, S->DT, S->V, S->rho, nN);
        if (fz_bodies) fz_bodies[b] = fz;
        if (b == 0) r0 = fz;
    }
    /* static reaction of body 0 from this step's forces */
    double rs = 0;
    if (S->nBodies > 0) {
        if (S->slot_of) {
            rs = or_static_reaction(S->F, S->families[0], S->nFams[0], S->V, nN);
        } else {
            for (int f = 0; f < S->nFams[0]; f++) rs = rs + S->F[(int)S->families[0][f] - 1 + 2 * (size_t)nN] * S->V;
        }
    }
    S->last_static = rs;
    S->last_ratio = or_equilibrium_ratio(res2, S->V, rs);
    S->last_inc_end = 0;
    if (S->schedule == 1) {
        S->last_inc_end = (iTimeStep % S->ipi) == 0;
    } else if (S->schedule == 2) {
        S->inc_iters += 1;
        if ((S->inc_iters >= S->ipi_min && S->last_ratio < S->eq_tol) || S->inc_iters >= S->ipi) {
            S->last_inc_end = 1;
            S->inc += 1;
            S->inc_iters = 0;
        }
    }
    if (S->integrator == 1) memcpy(S->Fp, S->F, sizeof(double) * 3 * nN);
    return r0;
}

/* ================================================================================================
 * Timed CPU baseline: passes A-E as separate OpenMP loops over nBonds-length intermediates with a
 * static schedule and the per-thread scratch reduction (calculatenodalforcesopenmpMex.c:41-82).
 * Same arithmetic as above; used only to time the reference's CPU path.
 * ============================================================================================== */
void or_step_omp(or_state *S, int T)
{
    int nN = S->nNodes, nB = S->nBonds;
    const double *x = S->x, *BL = S->BONDLIST, *L0 = S->L0;
#ifdef _OPENMP
    omp_set_num_threads(T);
#endif
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nB; k++) {
        int i = (int)BL[k], j = (int)BL[nB + k];
        S->dX[k] = x[j - 1] - x[i - 1];
        S->dY[k] = x[j + nN - 1] - x[i + nN - 1];
        S->dZ[k] = x[j + 2 * nN - 1] - x[i + 2 * nN - 1];
        S->l[k] = sqrt(S->dX[k] * S->dX[k] + S->dY[k] * S->dY[k] + S->dZ[k] * S->dZ[k]);
        S->s[k] = (S->l[k] - L0[k]) / L0[k];
    }
    if (S->law == 1) {
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nB; b++) {
            double a = S->s0[(size_t)b * S->s0_stride], c = S->sc[(size_t)b * S->sc_stride];
            if (S->BONDTYPE[b] == 0 && S->s[b] > a) S->flag[b] = 1;
            double cand = (((S->s[b] - a) / S->s[b]) * (c / (c - a))) * S->flag[b];
            if (cand == cand && cand > S->H[b]) S->H[b] = cand;
            if (S->H[b] > 1) S->H[b] = 1;
        }
    } else if (S->law == 3) {
#pragma omp parallel for schedule(static)
        for (int b = 0; b < nB; b++) S->bsf[b] = 0;
#pragma omp parallel
        {
#ifdef _OPENMP
            int t = omp_get_thread_num(), TT = omp_get_num_threads();
#else
            int t = 0, TT = 1;
#endif
            int lo, hi;
            static_block(nB, TT, t, &lo, &hi);
            or_calculatebsfexponential(S->s + lo, S->s0_full + lo, S->sc_full + lo, S->k, S->alpha, S->bsf + lo,
                                       S->H + lo, S->BONDTYPE + lo, S->flag + lo, hi - lo);
        }
    } else if (S->law == 2) {
#pragma omp parallel
        {
#ifdef _OPENMP
            int t = omp_get_thread_num(), TT = omp_get_num_threads();
#else
            int t = 0, TT = 1;
#endif
            int lo, hi;
            static_block(nB, TT, t, &lo, &hi);
            or_calculatebsftrilinear(S->s + lo, S->s0[0], S->s1, S->sc[0], S->H + lo, S->BONDTYPE + lo,
                                     S->flag + lo, hi - lo);
        }
    }
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nB; b++) {
        double p = S->fail[b] * S->c[b] * (1 - S->H[b]) * (S->s[b] - S->sp[b]) * S->V * S->vcf[b];
        S->fX[b] = p * (S->dX[b] / S->l[b]);
        S->fY[b] = p * (S->dY[b] / S->l[b]);
        S->fZ[b] = p * (S->dZ[b] / S->l[b]);
        if (S->fX[b] != S->fX[b]) S->fX[b] = 0.0;
        if (S->fY[b] != S->fY[b]) S->fY[b] = 0.0;
        if (S->fZ[b] != S->fZ[b]) S->fZ[b] = 0.0;
    }
    size_t nsc = (size_t)nN * T;
#pragma omp parallel for schedule(static)
    for (size_t q = 0; q < nsc; q++) {
        S->sXs[q] = 0;
        S->sYs[q] = 0;
        S->sZs[q] = 0;
    }
#pragma omp parallel for schedule(static)
    for (int q = 0; q < 3 * nN; q++) S->F[q] = 0;
#pragma omp parallel
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
#pragma omp for schedule(static)
        for (int k = 0; k < nB; k++) {
            int i = (int)BL[k], j = (int)BL[nB + k];
            S->sXs[(i - 1) + (size_t)nN * t] += S->fX[k];
            S->sXs[(j - 1) + (size_t)nN * t] -= S->fX[k];
            S->sYs[(i - 1) + (size_t)nN * t] += S->fY[k];
            S->sYs[(j - 1) + (size_t)nN * t] -= S->fY[k];
            S->sZs[(i - 1) + (size_t)nN * t] += S->fZ[k];
            S->sZs[(j - 1) + (size_t)nN * t] -= S->fZ[k];
        }
#pragma omp for schedule(static)
        for (int n = 0; n < nN; n++)
            for (int tt = 0; tt < T; tt++) {
                S->F[n] += S->sXs[n + (size_t)nN * tt];
                S->F[n + nN] += S->sYs[n + (size_t)nN * tt];
                S->F[n + 2 * nN] += S->sZs[n + (size_t)nN * tt];
            }
    }
#pragma omp parallel for schedule(static)
    for (int n = 0; n < nN; n++)
        for (int d = 0; d < 3; d++) {
            size_t q = n + (size_t)nN * d;
            S->a[q] = (S->F[q] - S->DAMPING * S->v[q]) / S->rho[n];
            if (S->cons[q] == 1) S->a[q] = 0;
            S->v[q] = S->v[q] + (S->a[q] * S->DT);
            S->u[q] = S->u[q] + (S->v[q] * S->DT);
            S->x[q] = S->X[q] + S->u[q];
        }
}

int or_sizeof_state(void) { return (int)sizeof(or_state); }
