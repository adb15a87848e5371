// bbpd_solver.hip -- layer 2: the fused, device-resident explicit step.
//
// One step restates the reference driver solvers/scripts/dynamic/dynamicsolverdisplacementcontrolled.m:71-121:
//   smoothstepdata (:76) -> deformed length + stretch (:79) -> softening law (:82-91) -> bond force (:94)
//   -> nodal force (:97-101) -> time integration (:105) -> contact x3 (:108-112) -> reaction (:115).
//
// Device layout
//   * Nodes live in "slots" ordered along a Morton curve of their coordinates, padded to a multiple of
//     2048.  Nodal arrays are structure-of-arrays FP64 (one array per component).
//   * Every bond is held from both ends in a node-centric table in SELL-64 form: the 64 slots of a
//     slice (one wavefront) store entry k of their rows side by side, so entry loads coalesce.  A
//     row lists its bonds in BONDLIST order, which is exactly the order the reference's serial nodal
//     sum (calculatenodalforcesMex.c:27-44) adds them to that node: the gather is deterministic and
//     reproduces it bit for bit with no atomics.
//   * Per entry: u32 word = neighbour slot | BONDTYPE << 28 | failed << 30 | softening flag << 31,
//     the history bondSofteningFactor (FP64, or FP32 in the mixed form), the bond coefficient
//     (BONDSTIFFNESS in FP64; c*V*vcf in FP32 for the mixed form) and optional per-bond s0 / sc.
//     Both copies of a bond see the same stretch and update their history identically.
//   * UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS are recomputed from the coordinates with the
//     reference formulas (buildhorizons.m:93-97, calculatevolumecorrectionfactors.m:16-34); set-up
//     checks bit for bit that the given arrays equal them.
//   * Positions are double-buffered (deformed coordinates in FP64 form, displacements in the mixed
//     form): a step reads buffer step&1 and writes the other, so the Euler-Cromer update of a node
//     runs in the same kernel that gathers its neighbours.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <cstring>
#include <numeric>
#include <string>
#include <vector>

#include "bbpd_internal.hpp"

namespace {

constexpr int BLOCK = 256;
constexpr int SLICE = 64;
constexpr int SLOT_ALIGN = 2048; // 8 XCDs x BLOCK
constexpr uint32_t NODE_MASK = 0x0FFFFFFFu;
constexpr uint32_t TYPE_SHIFT = 28;
constexpr uint32_t FAILED_BIT = 1u << 30;
constexpr uint32_t FLAG_BIT = 1u << 31;
constexpr int HIST_CAP = 1 << 20;
constexpr int N_HIST = 4; // reaction, static reaction, deflection, CMOD

struct DevScalars {
    long long step; // steps completed
    unsigned int done; // blocks of the step-end kernel finished this step
    double cn;
    double cn1, cn2;
    double reaction[BBPD_MAX_BODIES];
    double reaction_static;
};

struct Body {
    int *fam = nullptr; // slots, in family order
    int n = 0;
    double cx = 0, cz = 0, R = 0;
    int moves = 0;
};

struct Params {
    // nodes
    int nN, nSlots;
    const double *X0, *X1, *X2;
    double *p0[2], *p1[2], *p2[2]; // positions (x in FP64 form, u in mixed form), double-buffered
    double *u0, *u1, *u2;          // displacement (FP64 form only)
    double *v0, *v1, *v2;
    double *F0, *F1, *F2;
    double *Fp0, *Fp1, *Fp2; // ADR previous force
    const double *rho, *mass;
    const unsigned char *cons; // bit d set = DOF d constrained
    // entries
    const int *rowlen;
    const long long *slice_off;
    uint32_t *nbr;
    void *hist;
    const void *coef;
    const void *s0e, *sce;
    // scalars
    double V, horizon, hmr, hpr, two_r; // hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2*RADIJ
    double s0, s1, sc, k, alpha, denom_exp, sc_steel, ff;
    double DT, D;
    double *partials;
    DevScalars *ds;
};

// ------------------------------------------------------------------------------------------------
// Softening laws, per entry.  Each mirrors its reference routine; `w` carries the flags.
// ------------------------------------------------------------------------------------------------
template <int LAW, typename T>
__device__ inline void apply_law(T s, T s0, T sc, T s1, T kk, T alpha, T denom, T sc_steel, T ff, T &H, uint32_t &w)
{
    const uint32_t type = (w >> TYPE_SHIFT) & 3u;
    if (LAW == BBPD_LAW_LINEAR) {
        // calculatebondfailureMex.c:36-38
        if (!(w & FAILED_BIT) && ff == T(0)) {
            bool brk = (type == 0 && s > sc) || (type == 1 && s > T(3) * sc) || (type == 2 && s > sc_steel);
            if (brk) w |= FAILED_BIT;
        }
    } else if (LAW == BBPD_LAW_BILINEAR) {
        // calculatebondsofteningfactor.m:42-48.  With the flag clear the candidate is +-0 or NaN and
        // never raises H (kept in [0,1]), so the arithmetic only runs for flagged bonds.
        if (type == 0 && s > s0) w |= FLAG_BIT;
        if (w & FLAG_BIT) {
            T cand = ((s - s0) / s) * (sc / (sc - s0));
            if (cand == cand && cand > H) H = cand;
            if (H > T(1)) H = T(1);
        }
    } else if (LAW == BBPD_LAW_TRILINEAR) {
        // calculatebsftrilinear.m:80-112 (eta and beta fixed per call)
        if (type == 0) {
            const T beta = T(0.25);
            const T eta = s1 / s0;
            T bsf = T(0);
            if ((s > s0) && (s <= s1)) {
                bsf = (T(1) - ((eta - beta) / (eta - T(1)) * (s0 / s))) + ((T(1) - beta) / (eta - T(1)));
                w |= FLAG_BIT;
            } else if ((s > s1) && (s <= sc)) {
                bsf = T(1) - ((s0 * beta) / s) * ((sc - s) / (sc - s1));
            } else if (s > sc) {
                bsf = T(1);
            }
            if (bsf > H) H = bsf;
        }
        if (H != H) H = T(0);
    } else {
        // calculatebsfexponentialMex.c:35-58 with the per-step scratch bsf zeroed by the caller (:86)
        if (type == 0) {
            T bsf = T(0);
            if (s0 < s && s <= sc) {
                w |= FLAG_BIT;
                T numerator = T(1) - exp(-kk * (s - s0) / (sc - s0));
                T residual = alpha * (T(1) - (s - s0) / (sc - s0));
                bsf = T(1) - (s0 / s) * ((T(1) - (numerator / denom)) + residual) / (T(1) + alpha);
            } else if (s > sc) {
                bsf = T(1);
            }
            if (bsf > H) H = bsf;
            if (H > T(1)) H = T(1);
            if (H != H) H = T(0);
        }
    }
}

__device__ inline int remap_block(int b, int nb)
{
    // consecutive Morton slots stay on one XCD: hardware places block b on XCD b % 8
    int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

// ------------------------------------------------------------------------------------------------
// Bond kernel, plain FP64 form.  Thread = node slot.  Gathers the neighbours' deformed coordinates,
// updates every entry's law state, sums the row-oriented bond forces in BONDLIST order, then either
// applies the Euler-Cromer update (timeintegrationeulercromerMex.c:38-46) or, for ADR, stores F and
// the block's Rayleigh-quotient partial sums (calculatedampingcoefficient.m:61-85).
// ------------------------------------------------------------------------------------------------
template <int LAW, bool S0A, bool SCA, bool ADR>
__global__ __launch_bounds__(BLOCK) void k_bond_fp64(Params P)
{
    __shared__ double sh[BLOCK];
    const int b = remap_block(blockIdx.x, gridDim.x);
    const int n = b * BLOCK + threadIdx.x;
    const int cur = (int)(P.ds->step & 1);
    const double *__restrict__ x0 = P.p0[cur];
    const double *__restrict__ x1 = P.p1[cur];
    const double *__restrict__ x2 = P.p2[cur];
    const double *__restrict__ X0 = P.X0;
    const double *__restrict__ X1 = P.X1;
    const double *__restrict__ X2 = P.X2;
    double *__restrict__ hist = (double *)P.hist;
    const double *__restrict__ cc = (const double *)P.coef;
    const double *__restrict__ s0e = (const double *)P.s0e;
    const double *__restrict__ sce = (const double *)P.sce;

    double f0 = 0, f1 = 0, f2 = 0;
    const int len = P.rowlen[n];
    const double xn0 = x0[n], xn1 = x1[n], xn2 = x2[n];
    const double Xn0 = X0[n], Xn1 = X1[n], Xn2 = X2[n];
    long long e = P.slice_off[n >> 6] + (n & (SLICE - 1));
    // software pipeline: entry kk+1's stream words and neighbour gathers are in flight while
    // entry kk is computed (loads only move; the arithmetic and its order are unchanged)
    uint32_t wN = 0;
    double HN = 0, cN = 0, s0N = 0, scN = 0, xm0N = 0, xm1N = 0, xm2N = 0, Xm0N = 0, Xm1N = 0, Xm2N = 0;
    if (len > 0) {
        wN = P.nbr[e];
        HN = hist[e];
        cN = cc[e];
        if (S0A) s0N = s0e[e];
        if (SCA) scN = sce[e];
        const int mN = (int)(wN & NODE_MASK);
        xm0N = x0[mN], xm1N = x1[mN], xm2N = x2[mN];
        Xm0N = X0[mN], Xm1N = X1[mN], Xm2N = X2[mN];
    }
    for (int kk = 0; kk < len; kk++, e += SLICE) {
        uint32_t w = wN;
        const uint32_t w0 = w;
        double H = HN;
        const double ccur = cN, s0c = s0N, scc = scN;
        const double xm0 = xm0N, xm1 = xm1N, xm2 = xm2N, Xm0 = Xm0N, Xm1 = Xm1N, Xm2 = Xm2N;
        if (kk + 1 < len) {
            const long long e1 = e + SLICE;
            wN = P.nbr[e1];
            HN = hist[e1];
            cN = cc[e1];
            if (S0A) s0N = s0e[e1];
            if (SCA) scN = sce[e1];
            const int mN = (int)(wN & NODE_MASK);
            xm0N = x0[mN], xm1N = x1[mN], xm2N = x2[mN];
            Xm0N = X0[mN], Xm1N = X1[mN], Xm2N = X2[mN];
        }
        // calculatedeformedlengthMex.c:37-43, row-oriented (x_m - x_n): exact negation of the
        // reference's x_j - x_i at the j end, so |.|, s and the force magnitude are identical
        const double dx = xm0 - xn0, dy = xm1 - xn1, dz = xm2 - xn2;
        const double l = sqrt(dx * dx + dy * dy + dz * dz);
        const double ex = Xn0 - Xm0, ey = Xn1 - Xm1, ez = Xn2 - Xm2;
        const double L0 = sqrt((ex * ex + ey * ey) + ez * ez);
        const double s = (l - L0) / L0;
        const double H0 = H;
        const double s0 = S0A ? s0c : P.s0;
        const double sc = SCA ? scc : P.sc;
        apply_law<LAW, double>(s, s0, sc, P.s1, P.k, P.alpha, P.denom_exp, P.sc_steel, P.ff, H, w);
        if (w != w0) P.nbr[e] = w;
        if (__double_as_longlong(H) != __double_as_longlong(H0)) hist[e] = H;
        // calculatevolumecorrectionfactors.m:20-32
        double vcf;
        if (L0 <= P.hmr) vcf = 1;
        else if ((L0 > P.hmr) && (L0 <= P.horizon)) vcf = (P.hpr - L0) / P.two_r;
        else vcf = 0;
        // calculatebondforcesMex.c:34-40 with stretchPlastic = 0 (s - 0 == s exactly)
        const double fail = (w & FAILED_BIT) ? 0.0 : 1.0;
        const double p = fail * ccur * (1 - H) * s * P.V * vcf;
        double gx = p * (dx / l), gy = p * (dy / l), gz = p * (dz / l);
        if (gx != gx) gx = 0.0;
        if (gy != gy) gy = 0.0;
        if (gz != gz) gz = 0.0;
        f0 += gx;
        f1 += gy;
        f2 += gz;
    }
    const bool real = n < P.nN;
    if (!ADR) {
        if (!real) return;
        const int nxt = cur ^ 1;
        P.F0[n] = f0;
        P.F1[n] = f1;
        P.F2[n] = f2;
        const unsigned char cf = P.cons[n];
        const double r = P.rho[n];
        double *uu[3] = {P.u0, P.u1, P.u2};
        double *vv[3] = {P.v0, P.v1, P.v2};
        double *xo[3] = {P.p0[nxt], P.p1[nxt], P.p2[nxt]};
        const double Xn[3] = {Xn0, Xn1, Xn2};
        const double ff[3] = {f0, f1, f2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double a = (ff[d] - P.D * vv[d][n]) / r;
            if (cf & (1u << d)) a = 0;
            const double nv = vv[d][n] + (a * P.DT);
            const double nu = uu[d][n] + (nv * P.DT);
            vv[d][n] = nv;
            uu[d][n] = nu;
            xo[d][n] = Xn[d] + nu;
        }
    } else {
        double c1 = 0, c2 = 0;
        if (real) {
            P.F0[n] = f0;
            P.F1[n] = f1;
            P.F2[n] = f2;
            const double mi = P.mass[n];
            const double ff[3] = {f0, f1, f2};
            const double fp[3] = {P.Fp0[n], P.Fp1[n], P.Fp2[n]};
            const double vp[3] = {P.v0[n], P.v1[n], P.v2[n]};
            const double uq[3] = {P.u0[n], P.u1[n], P.u2[n]};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (vp[d] != 0) c1 = c1 - (uq[d] * uq[d]) * ((ff[d] / mi - fp[d] / mi) / (P.DT * vp[d]));
                c2 = c2 + uq[d] * uq[d];
            }
        }
        const double s1 = bbpd::block_sum_fixed<BLOCK>(c1, sh);
        const double s2 = bbpd::block_sum_fixed<BLOCK>(c2, sh);
        if (threadIdx.x == 0) {
            P.partials[2 * b] = s1;
            P.partials[2 * b + 1] = s2;
        }
    }
}

// ------------------------------------------------------------------------------------------------
// Bond kernel, FP32-mixed form.  Relative displacement eta = u_m - u_n and bond vector
// xi = X_m - X_n are formed in FP64 and rounded; the stretch uses the cancellation-free form
//   s = (2 xi.eta + eta.eta) / (|xi| (|xi + eta| + |xi|)),
// the law and the force run in FP32, the row sum and all nodal state stay FP64.
// ------------------------------------------------------------------------------------------------
template <int LAW, bool S0A, bool SCA, bool ADR>
__global__ __launch_bounds__(BLOCK) void k_bond_mixed(Params P)
{
    __shared__ double sh[BLOCK];
    const int b = remap_block(blockIdx.x, gridDim.x);
    const int n = b * BLOCK + threadIdx.x;
    const int cur = (int)(P.ds->step & 1);
    const double *__restrict__ u0 = P.p0[cur];
    const double *__restrict__ u1 = P.p1[cur];
    const double *__restrict__ u2 = P.p2[cur];
    const double *__restrict__ X0 = P.X0;
    const double *__restrict__ X1 = P.X1;
    const double *__restrict__ X2 = P.X2;
    float *__restrict__ hist = (float *)P.hist;
    const float *__restrict__ kap = (const float *)P.coef;
    const float *__restrict__ s0e = (const float *)P.s0e;
    const float *__restrict__ sce = (const float *)P.sce;
    const float s0c = (float)P.s0, scc = (float)P.sc, s1c = (float)P.s1, kc = (float)P.k, ac = (float)P.alpha;
    const float dc = (float)P.denom_exp, sstc = (float)P.sc_steel, ffc = (float)P.ff;

    double f0 = 0, f1 = 0, f2 = 0;
    const int len = P.rowlen[n];
    const double un0 = u0[n], un1 = u1[n], un2 = u2[n];
    const double Xn0 = X0[n], Xn1 = X1[n], Xn2 = X2[n];
    long long e = P.slice_off[n >> 6] + (n & (SLICE - 1));
    // software pipeline as in the FP64 kernel: entry kk+1 is loaded while entry kk is computed
    uint32_t wN = 0;
    float HN = 0.f, kN = 0.f, s0N = 0.f, scN = 0.f;
    double um0N = 0, um1N = 0, um2N = 0, Xm0N = 0, Xm1N = 0, Xm2N = 0;
    if (len > 0) {
        wN = P.nbr[e];
        HN = hist[e];
        kN = kap[e];
        if (S0A) s0N = s0e[e];
        if (SCA) scN = sce[e];
        const int mN = (int)(wN & NODE_MASK);
        um0N = u0[mN], um1N = u1[mN], um2N = u2[mN];
        Xm0N = X0[mN], Xm1N = X1[mN], Xm2N = X2[mN];
    }
    for (int kk = 0; kk < len; kk++, e += SLICE) {
        uint32_t w = wN;
        const uint32_t w0 = w;
        float H = HN;
        const float kcur = kN, s0cur = s0N, sccur = scN;
        const double um0 = um0N, um1 = um1N, um2 = um2N, Xm0 = Xm0N, Xm1 = Xm1N, Xm2 = Xm2N;
        if (kk + 1 < len) {
            const long long e1 = e + SLICE;
            wN = P.nbr[e1];
            HN = hist[e1];
            kN = kap[e1];
            if (S0A) s0N = s0e[e1];
            if (SCA) scN = sce[e1];
            const int mN = (int)(wN & NODE_MASK);
            um0N = u0[mN], um1N = u1[mN], um2N = u2[mN];
            Xm0N = X0[mN], Xm1N = X1[mN], Xm2N = X2[mN];
        }
        const float xi0 = (float)(Xm0 - Xn0), xi1 = (float)(Xm1 - Xn1), xi2 = (float)(Xm2 - Xn2);
        const float et0 = (float)(um0 - un0), et1 = (float)(um1 - un1), et2 = (float)(um2 - un2);
        const float y0 = xi0 + et0, y1 = xi1 + et1, y2 = xi2 + et2;
        const float L0 = sqrtf(xi0 * xi0 + xi1 * xi1 + xi2 * xi2);
        const float l = sqrtf(y0 * y0 + y1 * y1 + y2 * y2);
        const float num = 2.0f * (xi0 * et0 + xi1 * et1 + xi2 * et2) + (et0 * et0 + et1 * et1 + et2 * et2);
        const float s = num / (L0 * (l + L0));
        const float H0 = H;
        const float s0 = S0A ? s0cur : s0c;
        const float sc = SCA ? sccur : scc;
        apply_law<LAW, float>(s, s0, sc, s1c, kc, ac, dc, sstc, ffc, H, w);
        if (w != w0) P.nbr[e] = w;
        if (__float_as_uint(H) != __float_as_uint(H0)) hist[e] = H;
        const float fail = (w & FAILED_BIT) ? 0.0f : 1.0f;
        const float p = fail * kcur * (1.0f - H) * s / l;
        float gx = p * y0, gy = p * y1, gz = p * y2;
        if (gx != gx) gx = 0.0f;
        if (gy != gy) gy = 0.0f;
        if (gz != gz) gz = 0.0f;
        f0 += (double)gx;
        f1 += (double)gy;
        f2 += (double)gz;
    }
    const bool real = n < P.nN;
    if (!ADR) {
        if (!real) return;
        const int nxt = cur ^ 1;
        P.F0[n] = f0;
        P.F1[n] = f1;
        P.F2[n] = f2;
        const unsigned char cf = P.cons[n];
        const double r = P.rho[n];
        double *vv[3] = {P.v0, P.v1, P.v2};
        double *uo[3] = {P.p0[nxt], P.p1[nxt], P.p2[nxt]};
        const double un[3] = {un0, un1, un2};
        const double ff[3] = {f0, f1, f2};
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double a = (ff[d] - P.D * vv[d][n]) / r;
            if (cf & (1u << d)) a = 0;
            const double nv = vv[d][n] + (a * P.DT);
            vv[d][n] = nv;
            uo[d][n] = un[d] + (nv * P.DT);
        }
    } else {
        double c1 = 0, c2 = 0;
        if (real) {
            P.F0[n] = f0;
            P.F1[n] = f1;
            P.F2[n] = f2;
            const double mi = P.mass[n];
            const double ff[3] = {f0, f1, f2};
            const double fp[3] = {P.Fp0[n], P.Fp1[n], P.Fp2[n]};
            const double vp[3] = {P.v0[n], P.v1[n], P.v2[n]};
            const double uq[3] = {un0, un1, un2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (vp[d] != 0) c1 = c1 - (uq[d] * uq[d]) * ((ff[d] / mi - fp[d] / mi) / (P.DT * vp[d]));
                c2 = c2 + uq[d] * uq[d];
            }
        }
        const double s1 = bbpd::block_sum_fixed<BLOCK>(c1, sh);
        const double s2 = bbpd::block_sum_fixed<BLOCK>(c2, sh);
        if (threadIdx.x == 0) {
            P.partials[2 * b] = s1;
            P.partials[2 * b + 1] = s2;
        }
    }
}

// ADR: damping coefficient from the block partials (calculatedampingcoefficient.m:90-112)
__global__ __launch_bounds__(BLOCK) void k_adr_cn(Params P, int nPart)
{
    __shared__ double sh[BLOCK];
    double c1 = 0, c2 = 0;
    for (int q = threadIdx.x; q < nPart; q += BLOCK) {
        c1 = c1 + P.partials[2 * q];
        c2 = c2 + P.partials[2 * q + 1];
    }
    const double s1 = bbpd::block_sum_fixed<BLOCK>(c1, sh);
    const double s2 = bbpd::block_sum_fixed<BLOCK>(c2, sh);
    if (threadIdx.x == 0) {
        double cn = 0;
        if (s2 != 0) cn = ((s1 / s2) > 0) ? 2 * sqrt(s1 / s2) : 0;
        if (cn > 2) cn = 1.9;
        P.ds->cn = cn;
        P.ds->cn1 = s1;
        P.ds->cn2 = s2;
    }
}

// ADR: leapfrog update (timeintegrationADR.m:53-80), then x = X + u and F_prev = F
// (the hand-over of dynamicsolverdisplacementcontrolledADR.m:153-155)
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_adr_update(Params P)
{
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    if (n >= P.nN) return;
    const int cur = (int)(P.ds->step & 1), nxt = cur ^ 1;
    const int it = (int)(P.ds->step + 1);
    const double cn = P.ds->cn, DT = P.DT;
    const double minv = 1 / P.mass[n];
    const unsigned char cf = P.cons[n];
    double *vv[3] = {P.v0, P.v1, P.v2};
    double *FF[3] = {P.F0, P.F1, P.F2};
    double *FP[3] = {P.Fp0, P.Fp1, P.Fp2};
    const double *XX[3] = {P.X0, P.X1, P.X2};
    double *uu[3] = {P.u0, P.u1, P.u2};
    double *pc[3] = {P.p0[cur], P.p1[cur], P.p2[cur]};
    double *po[3] = {P.p0[nxt], P.p1[nxt], P.p2[nxt]};
#pragma unroll
    for (int d = 0; d < 3; d++) {
        const double f = FF[d][n];
        double nv;
        if (it == 1) nv = 0.5 * DT * minv * f;
        else nv = (((2 - cn * DT) * vv[d][n]) + (2 * DT * minv * f)) / (2 + cn * DT);
        const double uold = MIXED ? pc[d][n] : uu[d][n];
        double nu = uold + (DT * nv);
        if (cf & (1u << d)) {
            nu = 0;
            nv = 0;
        }
        vv[d][n] = nv;
        if (MIXED) {
            po[d][n] = nu;
        } else {
            uu[d][n] = nu;
            po[d][n] = XX[d][n] + nu;
        }
        FP[d][n] = f;
    }
}

// ------------------------------------------------------------------------------------------------
// Step end: contact with every rigid body in order (calculatecontactforce.m:62-121), the load
// schedule (smoothstepdata.m:40-42 or a staircase of increments), the observables of the step and
// the step counter.  One block; family members are projected in parallel and each body's reaction
// is summed in family order by one thread.
// ------------------------------------------------------------------------------------------------
struct StepEnd {
    int nBodies;
    const int *fam[BBPD_MAX_BODIES];
    int nFam[BBPD_MAX_BODIES];
    double cx[BBPD_MAX_BODIES], cz[BBPD_MAX_BODIES], R[BBPD_MAX_BODIES];
    int moves[BBPD_MAX_BODIES];
    int schedule, nTimeSteps, ipi;
    double applied, increment;
    int ref_slot, cmod1, cmod2; // -1 = none
    double *hist;               // HIST_CAP x N_HIST ring
};

__device__ inline double schedule_disp(const StepEnd &S, long long i)
{
    if (S.schedule == 0) {
        // smoothstepdata(i, 1, N, 0, a): total displacement at step i
        double xi = ((double)i - 1) / ((double)S.nTimeSteps - 1);
        return 0 + (S.applied - 0) * (xi * xi * xi) * (10 - 15 * xi + 6 * (xi * xi));
    }
    long long inc = (i + S.ipi - 1) / S.ipi;
    return S.increment * (double)inc;
}

template <bool MIXED>
__global__ __launch_bounds__(1024) void k_step_end(Params P, StepEnd S)
{
    // one block per contact body (families are disjoint); the last block to finish records the
    // step's observables and advances the step counter
    __shared__ double vals[1024];
    __shared__ int wofs[17];
    __shared__ int last_block;
    const int nxt = (int)((P.ds->step + 1) & 1);
    const long long it = P.ds->step + 1;
    double *pos[3] = {P.p0[nxt], P.p1[nxt], P.p2[nxt]};
    double *uu[3] = {P.u0, P.u1, P.u2};
    double *vv[3] = {P.v0, P.v1, P.v2};
    const double *XX[3] = {P.X0, P.X1, P.X2};
    const int bdy = blockIdx.x;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double total = 0, rs = 0;
    if (bdy < S.nBodies) {
        const double disp = schedule_disp(S, it);
        const double cx = S.cx[bdy], czm = S.cz[bdy] + (S.moves[bdy] ? disp : 0.0), R = S.R[bdy];
        const int nf = S.nFam[bdy];
        for (int base = 0; base < nf; base += blockDim.x) {
            const int t = base + threadIdx.x;
            bool hit = false;
            double d_z = 0;
            if (t < nf) {
                const int i = S.fam[bdy][t];
                double xc[3];
#pragma unroll
                for (int d = 0; d < 3; d++) xc[d] = MIXED ? XX[d][i] + pos[d][i] : pos[d][i];
                const double ddx = xc[0] - cx, ddz = xc[2] - czm;
                const double dist = sqrt((ddx * ddx) + (ddz * ddz));
                if (dist < R) {
                    hit = true;
                    xc[0] = cx + (ddx / dist) * R;
                    xc[2] = czm + (ddz / dist) * R;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double uo = MIXED ? pos[d][i] : uu[d][i];
                        const double vo = vv[d][i];
                        const double nu = xc[d] - XX[d][i];
                        const double nv = (nu - uo) / P.DT;
                        if (MIXED) {
                            pos[d][i] = nu;
                        } else {
                            uu[d][i] = nu;
                            pos[d][i] = xc[d];
                        }
                        vv[d][i] = nv;
                        if (d == 2) d_z = -1 * P.rho[i] * (nv - vo) / P.DT * P.V;
                    }
                }
            }
            // compact the members in contact, in family order, then sum them in that order
            // (the order of MATLAB's sum over dpenfz; members not in contact add nothing)
            const unsigned long long mask = __ballot(hit);
            if (lane == 0) wofs[wv + 1] = __popcll(mask);
            __syncthreads();
            if (threadIdx.x == 0) {
                wofs[0] = 0;
                for (int q = 1; q <= nw; q++) wofs[q] += wofs[q - 1];
            }
            __syncthreads();
            if (hit) vals[wofs[wv] + __popcll(mask & ((1ull << lane) - 1))] = d_z;
            __syncthreads();
            if (threadIdx.x == 0)
                for (int q = 0; q < wofs[nw]; q++) total = total + vals[q];
            __syncthreads();
        }
        if (bdy == 0) {
            // calculatereactionforceFast.m:3-5 over body 0's family: F_z * cellVolume (fixed tree)
            double part = 0;
            for (int t = threadIdx.x; t < nf; t += blockDim.x) part = part + P.F2[S.fam[0][t]] * P.V;
            rs = bbpd::block_sum_fixed<1024>(part, vals);
        }
    }
    if (threadIdx.x == 0) {
        if (bdy < S.nBodies) P.ds->reaction[bdy] = total;
        if (bdy == 0) P.ds->reaction_static = rs;
        __threadfence();
        const unsigned ticket = atomicAdd(&P.ds->done, 1u);
        last_block = (ticket == gridDim.x - 1);
    }
    __syncthreads();
    if (last_block && threadIdx.x == 0) {
        __threadfence();
        volatile DevScalars *vds = P.ds;
        const volatile double *uz = MIXED ? pos[2] : uu[2];
        const volatile double *ux = MIXED ? pos[0] : uu[0];
        const double defl = S.ref_slot >= 0 ? uz[S.ref_slot] : 0.0;
        const double cmod = (S.cmod1 >= 0 && S.cmod2 >= 0) ? ux[S.cmod2] - ux[S.cmod1] : 0.0;
        double *h = S.hist + (size_t)(it - 1) % HIST_CAP * N_HIST;
        h[0] = S.nBodies > 0 ? vds->reaction[0] : 0.0;
        h[1] = vds->reaction_static;
        h[2] = defl;
        h[3] = cmod;
        vds->done = 0u;
        vds->step = it;
    }
}

// observables over entries: softening flags, broken (H == 1 or failed), failed; each bond twice
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_count(Params P, unsigned long long *out)
{
    const int n = blockIdx.x * BLOCK + threadIdx.x;
    unsigned long long cs = 0, cb = 0, cf = 0;
    if (n < P.nSlots) {
        const int len = P.rowlen[n];
        long long e = P.slice_off[n >> 6] + (n & (SLICE - 1));
        for (int kk = 0; kk < len; kk++, e += SLICE) {
            uint32_t w = P.nbr[e];
            bool one = MIXED ? (((const float *)P.hist)[e] == 1.0f) : (((const double *)P.hist)[e] == 1.0);
            cs += (w & FLAG_BIT) ? 1 : 0;
            cf += (w & FAILED_BIT) ? 1 : 0;
            cb += ((w & FAILED_BIT) || one) ? 1 : 0;
        }
    }
    // wavefront-level integer reduction then one atomic per wave (integer sums are exact)
    for (int off = 32; off > 0; off >>= 1) {
        cs += __shfl_down(cs, off);
        cb += __shfl_down(cb, off);
        cf += __shfl_down(cf, off);
    }
    if ((threadIdx.x & 63) == 0) {
        atomicAdd(out + 0, cs);
        atomicAdd(out + 1, cb);
        atomicAdd(out + 2, cf);
    }
}

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z)
{
    auto spread = [](uint64_t v) {
        v &= 0x1fffff;
        v = (v | v << 32) & 0x1f00000000ffffull;
        v = (v | v << 16) & 0x1f0000ff0000ffull;
        v = (v | v << 8) & 0x100f00f00f00f00full;
        v = (v | v << 4) & 0x10c30c30c30c30c3ull;
        v = (v | v << 2) & 0x1249249249249249ull;
        return v;
    };
    return spread(x) | (spread(y) << 1) | (spread(z) << 2);
}

template <typename T>
int dalloc(T **p, size_t n, std::vector<void *> &owned)
{
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * n);
    if (e != hipSuccess) return bbpd::hip_fail(e, "hipMalloc (solver state)");
    owned.push_back(*p);
    return BBPD_OK;
}

} // namespace

struct bbpd_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = 0, law = 0, integrator = 0;
    bool s0arr = false, scarr = false;
    int nN = 0, nB = 0, nSlots = 0, nSlices = 0;
    long long nEntries = 0, nPadded = 0;
    std::vector<int> slot_of;               // original node (0-based) -> slot
    std::vector<long long> entry_i, entry_j; // bond -> entry index at its two ends
    std::vector<void *> owned;
    Params P{};
    StepEnd S{};
    DevScalars *ds = nullptr;
    unsigned long long *counts = nullptr;
    double *hist = nullptr;
    long long hist_read = 0; // steps already returned by bbpd_solver_history
    size_t device_bytes = 0;
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double t_acc_ms = 0;
    long long t_launches = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
};

namespace {

using KernelFn = void (*)(Params);

template <bool MIXED, int LAW, bool S0A, bool SCA, bool ADR>
KernelFn pick4()
{
    if (MIXED) return k_bond_mixed<LAW, S0A, SCA, ADR>;
    return k_bond_fp64<LAW, S0A, SCA, ADR>;
}

template <bool MIXED, int LAW>
KernelFn pick3(bool s0a, bool sca, bool adr)
{
    if (s0a) {
        if (sca) return adr ? pick4<MIXED, LAW, true, true, true>() : pick4<MIXED, LAW, true, true, false>();
        return adr ? pick4<MIXED, LAW, true, false, true>() : pick4<MIXED, LAW, true, false, false>();
    }
    if (sca) return adr ? pick4<MIXED, LAW, false, true, true>() : pick4<MIXED, LAW, false, true, false>();
    return adr ? pick4<MIXED, LAW, false, false, true>() : pick4<MIXED, LAW, false, false, false>();
}

template <bool MIXED>
KernelFn pick2(int law, bool s0a, bool sca, bool adr)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return pick3<MIXED, BBPD_LAW_LINEAR>(s0a, sca, adr);
    case BBPD_LAW_BILINEAR: return pick3<MIXED, BBPD_LAW_BILINEAR>(s0a, sca, adr);
    case BBPD_LAW_TRILINEAR: return pick3<MIXED, BBPD_LAW_TRILINEAR>(s0a, sca, adr);
    default: return pick3<MIXED, BBPD_LAW_EXPONENTIAL>(s0a, sca, adr);
    }
}

KernelFn pick_bond_kernel(const bbpd_solver *s)
{
    bool adr = s->integrator == BBPD_INTEGRATOR_ADR;
    if (s->precision == BBPD_PRECISION_MIXED) return pick2<true>(s->law, s->s0arr, s->scarr, adr);
    return pick2<false>(s->law, s->s0arr, s->scarr, adr);
}

int launch_step(bbpd_solver *s, bool allow_timing)
{
    const unsigned grid = (unsigned)(s->nSlots / BLOCK);
    KernelFn kb = pick_bond_kernel(s);
    const bool timed = allow_timing && s->timing;
    if (timed) {
        if (s->ev_used + 2 > (int)s->ev.size()) {
            // drain the pool: accumulate the recorded pairs
            BBPD_HIP(hipStreamSynchronize(s->stream));
            for (int q = 0; q + 1 < s->ev_used; q += 2) {
                float ms = 0;
                BBPD_HIP(hipEventElapsedTime(&ms, s->ev[q], s->ev[q + 1]));
                s->t_acc_ms += ms;
                s->t_launches++;
            }
            s->ev_used = 0;
        }
        BBPD_HIP(hipEventRecord(s->ev[s->ev_used], s->stream));
    }
    hipLaunchKernelGGL(kb, dim3(grid), dim3(BLOCK), 0, s->stream, s->P);
    BBPD_HIP(hipGetLastError());
    if (timed) {
        BBPD_HIP(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
        s->ev_used += 2;
    }
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    if (s->integrator == BBPD_INTEGRATOR_ADR) {
        hipLaunchKernelGGL(k_adr_cn, dim3(1), dim3(BLOCK), 0, s->stream, s->P, (int)grid);
        BBPD_HIP(hipGetLastError());
        const unsigned g2 = bbpd::grid_for(s->nN, BLOCK);
        if (mixed) hipLaunchKernelGGL(k_adr_update<true>, dim3(g2), dim3(BLOCK), 0, s->stream, s->P);
        else hipLaunchKernelGGL(k_adr_update<false>, dim3(g2), dim3(BLOCK), 0, s->stream, s->P);
        BBPD_HIP(hipGetLastError());
    }
    const unsigned nb = (unsigned)std::max(s->S.nBodies, 1);
    if (mixed) hipLaunchKernelGGL(k_step_end<true>, dim3(nb), dim3(1024), 0, s->stream, s->P, s->S);
    else hipLaunchKernelGGL(k_step_end<false>, dim3(nb), dim3(1024), 0, s->stream, s->P, s->S);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

bool bits_equal(double a, double b)
{
    uint64_t x, y;
    std::memcpy(&x, &a, 8);
    std::memcpy(&y, &b, 8);
    return x == y;
}

} // namespace

extern "C" int bbpd_solver_create(const bbpd_problem *pr, int precision, int device, bbpd_solver **out)
{
    BBPD_REQUIRE(pr && out, "null argument");
    *out = nullptr;
    const int nN = pr->nNodes, nB = pr->nBonds;
    BBPD_REQUIRE(nN > 0 && nB >= 0, "need nNodes > 0 and nBonds >= 0");
    BBPD_REQUIRE((long long)nN < (long long)NODE_MASK - SLOT_ALIGN, "too many nodes for 28-bit slots");
    BBPD_REQUIRE(pr->undeformedCoordinates && pr->BONDLIST && pr->UNDEFORMEDLENGTH && pr->VOLUMECORRECTIONFACTORS &&
                     pr->BONDSTIFFNESS && pr->DENSITY,
                 "missing input array");
    BBPD_REQUIRE(precision == BBPD_PRECISION_FP64 || precision == BBPD_PRECISION_MIXED, "unknown precision");
    BBPD_REQUIRE(pr->law >= BBPD_LAW_LINEAR && pr->law <= BBPD_LAW_EXPONENTIAL, "unknown law");
    BBPD_REQUIRE(pr->integrator == BBPD_INTEGRATOR_EULER_CROMER || pr->integrator == BBPD_INTEGRATOR_ADR,
                 "unknown integrator");
    BBPD_REQUIRE(pr->integrator != BBPD_INTEGRATOR_ADR || pr->massVector, "ADR needs massVector");
    BBPD_REQUIRE(pr->nBodies >= 0 && pr->nBodies <= BBPD_MAX_BODIES, "too many contact bodies");
    BBPD_REQUIRE(pr->law != BBPD_LAW_TRILINEAR || (!pr->s0 && !pr->sc), "trilinear law takes scalar s0, s1, sc");
    BBPD_REQUIRE(pr->law != BBPD_LAW_LINEAR || !pr->sc, "linear law takes a scalar critical stretch");
    BBPD_REQUIRE(pr->law != BBPD_LAW_LINEAR || pr->failureFunctionality == 0 || pr->failureFunctionality == 1,
                 "failureFunctionality must be 0 or 1");
    BBPD_REQUIRE(pr->schedule == 0 ? pr->nTimeSteps > 1 : pr->itersPerIncrement > 0, "bad load schedule");
    BBPD_REQUIRE(pr->DT > 0, "DT must be positive");

    int ndev = 0;
    BBPD_HIP(hipGetDeviceCount(&ndev));
    BBPD_REQUIRE(device >= 0 && device < ndev, "device ordinal out of range");
    BBPD_HIP(hipSetDevice(device));

    auto *s = new bbpd_solver();
    auto fail = [&](int rc) {
        bbpd_solver_destroy(s);
        return rc;
    };
    s->device = device;
    s->precision = precision;
    s->law = pr->law;
    s->integrator = pr->integrator;
    s->nN = nN;
    s->nB = nB;
    const bool mixed = precision == BBPD_PRECISION_MIXED;
    const double *X = pr->undeformedCoordinates;
    const double *BL = pr->BONDLIST;

    // ---- check L0 and vcf against their formulas (they are recomputed on the device)
    const double horizon = pr->horizon, RADIJ = pr->RADIJ;
    const double hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2 * RADIJ;
    std::vector<int> bi(nB), bj(nB);
    for (int k = 0; k < nB; k++) {
        double di = BL[k], dj = BL[(size_t)nB + k];
        int i = (int)di, j = (int)dj;
        if (di != (double)i || dj != (double)j || i < 1 || j < 1 || i > nN || j > nN || i == j) {
            bbpd::set_error("BONDLIST row " + std::to_string(k + 1) + " is not a pair of distinct ids in 1..nNodes");
            return fail(BBPD_ERR_ARG);
        }
        bi[k] = i - 1;
        bj[k] = j - 1;
        double ex = X[i - 1] - X[j - 1], ey = X[i - 1 + nN] - X[j - 1 + nN], ez = X[i - 1 + 2 * (size_t)nN] - X[j - 1 + 2 * (size_t)nN];
        double L0 = std::sqrt((ex * ex + ey * ey) + ez * ez);
        if (!bits_equal(L0, pr->UNDEFORMEDLENGTH[k])) {
            bbpd::set_error("UNDEFORMEDLENGTH(" + std::to_string(k + 1) +
                            ") differs from the length recomputed from undeformedCoordinates (buildhorizons.m:93-97)");
            return fail(BBPD_ERR_UNSUPPORTED);
        }
        double vcf;
        if (L0 <= hmr) vcf = 1;
        else if ((L0 > hmr) && (L0 <= horizon)) vcf = (hpr - L0) / two_r;
        else vcf = 0;
        if (!bits_equal(vcf, pr->VOLUMECORRECTIONFACTORS[k])) {
            bbpd::set_error("VOLUMECORRECTIONFACTORS(" + std::to_string(k + 1) +
                            ") differs from calculatevolumecorrectionfactors(L0, horizon, RADIJ)");
            return fail(BBPD_ERR_UNSUPPORTED);
        }
        if (pr->fail && !(pr->fail[k] == 0 || pr->fail[k] == 1)) {
            bbpd::set_error("fail must hold 0 or 1");
            return fail(BBPD_ERR_ARG);
        }
        if (pr->BONDTYPE && !(pr->BONDTYPE[k] == 0 || pr->BONDTYPE[k] == 1 || pr->BONDTYPE[k] == 2)) {
            bbpd::set_error("BONDTYPE must hold 0, 1 or 2");
            return fail(BBPD_ERR_ARG);
        }
    }

    // ---- slots.  Coordinates are quantised by the shortest bond (lattice indices on a regular
    // lattice).  Default order: tiles of 64 x 2 x 2 nodes (64 along the longest axis) visited in
    // Morton order, lines of 64 inside a tile, so a wavefront is 64 consecutive nodes of a lattice
    // line and the k-th neighbours of its lanes are again 64 consecutive slots (coalesced gathers).
    // BBPD_SLOT_ORDER=morton selects a plain Morton order of the nodes instead.
    double xmin[3] = {1e300, 1e300, 1e300}, xmax[3] = {-1e300, -1e300, -1e300};
    for (int d = 0; d < 3; d++)
        for (int n = 0; n < nN; n++) {
            xmin[d] = std::min(xmin[d], X[n + (size_t)nN * d]);
            xmax[d] = std::max(xmax[d], X[n + (size_t)nN * d]);
        }
    double hq = 1e300;
    for (int k = 0; k < nB; k++) hq = std::min(hq, pr->UNDEFORMEDLENGTH[k]);
    if (!(hq > 0) || hq > 1e299) hq = 1;
    int ax[3] = {0, 1, 2}; // ax[0] = longest axis
    std::sort(ax, ax + 3, [&](int a, int b) { return (xmax[a] - xmin[a]) > (xmax[b] - xmin[b]); });
    const char *ord_env = std::getenv("BBPD_SLOT_ORDER");
    const bool morton_only = ord_env && std::string(ord_env) == "morton";
    std::vector<uint64_t> code(nN);
    for (int n = 0; n < nN; n++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) {
            double t = std::floor((X[n + (size_t)nN * ax[d]] - xmin[ax[d]]) / hq + 0.5);
            q[d] = (uint32_t)std::min(std::max(t, 0.0), 2097151.0);
        }
        if (morton_only) {
            code[n] = morton3(q[0], q[1], q[2]);
        } else {
            uint64_t tile = morton3(q[0] >> 6, q[1] >> 1, q[2] >> 1) & ((1ull << 55) - 1);
            uint64_t in = ((uint64_t)((q[2] & 1u) * 2u + (q[1] & 1u)) << 6) | (q[0] & 63u);
            code[n] = (tile << 8) | in;
        }
    }
    std::vector<int> order(nN);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });
    const int nSlots = (int)(((long long)nN + SLOT_ALIGN - 1) / SLOT_ALIGN * SLOT_ALIGN);
    s->nSlots = nSlots;
    s->nSlices = nSlots / SLICE;
    s->slot_of.assign(nN, 0);
    for (int q = 0; q < nN; q++) s->slot_of[order[q]] = q;

    // ---- rows in BONDLIST order
    std::vector<int> rowlen(nSlots, 0);
    for (int k = 0; k < nB; k++) {
        rowlen[s->slot_of[bi[k]]]++;
        rowlen[s->slot_of[bj[k]]]++;
    }
    std::vector<long long> slice_off(s->nSlices + 1, 0);
    for (int sl = 0; sl < s->nSlices; sl++) {
        int w = 0;
        for (int t = 0; t < SLICE; t++) w = std::max(w, rowlen[sl * SLICE + t]);
        slice_off[sl + 1] = slice_off[sl] + (long long)w * SLICE;
    }
    const long long nE = slice_off[s->nSlices];
    s->nEntries = 2LL * nB;
    s->nPadded = nE;
    std::vector<int> fillpos(nSlots, 0);
    s->entry_i.assign(nB, 0);
    s->entry_j.assign(nB, 0);
    auto entry_at = [&](int slot, int kk) { return slice_off[slot / SLICE] + (long long)kk * SLICE + (slot % SLICE); };
    std::vector<uint32_t> h_nbr(nE, 0);
    for (int k = 0; k < nB; k++) {
        int si = s->slot_of[bi[k]], sj = s->slot_of[bj[k]];
        uint32_t type = pr->BONDTYPE ? (uint32_t)pr->BONDTYPE[k] : 0u;
        uint32_t flags = (type << TYPE_SHIFT) | ((pr->fail && pr->fail[k] == 0) ? FAILED_BIT : 0u);
        long long ei = entry_at(si, fillpos[si]++);
        long long ej = entry_at(sj, fillpos[sj]++);
        h_nbr[ei] = (uint32_t)sj | flags;
        h_nbr[ej] = (uint32_t)si | flags;
        s->entry_i[k] = ei;
        s->entry_j[k] = ej;
    }
    // padding entries point at their own slot (never read: rows stop at rowlen)
    for (int sl = 0; sl < s->nSlices; sl++)
        for (int t = 0; t < SLICE; t++) {
            int slot = sl * SLICE + t;
            long long w = (slice_off[sl + 1] - slice_off[sl]) / SLICE;
            for (long long kk = rowlen[slot]; kk < w; kk++) h_nbr[entry_at(slot, (int)kk)] = (uint32_t)slot;
        }

    // ---- per-entry values
    s->s0arr = pr->s0 != nullptr;
    s->scarr = pr->sc != nullptr;
    const double V = pr->cellVolume;
    std::vector<double> h_hist64, h_coef64, h_s064, h_sc64;
    std::vector<float> h_hist32, h_coef32, h_s032, h_sc32;
    if (mixed) {
        h_hist32.assign(nE, 0.f);
        h_coef32.assign(nE, 0.f);
        if (s->s0arr) h_s032.assign(nE, 0.f);
        if (s->scarr) h_sc32.assign(nE, 0.f);
    } else {
        h_hist64.assign(nE, 0.0);
        h_coef64.assign(nE, 0.0);
        if (s->s0arr) h_s064.assign(nE, 0.0);
        if (s->scarr) h_sc64.assign(nE, 0.0);
    }
    for (int k = 0; k < nB; k++) {
        long long ends[2] = {s->entry_i[k], s->entry_j[k]};
        for (long long e : ends) {
            if (mixed) {
                h_coef32[e] = (float)(pr->BONDSTIFFNESS[k] * V * pr->VOLUMECORRECTIONFACTORS[k]);
                if (s->s0arr) h_s032[e] = (float)pr->s0[k];
                if (s->scarr) h_sc32[e] = (float)pr->sc[k];
            } else {
                h_coef64[e] = pr->BONDSTIFFNESS[k];
                if (s->s0arr) h_s064[e] = pr->s0[k];
                if (s->scarr) h_sc64[e] = pr->sc[k];
            }
        }
    }

    // ---- device allocations
    BBPD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    Params &P = s->P;
    P.nN = nN;
    P.nSlots = nSlots;
    int rc = 0;
    double *Xd, *pos[2], *ud = nullptr, *vd, *Fd, *Fpd = nullptr, *rhod, *massd = nullptr;
    unsigned char *consd;
    if ((rc = dalloc(&Xd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    for (int q = 0; q < 2; q++)
        if ((rc = dalloc(&pos[q], 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if (!mixed && (rc = dalloc(&ud, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&vd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&Fd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if (pr->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&Fpd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&rhod, nSlots, s->owned))) return fail(rc);
    if (pr->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&massd, nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&consd, nSlots, s->owned))) return fail(rc);
    int *rowlend;
    long long *soffd;
    uint32_t *nbrd;
    if ((rc = dalloc(&rowlend, nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&soffd, s->nSlices + 1, s->owned))) return fail(rc);
    if ((rc = dalloc(&nbrd, nE, s->owned))) return fail(rc);
    void *histd, *coefd, *s0d = nullptr, *scd = nullptr;
    const size_t ew = mixed ? 4 : 8;
    if ((rc = dalloc((char **)&histd, nE * ew, s->owned))) return fail(rc);
    if ((rc = dalloc((char **)&coefd, nE * ew, s->owned))) return fail(rc);
    if (s->s0arr && (rc = dalloc((char **)&s0d, nE * ew, s->owned))) return fail(rc);
    if (s->scarr && (rc = dalloc((char **)&scd, nE * ew, s->owned))) return fail(rc);
    double *partd;
    if ((rc = dalloc(&partd, 2 * (size_t)(nSlots / BLOCK) + 2, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->ds, 1, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->counts, 4, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->hist, (size_t)HIST_CAP * N_HIST, s->owned))) return fail(rc);

    // ---- host staging in slot order
    {
        std::vector<double> hX(3 * (size_t)nSlots, 0.0), hx(3 * (size_t)nSlots, 0.0), zero(3 * (size_t)nSlots, 0.0);
        std::vector<double> hrho(nSlots, 1.0), hm(nSlots, 1.0);
        std::vector<unsigned char> hc(nSlots, 0);
        for (int n = 0; n < nN; n++) {
            int q = s->slot_of[n];
            for (int d = 0; d < 3; d++) {
                hX[q + (size_t)nSlots * d] = X[n + (size_t)nN * d];
                hx[q + (size_t)nSlots * d] = X[n + (size_t)nN * d];
                if (pr->CONSTRAINTFLAG && pr->CONSTRAINTFLAG[n + (size_t)nN * d] == 1) hc[q] |= (unsigned char)(1u << d);
            }
            hrho[q] = pr->DENSITY[n];
            if (pr->massVector) hm[q] = pr->massVector[n];
        }
        BBPD_HIP(hipMemcpy(Xd, hX.data(), sizeof(double) * hX.size(), hipMemcpyHostToDevice));
        // FP64 form: positions start at deformedCoordinates = undeformedCoordinates
        // (dynamicsolverdisplacementcontrolled.m:33); mixed form: displacements start at 0
        for (int q = 0; q < 2; q++)
            BBPD_HIP(hipMemcpy(pos[q], mixed ? zero.data() : hx.data(), sizeof(double) * hx.size(), hipMemcpyHostToDevice));
        if (ud) BBPD_HIP(hipMemcpy(ud, zero.data(), sizeof(double) * zero.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(vd, zero.data(), sizeof(double) * zero.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(Fd, zero.data(), sizeof(double) * zero.size(), hipMemcpyHostToDevice));
        if (Fpd) BBPD_HIP(hipMemcpy(Fpd, zero.data(), sizeof(double) * zero.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(rhod, hrho.data(), sizeof(double) * nSlots, hipMemcpyHostToDevice));
        if (massd) BBPD_HIP(hipMemcpy(massd, hm.data(), sizeof(double) * nSlots, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(consd, hc.data(), nSlots, hipMemcpyHostToDevice));
    }
    BBPD_HIP(hipMemcpy(rowlend, rowlen.data(), sizeof(int) * nSlots, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy(soffd, slice_off.data(), sizeof(long long) * slice_off.size(), hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy(nbrd, h_nbr.data(), sizeof(uint32_t) * nE, hipMemcpyHostToDevice));
    if (mixed) {
        BBPD_HIP(hipMemcpy(histd, h_hist32.data(), 4 * nE, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(coefd, h_coef32.data(), 4 * nE, hipMemcpyHostToDevice));
        if (s0d) BBPD_HIP(hipMemcpy(s0d, h_s032.data(), 4 * nE, hipMemcpyHostToDevice));
        if (scd) BBPD_HIP(hipMemcpy(scd, h_sc32.data(), 4 * nE, hipMemcpyHostToDevice));
    } else {
        BBPD_HIP(hipMemcpy(histd, h_hist64.data(), 8 * nE, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(coefd, h_coef64.data(), 8 * nE, hipMemcpyHostToDevice));
        if (s0d) BBPD_HIP(hipMemcpy(s0d, h_s064.data(), 8 * nE, hipMemcpyHostToDevice));
        if (scd) BBPD_HIP(hipMemcpy(scd, h_sc64.data(), 8 * nE, hipMemcpyHostToDevice));
    }
    DevScalars zs{};
    BBPD_HIP(hipMemcpy(s->ds, &zs, sizeof(zs), hipMemcpyHostToDevice));
    BBPD_HIP(hipMemset(s->hist, 0, sizeof(double) * (size_t)HIST_CAP * N_HIST));

    auto comp = [&](double *base, int d) { return base ? base + (size_t)nSlots * d : nullptr; };
    P.X0 = comp(Xd, 0);
    P.X1 = comp(Xd, 1);
    P.X2 = comp(Xd, 2);
    for (int q = 0; q < 2; q++) {
        P.p0[q] = comp(pos[q], 0);
        P.p1[q] = comp(pos[q], 1);
        P.p2[q] = comp(pos[q], 2);
    }
    P.u0 = comp(ud, 0);
    P.u1 = comp(ud, 1);
    P.u2 = comp(ud, 2);
    P.v0 = comp(vd, 0);
    P.v1 = comp(vd, 1);
    P.v2 = comp(vd, 2);
    P.F0 = comp(Fd, 0);
    P.F1 = comp(Fd, 1);
    P.F2 = comp(Fd, 2);
    P.Fp0 = comp(Fpd, 0);
    P.Fp1 = comp(Fpd, 1);
    P.Fp2 = comp(Fpd, 2);
    P.rho = rhod;
    P.mass = massd;
    P.cons = consd;
    P.rowlen = rowlend;
    P.slice_off = soffd;
    P.nbr = nbrd;
    P.hist = histd;
    P.coef = coefd;
    P.s0e = s0d;
    P.sce = scd;
    P.V = V;
    P.horizon = horizon;
    P.hmr = hmr;
    P.hpr = hpr;
    P.two_r = two_r;
    P.s0 = pr->s0 ? 0.0 : pr->s0_scalar;
    P.sc = pr->sc ? 0.0 : pr->sc_scalar;
    P.s1 = pr->s1;
    P.k = pr->k;
    P.alpha = pr->alpha;
    P.denom_exp = 1 - std::exp(-pr->k);
    P.sc_steel = pr->sc_steel;
    P.ff = pr->failureFunctionality;
    P.DT = pr->DT;
    P.D = pr->DAMPING;
    P.partials = partd;
    P.ds = s->ds;

    // ---- contact bodies and observables
    StepEnd &S = s->S;
    S.nBodies = pr->nBodies;
    // the bodies are projected concurrently (one block each), so a node may belong to one family
    // only; the reference applies them in sequence, which is the same for disjoint families
    std::vector<unsigned char> in_family(nN, 0);
    for (int bdy = 0; bdy < pr->nBodies; bdy++) {
        int nf = pr->nFamily[bdy];
        BBPD_REQUIRE(nf >= 0 && (nf == 0 || pr->family[bdy]), "contact family missing");
        std::vector<int> fs(nf > 0 ? nf : 1, 0);
        for (int t = 0; t < nf; t++) {
            double idd = pr->family[bdy][t];
            int id = (int)idd;
            if (idd != (double)id || id < 1 || id > nN) {
                bbpd::set_error("contact family holds an id outside 1..nNodes");
                return fail(BBPD_ERR_ARG);
            }
            if (in_family[id - 1]) {
                bbpd::set_error("contact families overlap (node " + std::to_string(id) +
                                "): the fused solver needs disjoint families");
                return fail(BBPD_ERR_UNSUPPORTED);
            }
            in_family[id - 1] = 1;
            fs[t] = s->slot_of[id - 1];
        }
        int *fd;
        if ((rc = dalloc(&fd, fs.size(), s->owned))) return fail(rc);
        BBPD_HIP(hipMemcpy(fd, fs.data(), sizeof(int) * fs.size(), hipMemcpyHostToDevice));
        S.fam[bdy] = fd;
        S.nFam[bdy] = nf;
        S.cx[bdy] = pr->centreX[bdy];
        S.cz[bdy] = pr->centreZ[bdy];
        S.R[bdy] = pr->radius[bdy];
        S.moves[bdy] = pr->moves[bdy];
    }
    S.schedule = pr->schedule;
    S.nTimeSteps = pr->nTimeSteps;
    S.ipi = pr->itersPerIncrement > 0 ? pr->itersPerIncrement : 1;
    S.applied = pr->appliedDisplacement;
    S.increment = pr->increment;
    auto slot_or_none = [&](int id1) { return (id1 >= 1 && id1 <= nN) ? s->slot_of[id1 - 1] : -1; };
    S.ref_slot = slot_or_none(pr->referenceNode);
    S.cmod1 = slot_or_none(pr->cmodNode1);
    S.cmod2 = slot_or_none(pr->cmodNode2);
    S.hist = s->hist;

    s->device_bytes = 0;
    s->device_bytes += sizeof(double) * 3 * (size_t)nSlots * (mixed ? 5 : 6) + sizeof(double) * 2 * nSlots + nSlots;
    s->device_bytes += (4 + 2 * ew + (s->s0arr ? ew : 0) + (s->scarr ? ew : 0)) * (size_t)nE;
    if (Fpd) s->device_bytes += sizeof(double) * 3 * (size_t)nSlots;

    s->ev.resize(2048);
    for (auto &e : s->ev) BBPD_HIP(hipEventCreate(&e));
    BBPD_HIP(hipDeviceSynchronize());
    *out = s;
    return BBPD_OK;
}

extern "C" int bbpd_solver_destroy(bbpd_solver *s)
{
    if (!s) return BBPD_OK;
    (void)hipSetDevice(s->device);
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    if (s->gexec) (void)hipGraphExecDestroy(s->gexec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (auto &e : s->ev) (void)hipEventDestroy(e);
    for (void *p : s->owned) (void)hipFree(p);
    if (s->stream) (void)hipStreamDestroy(s->stream);
    delete s;
    return BBPD_OK;
}

extern "C" int bbpd_solver_step(bbpd_solver *s, int n)
{
    BBPD_REQUIRE(s && n >= 0, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    for (int i = 0; i < n; i++) {
        int rc = launch_step(s, true);
        if (rc) return rc;
    }
    return BBPD_OK;
}

extern "C" int bbpd_solver_run_graph(bbpd_solver *s, int n, int reps)
{
    BBPD_REQUIRE(s && n > 0 && reps >= 0, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    if (!s->gexec || s->graph_steps != n) {
        if (s->gexec) {
            BBPD_HIP(hipGraphExecDestroy(s->gexec));
            s->gexec = nullptr;
        }
        if (s->graph) {
            BBPD_HIP(hipGraphDestroy(s->graph));
            s->graph = nullptr;
        }
        BBPD_HIP(hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal));
        int rc = 0;
        for (int i = 0; i < n && !rc; i++) rc = launch_step(s, false);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(s->stream, &g);
        if (rc) {
            if (g) (void)hipGraphDestroy(g);
            return rc;
        }
        if (e != hipSuccess) return bbpd::hip_fail(e, "hipStreamEndCapture");
        s->graph = g;
        BBPD_HIP(hipGraphInstantiate(&s->gexec, g, nullptr, nullptr, 0));
        s->graph_steps = n;
    }
    for (int r = 0; r < reps; r++) BBPD_HIP(hipGraphLaunch(s->gexec, s->stream));
    return BBPD_OK;
}

extern "C" int bbpd_solver_sync(bbpd_solver *s)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    return BBPD_OK;
}

extern "C" void *bbpd_solver_stream(bbpd_solver *s) { return s ? (void *)s->stream : nullptr; }

extern "C" int bbpd_solver_observables(bbpd_solver *s, bbpd_observables *o)
{
    BBPD_REQUIRE(s && o, "null argument");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipMemsetAsync(s->counts, 0, sizeof(unsigned long long) * 4, s->stream));
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    unsigned g = (unsigned)(s->nSlots / BLOCK);
    if (mixed) hipLaunchKernelGGL(k_count<true>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, s->counts);
    else hipLaunchKernelGGL(k_count<false>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, s->counts);
    BBPD_HIP(hipGetLastError());
    unsigned long long c[4];
    DevScalars d;
    BBPD_HIP(hipMemcpyAsync(c, s->counts, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    BBPD_HIP(hipMemcpyAsync(&d, s->ds, sizeof(d), hipMemcpyDeviceToHost, s->stream));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    std::memset(o, 0, sizeof(*o));
    o->step = d.step;
    o->reaction = s->S.nBodies > 0 ? d.reaction[0] : 0.0;
    o->reaction_static = d.reaction_static;
    o->cn = d.cn;
    o->n_softening = (int64_t)(c[0] / 2);
    o->n_broken = (int64_t)(c[1] / 2);
    o->n_failed = (int64_t)(c[2] / 2);
    if (d.step > 0) {
        double h[N_HIST];
        BBPD_HIP(hipMemcpy(h, s->hist + (size_t)((d.step - 1) % HIST_CAP) * N_HIST, sizeof(h), hipMemcpyDeviceToHost));
        o->deflection = h[2];
        o->cmod = h[3];
    }
    return BBPD_OK;
}

extern "C" int bbpd_solver_get_nodal(bbpd_solver *s, double *u, double *v, double *F, double *x)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    DevScalars d;
    BBPD_HIP(hipMemcpy(&d, s->ds, sizeof(d), hipMemcpyDeviceToHost));
    const int cur = (int)(d.step & 1);
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int nN = s->nN, nS = s->nSlots;
    std::vector<double> buf((size_t)nS * 3), Xb;
    auto fetch = [&](double *const *comps, double *dst) -> int {
        for (int dd = 0; dd < 3; dd++) BBPD_HIP(hipMemcpy(buf.data() + (size_t)nS * dd, comps[dd], sizeof(double) * nS, hipMemcpyDeviceToHost));
        for (int n = 0; n < nN; n++)
            for (int dd = 0; dd < 3; dd++) dst[n + (size_t)nN * dd] = buf[s->slot_of[n] + (size_t)nS * dd];
        return BBPD_OK;
    };
    double *pc[3] = {s->P.p0[cur], s->P.p1[cur], s->P.p2[cur]};
    double *uc[3] = {s->P.u0, s->P.u1, s->P.u2};
    double *vc[3] = {s->P.v0, s->P.v1, s->P.v2};
    double *fc[3] = {s->P.F0, s->P.F1, s->P.F2};
    int rc;
    if (u && (rc = fetch(mixed ? pc : uc, u))) return rc;
    if (v && (rc = fetch(vc, v))) return rc;
    if (F && (rc = fetch(fc, F))) return rc;
    if (x) {
        if (!mixed) {
            if ((rc = fetch(pc, x))) return rc;
        } else {
            std::vector<double> uu((size_t)nN * 3), XX((size_t)nN * 3);
            double *Xc[3] = {(double *)s->P.X0, (double *)s->P.X1, (double *)s->P.X2};
            if ((rc = fetch(pc, uu.data()))) return rc;
            if ((rc = fetch(Xc, XX.data()))) return rc;
            for (size_t q = 0; q < uu.size(); q++) x[q] = XX[q] + uu[q];
        }
    }
    return BBPD_OK;
}

extern "C" int bbpd_solver_get_bonds(bbpd_solver *s, double *H, double *flag, double *failout)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const long long nE = s->nPadded;
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    std::vector<uint32_t> w(nE);
    BBPD_HIP(hipMemcpy(w.data(), s->P.nbr, sizeof(uint32_t) * nE, hipMemcpyDeviceToHost));
    std::vector<double> h64;
    std::vector<float> h32;
    if (H) {
        if (mixed) {
            h32.resize(nE);
            BBPD_HIP(hipMemcpy(h32.data(), s->P.hist, 4 * nE, hipMemcpyDeviceToHost));
        } else {
            h64.resize(nE);
            BBPD_HIP(hipMemcpy(h64.data(), s->P.hist, 8 * nE, hipMemcpyDeviceToHost));
        }
    }
    for (int k = 0; k < s->nB; k++) {
        long long e = s->entry_i[k], e2 = s->entry_j[k];
        if (H) {
            double a = mixed ? (double)h32[e] : h64[e];
            double b = mixed ? (double)h32[e2] : h64[e2];
            if (!bits_equal(a, b)) {
                bbpd::set_error("internal: the two ends of bond " + std::to_string(k + 1) + " disagree");
                return BBPD_ERR_UNSUPPORTED;
            }
            H[k] = a;
        }
        if (flag) flag[k] = (w[e] & FLAG_BIT) ? 1.0 : 0.0;
        if (failout) failout[k] = (w[e] & FAILED_BIT) ? 0.0 : 1.0;
    }
    return BBPD_OK;
}

namespace {

// nodal array in reference layout (nNodes x 3, column-major) <-> three slot-ordered device arrays
int put_nodal(bbpd_solver *s, double *const *dst, const double *src)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<double> buf((size_t)nS * 3, 0.0);
    for (int dd = 0; dd < 3; dd++) BBPD_HIP(hipMemcpy(buf.data() + (size_t)nS * dd, dst[dd], sizeof(double) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++)
        for (int dd = 0; dd < 3; dd++) buf[s->slot_of[n] + (size_t)nS * dd] = src[n + (size_t)nN * dd];
    for (int dd = 0; dd < 3; dd++) BBPD_HIP(hipMemcpy(dst[dd], buf.data() + (size_t)nS * dd, sizeof(double) * nS, hipMemcpyHostToDevice));
    return BBPD_OK;
}

int get_nodal3(bbpd_solver *s, double *const *srcd, double *dst)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<double> buf((size_t)nS * 3);
    for (int dd = 0; dd < 3; dd++) BBPD_HIP(hipMemcpy(buf.data() + (size_t)nS * dd, srcd[dd], sizeof(double) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++)
        for (int dd = 0; dd < 3; dd++) dst[n + (size_t)nN * dd] = buf[s->slot_of[n] + (size_t)nS * dd];
    return BBPD_OK;
}

} // namespace

extern "C" int bbpd_solver_checkpoint(bbpd_solver *s, int64_t *step, double *u, double *v, double *x, double *F,
                                      double *Fp, double *H, double *flag, double *fail)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    DevScalars d;
    BBPD_HIP(hipMemcpy(&d, s->ds, sizeof(d), hipMemcpyDeviceToHost));
    if (step) *step = d.step;
    int rc = bbpd_solver_get_nodal(s, u, v, F, x);
    if (rc) return rc;
    if (Fp) {
        if (s->P.Fp0) {
            double *fp[3] = {s->P.Fp0, s->P.Fp1, s->P.Fp2};
            if ((rc = get_nodal3(s, fp, Fp))) return rc;
        } else {
            std::memset(Fp, 0, sizeof(double) * 3 * (size_t)s->nN);
        }
    }
    return bbpd_solver_get_bonds(s, H, flag, fail);
}

extern "C" int bbpd_solver_restore(bbpd_solver *s, int64_t step, const double *u, const double *v, const double *x,
                                   const double *F, const double *Fp, const double *H, const double *flag,
                                   const double *fail)
{
    BBPD_REQUIRE(s && step >= 0, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int cur = (int)(step & 1);
    double *pc[3] = {s->P.p0[cur], s->P.p1[cur], s->P.p2[cur]};
    double *uc[3] = {s->P.u0, s->P.u1, s->P.u2};
    double *vc[3] = {s->P.v0, s->P.v1, s->P.v2};
    double *fc[3] = {s->P.F0, s->P.F1, s->P.F2};
    double *fpc[3] = {s->P.Fp0, s->P.Fp1, s->P.Fp2};
    int rc;
    if (u && (rc = put_nodal(s, mixed ? pc : uc, u))) return rc;
    if (!mixed && x && (rc = put_nodal(s, pc, x))) return rc;
    if (v && (rc = put_nodal(s, vc, v))) return rc;
    if (F && (rc = put_nodal(s, fc, F))) return rc;
    if (Fp && s->P.Fp0 && (rc = put_nodal(s, fpc, Fp))) return rc;
    const long long nE = s->nPadded;
    if (H) {
        for (int k = 0; k < s->nB; k++)
            if (!(H[k] >= 0 && H[k] <= 1)) {
                bbpd::set_error("bondSofteningFactor must lie in [0, 1]");
                return BBPD_ERR_ARG;
            }
        if (mixed) {
            std::vector<float> h(nE);
            BBPD_HIP(hipMemcpy(h.data(), s->P.hist, 4 * nE, hipMemcpyDeviceToHost));
            for (int k = 0; k < s->nB; k++) h[s->entry_i[k]] = h[s->entry_j[k]] = (float)H[k];
            BBPD_HIP(hipMemcpy(s->P.hist, h.data(), 4 * nE, hipMemcpyHostToDevice));
        } else {
            std::vector<double> h(nE);
            BBPD_HIP(hipMemcpy(h.data(), s->P.hist, 8 * nE, hipMemcpyDeviceToHost));
            for (int k = 0; k < s->nB; k++) h[s->entry_i[k]] = h[s->entry_j[k]] = H[k];
            BBPD_HIP(hipMemcpy(s->P.hist, h.data(), 8 * nE, hipMemcpyHostToDevice));
        }
    }
    if (flag || fail) {
        std::vector<uint32_t> w(nE);
        BBPD_HIP(hipMemcpy(w.data(), s->P.nbr, sizeof(uint32_t) * nE, hipMemcpyDeviceToHost));
        for (int k = 0; k < s->nB; k++) {
            for (long long e : {s->entry_i[k], s->entry_j[k]}) {
                if (flag) w[e] = (w[e] & ~FLAG_BIT) | (flag[k] == 1 ? FLAG_BIT : 0u);
                if (fail) w[e] = (w[e] & ~FAILED_BIT) | (fail[k] == 0 ? FAILED_BIT : 0u);
            }
        }
        BBPD_HIP(hipMemcpy(s->P.nbr, w.data(), sizeof(uint32_t) * nE, hipMemcpyHostToDevice));
    }
    DevScalars d;
    BBPD_HIP(hipMemcpy(&d, s->ds, sizeof(d), hipMemcpyDeviceToHost));
    d.step = step;
    BBPD_HIP(hipMemcpy(s->ds, &d, sizeof(d), hipMemcpyHostToDevice));
    s->hist_read = step;
    return BBPD_OK;
}

extern "C" int bbpd_solver_history(bbpd_solver *s, double *reaction, double *deflection, int cap, int *n_out)
{
    BBPD_REQUIRE(s && n_out && cap >= 0, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    DevScalars d;
    BBPD_HIP(hipMemcpy(&d, s->ds, sizeof(d), hipMemcpyDeviceToHost));
    long long from = std::max(s->hist_read, d.step - (long long)HIST_CAP);
    long long cnt = std::min((long long)cap, d.step - from);
    std::vector<double> h((size_t)N_HIST);
    for (long long q = 0; q < cnt; q++) {
        long long st = from + q;
        BBPD_HIP(hipMemcpy(h.data(), s->hist + (size_t)(st % HIST_CAP) * N_HIST, sizeof(double) * N_HIST,
                           hipMemcpyDeviceToHost));
        if (reaction) reaction[q] = h[0];
        if (deflection) deflection[q] = h[2];
    }
    s->hist_read = from + cnt;
    *n_out = (int)cnt;
    return BBPD_OK;
}

extern "C" int bbpd_solver_info(bbpd_solver *s, int64_t *device_bytes, int64_t *bond_bytes, int64_t *n_entries,
                                int64_t *n_padded)
{
    BBPD_REQUIRE(s, "null solver");
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int64_t w = mixed ? 4 : 8;
    // algorithmic bytes of one bond-kernel launch: per entry the index word, the history, the
    // coefficient and the per-bond law parameters; per node the state the kernel reads and writes
    int64_t per_entry = 4 + w + w + (s->s0arr ? w : 0) + (s->scarr ? w : 0);
    int64_t per_node;
    if (s->integrator == BBPD_INTEGRATOR_ADR)
        per_node = 24 /*X*/ + 24 /*pos*/ + 24 /*F w*/ + 24 /*Fp*/ + 24 /*v*/ + 24 /*u*/ + 8 /*m*/ + 4 /*rowlen*/;
    else
        per_node = 24 /*X*/ + 24 /*pos r*/ + 24 /*pos w*/ + 48 /*v rw*/ + (mixed ? 0 : 48) /*u rw*/ + 24 /*F w*/ +
                   8 /*rho*/ + 1 /*cons*/ + 4 /*rowlen*/;
    if (device_bytes) *device_bytes = (int64_t)s->device_bytes;
    if (bond_bytes) *bond_bytes = per_entry * s->nEntries + per_node * s->nN;
    if (n_entries) *n_entries = s->nEntries;
    if (n_padded) *n_padded = s->nPadded;
    return BBPD_OK;
}

extern "C" int bbpd_solver_set_timing(bbpd_solver *s, int on)
{
    BBPD_REQUIRE(s, "null solver");
    s->timing = on != 0;
    s->ev_used = 0;
    s->t_acc_ms = 0;
    s->t_launches = 0;
    return BBPD_OK;
}

extern "C" int bbpd_solver_kernel_time(bbpd_solver *s, double *avg_ms, int64_t *launches)
{
    BBPD_REQUIRE(s && avg_ms && launches, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    for (int q = 0; q + 1 < s->ev_used; q += 2) {
        float ms = 0;
        BBPD_HIP(hipEventElapsedTime(&ms, s->ev[q], s->ev[q + 1]));
        s->t_acc_ms += ms;
        s->t_launches++;
    }
    s->ev_used = 0;
    *launches = s->t_launches;
    *avg_ms = s->t_launches ? s->t_acc_ms / (double)s->t_launches : 0.0;
    s->t_acc_ms = 0;
    s->t_launches = 0;
    return BBPD_OK;
}
