// bbpd_kernels.hip -- layer 2: the kernels of the fused, device-resident explicit step, and their launches.
//
// One step restates the reference driver solvers/scripts/dynamic/dynamicsolverdisplacementcontrolled.m:71-121:
//   smoothstepdata (:76) -> deformed length + stretch (:79) -> softening law (:82-91) -> bond force (:94)
//   -> nodal force (:97-101) -> time integration (:105) -> contact x3 (:108-112) -> reaction (:115).
//
// Device layout
//   * Nodes live in "slots": tiles of 64 x 2 x 2 lattice nodes (64 along the longest axis) in Morton
//     order, 64-node lines inside a tile, padded to a multiple of 2048.  A wavefront is one line, and
//     the k-th neighbours of its lanes are again 64 consecutive slots.
//   * Node records (32 B, double-buffered): the position (deformed coordinates in the FP64 form,
//     displacement in the mixed form) and the node's packed integer lattice coordinates.  On a
//     lattice (X == h * k exactly, checked at set-up) the undeformed coordinates of a neighbour are
//     rebuilt from the integers bit for bit, so a neighbour costs two 16-byte loads.  Off-lattice
//     inputs read a second 32-byte record with X.  Other nodal arrays are structure-of-arrays FP64.
//   * Every bond is held from both ends in a node-centric SELL-64 table: entry k of the 64 rows of a
//     slice are 64 consecutive 16-byte records, one 16-byte load per lane.  A row lists its bonds in
//     BONDLIST order, which is the order the reference's serial nodal sum
//     (calculatenodalforcesMex.c:27-44) adds them to the node: the gather is deterministic and
//     reproduces it bit for bit with no atomics.
//   * Entry record: u32 index word (neighbour slot | BONDTYPE << 28 | failed << 30 | flag << 31),
//     u32 table indices, and the history bondSofteningFactor (FP64).  Mixed form: index word, an
//     inline per-bond parameter (FP32), history (FP32), c*V*vcf (FP32).  FP64 BONDSTIFFNESS and
//     per-bond s0 / sc with few distinct values live in small tables copied to LDS by each block;
//     otherwise they are streamed from their own arrays.
//   * UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS are recomputed from the coordinates with the
//     reference formulas (buildhorizons.m:93-97, calculatevolumecorrectionfactors.m:16-34) when set-up
//     finds, bit for bit, that the given arrays equal them; otherwise (FP64 form) both are streamed
//     per entry to the off-lattice record kernel's variant for given values.
//   * Stencil layout (lattice boxes; see k_bond_st): neighbours implicit by slot offset, per entry only
//     the history.
//   * A step reads position buffer step&1 and writes the other, so the Euler-Cromer update of a node
//     runs in the kernel that gathers its neighbours.
#include "bbpd_solver_impl.hpp"

namespace bbpd_sv {

// ------------------------------------------------------------------------------------------------
// Softening laws, per entry.  Each mirrors its reference routine; `w` carries the flags.
// ------------------------------------------------------------------------------------------------
// a / b: the IEEE quotient in FP64; in FP32 (the mixed form) the product with the hardware
// reciprocal (within 1 ulp of the FP32 reciprocal), which the mixed form's law takes
template <typename T>
__device__ __forceinline__ T ldiv(T a, T b)
{
    if constexpr (std::is_same<T, float>::value)
        return a * __builtin_amdgcn_rcpf(b);
    else
        return a / b;
}

// e^x in FP64 for the law's arguments (x in [-k, 0], k = 25 in the reference; any x in [-700, 700]):
// x = n ln2 + r with n = rint(x / ln2) and r by a two-part reduction (ln2 = LN2_HI + LN2_LO, LN2_HI with
// 21 significant bits, so n LN2_HI and x - n LN2_HI are exact), |r| <= ln2 / 2; e^r by its Taylor
// polynomial to degree 13 (truncation below 2^-60 of the value) in Horner form; then the scaling by
// 2^n.  Within about one ulp of e^x.  The coefficients come from constant memory, so the compiler
// keeps them in scalar registers rather than in vector registers it would spill under pressure.
__constant__ double EXP_TAYLOR[14] = {
    1.0, 1.0, 0.5, 0.16666666666666666, 0.041666666666666664, 0.008333333333333333, 0.001388888888888889,
    0.0001984126984126984, 2.48015873015873e-05, 2.7557319223985893e-06, 2.755731922398589e-07,
    2.505210838544172e-08, 2.08767569878681e-09, 1.6059043836821613e-10};
__device__ __forceinline__ double exp_law(double x)
{
    constexpr double LOG2E = 1.4426950408889634;    // 1 / ln2
    constexpr double LN2_HI = 0x1.62e42p-1;         // ln2 to 21 significant bits
    constexpr double LN2_LO = 0x1.fdf473de6af28p-22; // ln2 - LN2_HI
    const double n = rint(x * LOG2E);
    const double r = fma(-n, LN2_LO, fma(-n, LN2_HI, x));
    double p = EXP_TAYLOR[13];
#pragma unroll
    for (int i = 12; i >= 0; i--) p = fma(p, r, EXP_TAYLOR[i]);
    return ldexp(p, (int)n);
}

// e^x: in FP64 exp_law; in FP32 (the mixed form) 2^(x log2 e) by the hardware exponential.  For the
// law's arguments (x in [-k, 0]) the FP32 form's absolute error stays at the FP32 level.
template <typename T>
__device__ __forceinline__ T lexp(T x)
{
    if constexpr (std::is_same<T, float>::value)
        return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    else
        return exp_law(x);
}

// The laws.  A bond whose history is already 1 (broken) keeps it: every candidate is clamped to 1
// or does not exceed it, so the bilinear and exponential candidates are only computed below 1.
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
        if ((w & FLAG_BIT) && H < T(1)) {
            T cand = ldiv(s - s0, s) * ldiv(sc, sc - s0);
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
                if (H < T(1)) {
                    T numerator = T(1) - lexp(ldiv(-kk * (s - s0), sc - s0));
                    T residual = alpha * (T(1) - ldiv(s - s0, sc - s0));
                    bsf = T(1) - ldiv(ldiv(s0, s) * ((T(1) - ldiv(numerator, denom)) + residual), T(1) + alpha);
                }
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
    // consecutive slots stay on one XCD: hardware places block b on XCD b % 8
    int per = nb >> 3;
    return (b & 7) * per + (b >> 3);
}

__device__ inline uint32_t *rec_word(uint4 *rec, long long e, int word)
{
    return reinterpret_cast<uint32_t *>(rec + e) + word;
}

// ADR keeps the forces of the last two steps in the two force sets, by launch parity: the launch with
// parity `cur` writes its F to set `cur` and reads the previous step's from set `cur ^ 1` (so nothing
// copies F to F_prev).  Euler-Cromer uses set 0 only.
__device__ __forceinline__ double *adr_force(const Params &P, int set, int d)
{
    double *const a[3] = {P.F0, P.F1, P.F2};
    double *const b[3] = {P.Fp0, P.Fp1, P.Fp2};
    return set ? b[d] : a[d];
}

// After a node's row: the Euler-Cromer update of the node (timeintegrationeulercromerMex.c:38-46),
// writing its next position record, or, for ADR, F and the block's partial sums: the Rayleigh
// quotient's two (calculatedampingcoefficient.m:61-85) and the squared out-of-balance force over the
// free DOFs (the equilibrium ratio's numerator, after calculateequilibriumstate.m:6-8).  Every
// thread of the block calls it (block reduction; sh holds 3 * BLOCK doubles).
template <bool MIXED>
__device__ __forceinline__ void finish_node(const Params &P, int n, int gb, bool real, int cur, double f0, double f1,
                                            double f2, double qn0, double qn1, double qn2, double Xn0, double Xn1,
                                            double Xn2, double *sh)
{
    if (!P.adr) {
        if (!real) return;
        const int nxt = cur ^ 1;
        P.F0[n] = f0;
        P.F1[n] = f1;
        P.F2[n] = f2;
        const unsigned char cf = P.cons[n];
        const double r = P.rho[n];
        double *uu[3] = {P.u0, P.u1, P.u2};
        double *vv[3] = {P.v0, P.v1, P.v2};
        const double Xn[3] = {Xn0, Xn1, Xn2};
        const double qn[3] = {qn0, qn1, qn2};
        const double ff[3] = {f0, f1, f2};
        double out[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
            double a = (ff[d] - P.D * vv[d][n]) / r;
            if (cf & (1u << d)) a = 0;
            const double nv = vv[d][n] + (a * P.DT);
            vv[d][n] = nv;
            if constexpr (MIXED) {
                out[d] = qn[d] + (nv * P.DT);
            } else {
                const double nu = uu[d][n] + (nv * P.DT);
                uu[d][n] = nu;
                out[d] = Xn[d] + nu;
            }
        }
        NodeRec *po = P.pos[nxt] + n;
        po->x = out[0];
        po->y = out[1];
        po->z = out[2];
    } else {
        double c1 = 0, c2 = 0, c3 = 0;
        if (real) {
            adr_force(P, cur, 0)[n] = f0;
            adr_force(P, cur, 1)[n] = f1;
            adr_force(P, cur, 2)[n] = f2;
        }
        // the quotient runs over the nodes no contact body drives: a family member's velocity
        // after a projection is the projection's correction, and its term would outweigh the
        // structure's by orders of magnitude, pinning cn to 0 or the clamp
        if (real && !(P.cons[n] & CONTACT_BIT)) {
            const double mi = P.mass[n];
            const double ff[3] = {f0, f1, f2};
            const double fp[3] = {adr_force(P, cur ^ 1, 0)[n], adr_force(P, cur ^ 1, 1)[n],
                                  adr_force(P, cur ^ 1, 2)[n]};
            const double vp[3] = {P.v0[n], P.v1[n], P.v2[n]};
            double uq[3];
            if constexpr (MIXED) {
                uq[0] = qn0, uq[1] = qn1, uq[2] = qn2;
            } else {
                uq[0] = P.u0[n], uq[1] = P.u1[n], uq[2] = P.u2[n];
            }
            const unsigned char cf = P.cons[n];
#pragma unroll
            for (int d = 0; d < 3; d++) {
                if (vp[d] != 0) c1 = c1 - (uq[d] * uq[d]) * ((ff[d] / mi - fp[d] / mi) / (P.DT * vp[d]));
                c2 = c2 + uq[d] * uq[d];
                if (!(cf & (1u << d))) c3 = c3 + ff[d] * ff[d];
            }
        }
        bbpd::block_sum_fixed3<BLOCK>(c1, c2, c3, sh);
        if (threadIdx.x == 0) {
            double2 *pp = reinterpret_cast<double2 *>(P.partials) + 2 * gb;
            pp[0] = make_double2(c1, c2);
            pp[1] = make_double2(c3, 0.0);
        }
    }
}

// the global block a launch block works on (after the XCD remap), -1 = idle
__device__ inline int launch_block(const Params &P, bool remap = true)
{
    return P.grid[remap ? remap_block(blockIdx.x, gridDim.x) : (int)blockIdx.x];
}

// ------------------------------------------------------------------------------------------------
// Bond kernel.  Thread = node slot, wavefront = slice.  For each entry of its row the thread loads
// the entry record and the neighbour's node record, updates the entry's law state, and adds the
// row-oriented bond force; entries are summed in BONDLIST order.  After the row it either applies the
// Euler-Cromer update (timeintegrationeulercromerMex.c:38-46) or, for ADR, stores F and the block's
// Rayleigh-quotient partial sums (calculatedampingcoefficient.m:61-85).
//
// Plain FP64 form: the reference arithmetic in the reference order (bit-identical results).
// FP32-mixed form: eta = u_m - u_n and xi = X_m - X_n are formed in FP64 and rounded, the stretch
// uses the cancellation-free s = (2 xi.eta + eta.eta) / (|xi| (|xi + eta| + |xi|)), the law and the
// force run in FP32 (contracted multiply-adds, hardware reciprocal and square root), the row sum and
// all nodal state stay FP64.
//
// Memory: a wide load costs the address unit about as much as a narrow one, so each entry is one
// 16-byte record and each neighbour one 32-byte record.  Two operand sets A and B alternate: while
// entry k (A) is computed, entry k+1 (B) and the index words of k+2 and k+3 are in flight, and the
// reverse for k+1.  The loop runs to the slice width (uniform over the wavefront); padding entries
// are valid memory, so every load is unconditional and only the arithmetic past the row is masked.
// ------------------------------------------------------------------------------------------------
// Mixed form: the factor p of the bond force p * (xi + eta).  The reference zeroes NaN force
// components, which a zero deformed length produces (calculatebondforcesMex.c:38); with finite
// positions the factor is non-finite only then (or when it overflows), so one test of the factor
// stands for the three tests of the components.
__device__ __forceinline__ float mixed_force_factor(float p)
{
    return fabsf(p) <= 3.402823466e38f ? p : 0.0f;
}

template <typename T>
struct Ent {
    uint4 r;             // entry record
    T p0, p1;            // streamed per-bond s0, sc
    double c;            // streamed BONDSTIFFNESS (FP64 form)
    double L0, vcf;      // streamed UNDEFORMEDLENGTH, VOLUMECORRECTIONFACTORS (FP64 form, Params::L0str)
    NodeRec nb;          // neighbour record
    double b0, b1, b2;   // neighbour undeformed coordinates (off-lattice inputs)
};

template <bool MIXED, int LAW, bool LAT, bool GIVEN = false>
__global__ __launch_bounds__(BLOCK) void k_bond(Params P)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    __shared__ double sh[3 * BLOCK];
    extern __shared__ double tabs[]; // FP64 parameter tables
    const int gb = launch_block(P);
    if (gb < 0) return;
    const int n = gb * BLOCK + threadIdx.x;
    const Region &RG = P.reg[P.blkreg[gb]];
    const bool real = n - RG.base < RG.nN;
    const double rs0 = RG.s0, rs1 = RG.s1, rsc = RG.sc;
    const int cur = (int)(*P.tick & 1);
    const NodeRec *__restrict__ pos = P.pos[cur];
    uint4 *__restrict__ rec = P.rec;
    if (!MIXED) {
        const int nt = P.ntabc + P.ntabp;
        for (int i = threadIdx.x; i < nt; i += BLOCK) tabs[i] = P.tab[i];
        __syncthreads();
    }
    const double *tabc = tabs;
    const double *tabp = tabs + P.ntabc;

    double f0 = 0, f1 = 0, f2 = 0;
    const int len = P.rowlen[n];
    const NodeRec me = pos[n];
    const double qn0 = me.x, qn1 = me.y, qn2 = me.z;
    const double Xn0 = P.X0[n], Xn1 = P.X1[n], Xn2 = P.X2[n];
    const long long sbase = P.slice_off[n >> 6];
    const int W = (int)((P.slice_off[(n >> 6) + 1] - sbase) >> 6);
    const long long e0 = sbase + (n & (SLICE - 1));
    auto E = [&](int k) { return e0 + (long long)min(k, W - 1) * SLICE; };
    // the look-ahead index word shares a cache line with its record, which is read two entries
    // later: a normally cached load keeps that line for the record load (non-temporal here would
    // fetch the line from memory twice)
    auto index = [&](int k) { return *rec_word(rec, E(k), 0); };
    auto load = [&](Ent<T> &d, int k, uint32_t w) {
        const long long e = E(k);
        const u32x4 v = __builtin_nontemporal_load(reinterpret_cast<const u32x4 *>(rec + e));
        d.r = make_uint4(v.x, v.y, v.z, v.w);
        if (P.s0mode == PM_STREAM) d.p0 = __builtin_nontemporal_load((const T *)P.s0str + e);
        if (P.scmode == PM_STREAM) d.p1 = __builtin_nontemporal_load((const T *)P.scstr + e);
        if (!MIXED && P.cmode == PM_STREAM) d.c = __builtin_nontemporal_load(P.cstr + e);
        if (GIVEN) {
            d.L0 = __builtin_nontemporal_load(P.L0str + e);
            d.vcf = __builtin_nontemporal_load(P.vcfstr + e);
        }
        const int m = (int)(w & NODE_MASK);
        d.nb = pos[m];
        if (!LAT) {
            const NodeRec xr = P.Xrec[m];
            d.b0 = xr.x, d.b1 = xr.y, d.b2 = xr.z;
        }
    };
    auto compute = [&](const Ent<T> &d, long long e) {
        uint32_t w = d.r.x;
        double Xm0, Xm1, Xm2;
        if (LAT) {
            const uint64_t q = d.nb.q;
            Xm0 = P.hx * (double)(int)(q & LAT_MASK);
            Xm1 = P.hy * (double)(int)((q >> LAT_BITS) & LAT_MASK);
            Xm2 = P.hz * (double)(int)((q >> (2 * LAT_BITS)) & LAT_MASK);
        } else {
            Xm0 = d.b0, Xm1 = d.b1, Xm2 = d.b2;
        }
        if constexpr (!MIXED) {
            double H = __hiloint2double((int)d.r.w, (int)d.r.z);
            const double H0 = H;
            const double c = P.cmode == PM_TABLE ? tabc[d.r.y & 0xFFFFu] : d.c;
            const double s0 = P.s0mode == PM_TABLE ? tabp[d.r.y >> 16] : (P.s0mode == PM_STREAM ? d.p0 : rs0);
            const double sc = P.scmode == PM_TABLE ? tabp[d.r.y >> 16] : (P.scmode == PM_STREAM ? d.p1 : rsc);
            // calculatedeformedlengthMex.c:37-43, row-oriented (x_m - x_n): exact negation of the
            // reference's x_j - x_i at the j end, so |.|, s and the force magnitude are identical
            const double dx = d.nb.x - qn0, dy = d.nb.y - qn1, dz = d.nb.z - qn2;
            const double l = sqrt(dx * dx + dy * dy + dz * dz);
            // the undeformed length and volume correction factor: as given where the problem's arrays
            // differ from the formulas, else recomputed (buildhorizons.m:93-97,
            // calculatevolumecorrectionfactors.m:20-32)
            const double ex = Xn0 - Xm0, ey = Xn1 - Xm1, ez = Xn2 - Xm2;
            const double L0 = GIVEN ? d.L0 : sqrt((ex * ex + ey * ey) + ez * ez);
            const double s = (l - L0) / L0;
            apply_law<LAW, double>(s, s0, sc, rs1, P.k, P.alpha, P.denom_exp, P.sc_steel, P.ff, H, w);
            double vcf;
            if (GIVEN) vcf = d.vcf;
            else if (L0 <= P.hmr) vcf = 1;
            else if ((L0 > P.hmr) && (L0 <= P.horizon)) vcf = (P.hpr - L0) / P.two_r;
            else vcf = 0;
            // calculatebondforcesMex.c:34-40 with stretchPlastic = 0 (s - 0 == s exactly)
            const double fail = (w & FAILED_BIT) ? 0.0 : 1.0;
            const double p = fail * c * (1 - H) * s * P.V * vcf;
            // dx/l, dy/l, dz/l: with rl = 1/l correctly rounded, q = dx*rl, r = dx - l*q (exact by
            // FMA) and q' = q + r*rl rounded once is the correctly rounded quotient (Markstein's
            // theorem; no quotient here is subnormal), so this equals the IEEE division bit for bit
            const double rl = 1.0 / l;
            double qx = dx * rl, qy = dy * rl, qz = dz * rl;
            qx = fma(fma(-l, qx, dx), rl, qx);
            qy = fma(fma(-l, qy, dy), rl, qy);
            qz = fma(fma(-l, qz, dz), rl, qz);
            double gx = p * qx, gy = p * qy, gz = p * qz;
            if (gx != gx) gx = 0.0;
            if (gy != gy) gy = 0.0;
            if (gz != gz) gz = 0.0;
            f0 += gx;
            f1 += gy;
            f2 += gz;
            if (__double_as_longlong(H) != __double_as_longlong(H0))
                *reinterpret_cast<double *>(rec_word(rec, e, 2)) = H;
        } else {
            float H = __uint_as_float(d.r.z);
            const float H0 = H;
            const float kap = __uint_as_float(d.r.w);
            const float s0 = P.s0mode == PM_INLINE ? __uint_as_float(d.r.y)
                                                   : (P.s0mode == PM_STREAM ? d.p0 : (float)rs0);
            const float sc = P.scmode == PM_INLINE ? __uint_as_float(d.r.y)
                                                   : (P.scmode == PM_STREAM ? d.p1 : (float)rsc);
            const float xi0 = (float)(Xm0 - Xn0), xi1 = (float)(Xm1 - Xn1), xi2 = (float)(Xm2 - Xn2);
            const float et0 = (float)(d.nb.x - qn0), et1 = (float)(d.nb.y - qn1), et2 = (float)(d.nb.z - qn2);
            const float y0 = xi0 + et0, y1 = xi1 + et1, y2 = xi2 + et2;
            const float L0 = __builtin_amdgcn_sqrtf(fmaf(xi0, xi0, fmaf(xi1, xi1, xi2 * xi2)));
            const float l = __builtin_amdgcn_sqrtf(fmaf(y0, y0, fmaf(y1, y1, y2 * y2)));
            const float num =
                fmaf(2.0f, fmaf(xi0, et0, fmaf(xi1, et1, xi2 * et2)), fmaf(et0, et0, fmaf(et1, et1, et2 * et2)));
            const float rl = __builtin_amdgcn_rcpf(l);
            const float s = num * __builtin_amdgcn_rcpf(L0 * (l + L0));
            apply_law<LAW, float>(s, s0, sc, (float)rs1, (float)P.k, (float)P.alpha, (float)P.denom_exp,
                                  (float)P.sc_steel, (float)P.ff, H, w);
            const float fail = (w & FAILED_BIT) ? 0.0f : 1.0f;
            const float p = mixed_force_factor(fail * kap * (1.0f - H) * s * rl);
            f0 += (double)(p * y0);
            f1 += (double)(p * y1);
            f2 += (double)(p * y2);
            if (__float_as_uint(H) != __float_as_uint(H0)) *rec_word(rec, e, 2) = __float_as_uint(H);
        }
        if (w != d.r.x) *rec_word(rec, e, 0) = w;
    };
    if (W > 0) {
        Ent<T> A, B;
        uint32_t iB = index(1);
        load(A, 0, index(0));
        for (int k = 0; k < W; k += 2) {
            const uint32_t iA2 = index(k + 2);
            const uint32_t iB2 = index(k + 3);
            load(B, k + 1, iB);
            if (k < len) compute(A, E(k));
            load(A, k + 2, iA2);
            if (k + 1 < len) compute(B, E(k + 1));
            iB = iB2;
        }
    }
    finish_node<MIXED>(P, n, gb, real, cur, f0, f1, f2, qn0, qn1, qn2, Xn0, Xn1, Xn2, sh);
}

// ------------------------------------------------------------------------------------------------
// Bond kernel, stencil layout.  On a full lattice box whose rows list their bonds in ascending
// lattice-offset order (the order a BONDLIST sorted by node id gives), the neighbour at offset o of
// slot n is slot n + delta_o: nothing per entry but the history is stored.  Per node, a 128-bit
// mask says which offsets are bonds (boundaries, notch).  BONDSTIFFNESS is a function of the pair
// (nFAMILY_i, nFAMILY_j) (buildbonddata.m:48-51), read from a small table keyed by the pair of the
// two nodes' family-size indices; per-bond s0 / sc come from key tables as well when they are
// functions of that key, else from per-entry streams.
// The softening flag (failed flag for the brittle law) is the sign bit of the stored history.
// Entries are visited in offset order, so the sums per node are those of the record layout and of
// the reference's serial nodal loop, bit for bit (FP64).
// ------------------------------------------------------------------------------------------------
template <typename T>
struct StEnt {
    T h;      // stored history (sign = flag)
    T p0, p1; // streamed s0, sc
    StOff st; // per-offset data (the same for every lane)
    unsigned long long mk; // the wavefront's lanes that have the bond
    bool hb;  // the slice's history bit of the offset (clear: every entry holds +0, not loaded)
};

// one stencil row of neighbour records in flight: sites 0 .. 63 (a) and 64 .. 63 + 2 rx (b)
struct RowRecs {
    NodeRec a, b;
};

// Per-offset tables are the same for every lane of a wavefront and no kernel writes them: read
// through the constant address space, they become scalar loads into scalar registers.  Memory a
// kernel writes (the history words) is not read this way.
template <typename T>
using sconst_ptr = const __attribute__((address_space(4))) T *;
template <typename T>
__device__ __forceinline__ T sload(const T *p, int i)
{
    static_assert(sizeof(T) == 4 || sizeof(T) == 8 || sizeof(T) == 16 || sizeof(T) == 32, "32- to 256-bit elements");
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, ((sconst_ptr<u32x2>)p)[i]);
    else if constexpr (sizeof(T) == 32)
        return __builtin_bit_cast(T, ((sconst_ptr<u32x8>)p)[i]);
    else if constexpr (sizeof(T) == 16)
        return __builtin_bit_cast(T, ((sconst_ptr<u32x4>)p)[i]);
    else
        return __builtin_bit_cast(T, ((sconst_ptr<uint32_t>)p)[i]);
}

// Buffer resources (gfx9 descriptor: 32-bit data format, raw addressing).  The address unit adds a
// scalar base, a per-lane 32-bit offset and a scalar offset, so a stream whose per-lane offset is
// fixed costs no vector instruction per access; a read at a per-lane offset past num_records
// returns zeros without touching memory.
using brsrc = __amdgpu_buffer_rsrc_t;
constexpr int CP_NT = 2; // cache policy: non-temporal

__device__ __forceinline__ brsrc make_rsrc(const void *base, uint32_t bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes, 0x00020000);
}

template <typename T>
__device__ __forceinline__ T bload(brsrc r, uint32_t voff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, 0, 0));
}

template <typename T>
__device__ __forceinline__ T bload_nt(brsrc r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, voff, soff, CP_NT));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, CP_NT));
}

template <typename T>
__device__ __forceinline__ void bstore(T v, brsrc r, uint32_t voff, uint32_t soff)
{
    if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, voff, soff, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, voff, soff, 0);
}

// LDS byte address of a __shared__ object, and a pointer from such an address: an address that is
// a whole sum needs no add in the instruction that reads through it
__device__ __forceinline__ uint32_t lds_addr(const void *p)
{
    return (uint32_t)(uintptr_t)(const __attribute__((address_space(3))) void *)p;
}
template <typename T>
__device__ __forceinline__ const T *lds_ptr(uint32_t a)
{
    return (const T *)(const __attribute__((address_space(3))) T *)(uintptr_t)a;
}

struct RecWords {
    u32x4 a, b;
};

__device__ __forceinline__ NodeRec bload_rec(brsrc r, uint32_t voff)
{
    RecWords w;
    w.a = __builtin_amdgcn_raw_buffer_load_b128(r, voff, 0, 0);
    w.b = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 16u, 0, 0);
    return __builtin_bit_cast(NodeRec, w);
}

// FP64 sqrt(x) and 1 / x for x in [2^-767, 2^1000]: the same operations as the compiler's IEEE
// sequences for these (reciprocal square root with Goldschmidt iterations; reciprocal with two Newton
// steps and the FMA correction of the quotient), without the operand scaling and special-case
// handling those carry for exponents near the limits and for 0, infinity and NaN, which take no effect
// in this range: the same bits.
constexpr double CORE_LO = 0x1p-767, CORE_HI = 0x1p+1000;
__device__ __forceinline__ double sqrt_core(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = r * 0.5;
    const double e = fma(-h, g, 0.5);
    g = fma(g, e, g);
    double d = fma(-g, g, x);
    h = fma(h, e, h);
    g = fma(d, h, g);
    d = fma(-g, g, x);
    return fma(d, h, g);
}
__device__ __forceinline__ double rcp_core(double x)
{
    double r = __builtin_amdgcn_rcp(x);
    double e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    r = fma(r, e, r);
    e = fma(-x, r, 1.0);
    return fma(e, r, r);
}

// The law's lowest threshold.  Every value a law stores lies in [0, 1]
// (calculatebsfexponentialMex.c:55-58 clamps; the others are maxima of values in [0, 1]) and set-up
// and restore accept only such values; with BONDTYPE 0 (the stencil layout's only type), a stretch
// at or below this threshold then leaves H and the flag as they are (calculatebondfailureMex.c:36-38,
// calculatebondsofteningfactor.m:42-48, calculatebsftrilinear.m:80-112,
// calculatebsfexponentialMex.c:35-58).
template <int LAW, typename T>
__device__ __forceinline__ T law_threshold(T s0, T sc, T s1)
{
    if (LAW == BBPD_LAW_LINEAR) return sc;
    if (LAW == BBPD_LAW_BILINEAR) return s0;
    if (LAW == BBPD_LAW_TRILINEAR) return fmin(fmin(s0, s1), sc);
    return fmin(s0, sc);
}

// True when the law can change the history H or the flag at stretch s; `thr` = law_threshold(s0,
// sc, s1).  A bilinear bond, once flagged, is evaluated every step: in compression its candidate is
// positive.
template <int LAW, typename T>
__device__ __forceinline__ bool law_active(T s, T thr, bool flag)
{
    if (LAW == BBPD_LAW_BILINEAR) return flag || s > thr;
    return s > thr;
}

// Mixed form, stencil layout: with x = (|y|^2 - |xi|^2) / (2 |xi|^2), the stretch is
// s = sqrt(1 + 2x) - 1 and the force factor g = s / (1 + s) = 1 - (1 + 2x)^(-1/2).  For |x| at most
// this both come from their series in x, else from a reciprocal square root and a reciprocal.  Both
// round to within a few ulps of the exact values.
constexpr float MIX_SERIES_X = 0x1p-10f;

// Mixed form: a law threshold on the stretch mapped to x = s + s^2 / 2 (increasing in s), lowered by a
// relative 2^-20 so that the test on x runs the law for every entry whose stretch passes the threshold
// (and for a few at it, where the law changes nothing)
__device__ __forceinline__ float thr_to_x(float t)
{
    const float x = fmaf(0.5f * t, t, t);
    return x - fabsf(x) * 0x1p-20f;
}

// Launch variants, picked per solver by an autotuner: OCC is the minimum waves per SIMD the compiler
// must allow (1 = its own choice), which bounds the registers per wave, trading instruction-level
// for wave-level parallelism; REMAP keeps consecutive blocks on one XCD (its L2) or not.  PV says
// where s0 and sc come from: PV_SCALAR the region's scalars, PV_TABLE both from key tables,
// PV_GENERAL each from its scalar or per-entry stream, PV_GENTAB the same with a key table for one of
// them, as the solver's modes say.
template <bool MIXED, int LAW, int OCC, bool REMAP, int PV, bool TILE>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(TILE && OCC > 1 ? 8 : OCC))) void k_bond_st(Params P)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    constexpr uint32_t TW = sizeof(T);
    // per wavefront, a row buffer: x, y, z and the key word (low word) of the staged sites, RS sites
    // per component: one row of ST_ROW sites (slices), or TILE_L lines of TILE_W sites (tiles).  One
    // buffer suffices: a wavefront's LDS accesses complete in program order, so staging the next row
    // (after the last read of this one) cannot overtake a read.
    constexpr int RS = TILE ? TILE_W : ST_ROW;
    __shared__ double s_row[BLOCK / SLICE][TILE ? TILE_L * 4 * TILE_W : 4 * ST_ROW];
    extern __shared__ double tabs[]; // key tables (FP64, or FP32 in the mixed form)
    const int gb = launch_block(P, REMAP);
    if (gb < 0) return;
    const Region &RG = P.reg[P.blkreg[gb]];
    const int lane = (int)(threadIdx.x % SLICE);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / SLICE));
    // tiles: wavefront wv of the region's block takes tile (block * 4 + wv); a wavefront past the last
    // tile has no node (it still joins the block's reductions)
    const int tile = TILE ? (gb - RG.blk0) * (BLOCK / SLICE) + wv : 0;
    const bool wave_on = !TILE || tile < RG.ntile;
    const int t0 = TILE && wave_on ? RG.base + tile_slot(tile, 0, RG.tnx) : 0; // the tile's first slot
    const int n = TILE ? (wave_on ? t0 + lane % TILE_X + RG.tnx * (lane / TILE_X) : RG.base)
                       : gb * BLOCK + (int)threadIdx.x;
    const int base = RG.base, noff = RG.noff, off0 = RG.off0, kg = RG.kg, tab0 = RG.tab0, ntab = RG.ntab;
    const int ts0 = RG.ts0, tsc = RG.tsc;
    // a region's flag, made a scalar so that a test of it is a scalar branch
    const bool fast64 = !MIXED && __builtin_amdgcn_readfirstlane(RG.fast64) != 0;
    const bool real = TILE ? wave_on : n - base < RG.nN;
    // the law's threshold: on the stretch (FP64 form), on x (mixed form, thr_to_x)
    auto thr_of = [](T t) -> T {
        if constexpr (MIXED) return thr_to_x(t);
        else return t;
    };
    const T rs0 = (T)RG.s0, rs1 = (T)RG.s1, rsc = (T)RG.sc, rthr = thr_of(law_threshold<LAW, T>(rs0, rsc, rs1));
    const T rs0lo = (T)RG.s0lo, rsclo = (T)RG.sclo;
    // PV_GENERAL: region scalars, a streamed value entering as the region's smallest (see below)
    constexpr bool GEN = PV == PV_GENERAL || PV == PV_GENTAB;
    const T gs0 = P.s0mode == PM_STREAM ? rs0lo : rs0, gsc = P.scmode == PM_STREAM ? rsclo : rsc;
    const T gthr = thr_of(law_threshold<LAW, T>(gs0, gsc, rs1));
    const T lk = (T)P.k, lalpha = (T)P.alpha, ldenom = (T)P.denom_exp, lsteel = (T)P.sc_steel, lff = (T)P.ff;
    const int cur = (int)(*P.tick & 1);
    const NodeRec *__restrict__ pos = P.pos[cur];
    // key tables in LDS.  PV_TABLE keeps one (c, law threshold) record per key, and reads s0 and sc
    // from the global tables when the law runs; otherwise the tables follow one another (c, then s0
    // and sc where tabled).
    const T *tsrc = MIXED ? reinterpret_cast<const T *>(P.tabf) : reinterpret_cast<const T *>(P.tab);
    T *tdst = reinterpret_cast<T *>(tabs);
    if (PV == PV_TABLE) {
        for (int i = threadIdx.x; i < kg * kg; i += BLOCK) {
            tdst[2 * i] = tsrc[tab0 + i];
            tdst[2 * i + 1] = thr_of(law_threshold<LAW, T>(tsrc[tab0 + ts0 + i], tsrc[tab0 + tsc + i], (T)RG.s1));
        }
    } else {
        for (int i = threadIdx.x; i < ntab; i += BLOCK) tdst[i] = tsrc[tab0 + i];
    }
    __syncthreads();
    const T *tab = tdst;
    // the region's key tables in global memory (PV_TABLE reads s0 and sc there when a law runs)
    const brsrc rtab = make_rsrc(tsrc + tab0, (uint32_t)ntab * TW);
    // a neighbour's staged key word: its family-size index, or (PV_SCALAR / PV_TABLE) that index's
    // byte offset in a row of the key table; the node's row `krow` is then an LDS byte address, so one
    // add gives the address of the entry's table record
    constexpr int KSCALE = GEN ? 1 : (PV == PV_TABLE ? 2 : 1) * (int)sizeof(T);
    const int tabl = GEN ? 0 : (int)lds_addr(tabs);

    const NodeRec me = pos[n];
    const double qn0 = me.x, qn1 = me.y, qn2 = me.z;
    const int kx = (int)(me.q & LAT_MASK), ky = (int)((me.q >> LAT_BITS) & LAT_MASK),
              kz = (int)((me.q >> (2 * LAT_BITS)) & LAT_MASK);
    const int gme = (int)(me.q >> G_SHIFT); // index of the node's family size
    const int krow = gme * kg * KSCALE + tabl; // the node's row of the key table (index or LDS address)
    // X == h * k exactly (checked at set-up)
    const double Xn0 = P.hx * (double)kx, Xn1 = P.hy * (double)ky, Xn2 = P.hz * (double)kz;
    // the entry arrays: entry (slice, o, lane) of the region at ((slice * noff + o) * SLICE + lane) * TW.
    // Slices: the wavefront's slice, lane offset lane * TW.  Tiles: the region's entries, the lane's
    // own slice and lane in the offset (set-up keeps them below 2^31 bytes).
    const long long eslice = TILE ? RG.e0 : RG.e0 + ((long long)((gb * BLOCK - base) / SLICE) + wv) * noff * SLICE;
    const uint32_t sbytes = TILE ? (uint32_t)(((long long)RG.nSlots * noff + 2 * SLICE) * TW) : (uint32_t)noff * SLICE * TW;
    const brsrc rh = make_rsrc((const T *)P.hs + eslice, sbytes);
    const brsrc r0 = GEN && P.s0mode == PM_STREAM ? make_rsrc((const T *)P.s0str + eslice, sbytes) : rh;
    const brsrc r1 = GEN && P.scmode == PM_STREAM ? make_rsrc((const T *)P.scstr + eslice, sbytes) : rh;
    const uint32_t lb = TILE ? (uint32_t)(((n - base) / SLICE * noff * SLICE + (n - base) % SLICE) * (int)TW)
                             : (uint32_t)lane * TW;
    // Neighbour records, staged per stencil row (the offsets that share dy and dz, consecutive in
    // offset order).  For the slice starting at slot n0, row r covers the slots
    // n0 + delta_r - rx .. n0 + delta_r + 63 + rx; the wavefront loads them once into its LDS row
    // buffer, and the offset (dx, dy, dz) of lane i reads site i + rx + dx there.  The next row's
    // loads are in flight while a row is computed.  Slot offsets are 32-bit: a site before slot 0
    // wraps to an offset past every valid one (set-up keeps slots below 2^26), and a read outside
    // the slot array returns zeros (such a site is never a bond).
    // Tiles stage each stencil row as TILE_L lines of TILE_W sites, lines at the slot stride tnx: line j
    // of row r holds the sites t0 + delta_r + j * tnx - TILE_RX + i, and lane (x, y) of the tile reads
    // line y, site x + TILE_RX + dx.  The lines sit in LDS line slots (sh + j) & 3: when row r + 1 is
    // row r moved by one line (delta_{r+1} - delta_r == tnx, the next dy of the same dz), it shares
    // three of its lines, so only its last line is loaded and sh advances by one; otherwise all four
    // lines are loaded.
    const brsrc rp = make_rsrc(pos, (uint32_t)P.nSlots * (uint32_t)sizeof(NodeRec));
    const int n0 = gb * BLOCK + wv * SLICE;
    const int rx = RG.rx, nrow = RG.nrow, tnx = RG.tnx;
    const int *rows = reinterpret_cast<const int *>(P.srow + RG.row0); // (count, delta) per row
    auto fetch = [&](RowRecs &R, int dl) {
        const uint32_t sb = (uint32_t)(n0 + dl - rx);
        R.a = bload_rec(rp, (sb + (uint32_t)lane) * (uint32_t)sizeof(NodeRec));
        if (lane < 2 * rx) R.b = bload_rec(rp, (sb + SLICE + (uint32_t)lane) * (uint32_t)sizeof(NodeRec));
    };
    auto put = [&](double *w, int site, const NodeRec q) {
        w[site] = q.x;
        w[RS + site] = q.y;
        w[2 * RS + site] = q.z;
        *reinterpret_cast<int *>(w + 3 * RS + site) = (int)(q.q >> G_SHIFT) * KSCALE;
        // the record's low lattice word is not used; keeping it live until here stops the register
        // allocator from reusing its register while the load is in flight (which would make the
        // wavefront wait for every outstanding load)
        __asm__ volatile("" ::"v"((uint32_t)q.q));
    };
    auto stage = [&](const RowRecs &R) {
        double *w = &s_row[wv][0];
        put(w, lane, R.a);
        if (lane < 2 * rx) put(w, SLICE + lane, R.b);
    };
    // tiles: the loads of a row (all lines, or only the last), and their LDS stores
    constexpr int TSITES = TILE_L * TILE_W; // sites of the four lines: lanes take site lane and lane + 64
    // the lane index, computed where it is used: the site and line numbers derived from it would
    // otherwise be kept in registers through the entry loop, which costs a wavefront per SIMD
    auto lane_now = []() {
        int l;
        __asm__ volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
        return l;
    };
    auto fetch_t = [&](RowRecs &R, int dl, bool one) {
        const int lane = lane_now();
        const uint32_t sb = (uint32_t)(t0 + dl - TILE_RX);
        // record a: line TILE_L - 1, site lane (only the last line), or site lane of the four lines
        const int ja = one ? TILE_L - 1 : lane / TILE_W, ia = one ? lane : lane % TILE_W;
        if (!one || lane < TILE_W) R.a = bload_rec(rp, (sb + (uint32_t)(ja * tnx + ia)) * (uint32_t)sizeof(NodeRec));
        // record b: site lane + 64 of the four lines
        const int s2 = lane + SLICE;
        if (!one && s2 < TSITES)
            R.b = bload_rec(rp, (sb + (uint32_t)(s2 / TILE_W * tnx + s2 % TILE_W)) * (uint32_t)sizeof(NodeRec));
    };
    int sh = 0; // tiles: LDS line slot of the current row's line 0
    auto stage_t = [&](const RowRecs &R, bool one) {
        const int lane = lane_now();
        double *w = &s_row[wv][0];
        constexpr int LW = 4 * TILE_W; // doubles per line slot
        const int ja = one ? TILE_L - 1 : lane / TILE_W, ia = one ? lane : lane % TILE_W;
        if (!one || lane < TILE_W) put(w + ((sh + ja) & 3) * LW, ia, R.a);
        const int s2 = lane + SLICE;
        if (!one && s2 < TSITES) put(w + ((sh + s2 / TILE_W) & 3) * LW, s2 % TILE_W, R.b);
    };
    // the lane's first site in the row buffer (site x + TILE_RX of its line, for tiles)
    auto row_base = [&]() -> const double * {
        if constexpr (TILE) {
            const int lane = lane_now();
            return &s_row[wv][0] + ((sh + lane / TILE_X) & 3) * (4 * TILE_W) + lane % TILE_X + TILE_RX;
        }
        else
            return &s_row[wv][0] + lane + rx;
    };
    const double *nbase = row_base();
    double f0 = 0, f1 = 0, f2 = 0;
    const StOff *sst = P.sst + RG.sst0;
    // lane masks per (slice, offset), or per (tile, offset) for tiles
    const unsigned long long *mks = TILE ? P.smaskT + RG.e0 / SLICE + (long long)tile * noff : P.smaskT + eslice / SLICE;
    // the slice's history words (noff <= 128): an offset whose bit is clear holds +0 in every entry,
    // which is taken without a load.  This kernel sets bits of them (atomicOr below), so they are read
    // once, here, with an ordinary global load (every lane the same address) made wavefront-uniform
    // with readfirstlane, not through the constant address space; the bits set later are for offsets
    // this wavefront has already passed.
    unsigned long long *hzp = P.snz + 2 * (TILE ? base / SLICE + tile : n0 / SLICE);
    auto uniform64 = [](unsigned long long w) {
        const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)w);
        const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(w >> 32));
        return ((unsigned long long)hi << 32) | lo;
    };
    // the two words as a 128-bit shift register, shifted right by 4 per group of four offsets: bit k
    // is then the bit of the group's offset k (k = 4: the next group's first)
    unsigned long long hlo = uniform64(hzp[0]), hhi = uniform64(hzp[1]);
    // The offsets in groups of four: the group's per-offset records, lane masks and history bytes
    // are at constant offsets from three bases that advance once per group (sp, mp, hso), so a load
    // needs no scalar address arithmetic of its own.
    const StOff *sp = sst;
    const unsigned long long *mp = mks;
    uint32_t hso = 0; // byte offset of the group's first offset in the slice's history
    // the loads of the group's offset k (0 .. 4); the offsets run to noff + 1: the entry arrays end at
    // least two slices' worth past the last entry and the per-offset tables ST_PAD records past the
    // last offset
    auto load = [&](StEnt<T> &d, int k, uint32_t bits) {
        d.hb = ((bits >> k) & 1u) != 0u;
        if (d.hb) d.h = bload_nt<T>(rh, lb + (uint32_t)k * (SLICE * TW), hso);
        else d.h = (T)0;
        d.st = sload(sp, k);
        d.mk = sload(mp, k);
    };
    // One entry.  The FP64 form adds its force to the row sum here; the mixed form adds it to the
    // FP32 sum `a` of a group of up to four entries, which the loop adds to the FP64 row sum.
    // FP64 form: the y and z terms of the undeformed length's square, (Y_n - Y_m)^2 and (Z_n - Z_m)^2,
    // depend on the offset's (dy, dz) only, which a stencil row shares: computed once per row, for
    // every lane (before the lane-mask test).  Not in the variant with streamed parameters, where the
    // two registers they hold across the row cost more in spills than the row saves.
    int cyz = INT_MIN;
    double eyy = 0, ezz = 0;
    auto compute = [&](const StEnt<T> &d, int o, int k, bool bit, float3 &a) {
        if (!MIXED && !GEN && d.st.dyz != cyz) {
            cyz = d.st.dyz;
            const double ey = Xn1 - P.hy * (double)(ky + d.st.dy()), ez = Xn2 - P.hz * (double)(kz + d.st.dz());
            eyy = ey * ey;
            ezz = ez * ez;
        }
        // the record's last word (unused by the mixed form) kept live where the record is used: the
        // record then arrives by one 32-byte scalar load rather than by three narrower ones
        if constexpr (MIXED) __asm__ volatile("" ::"s"(d.st.dyz));
        if (!bit) return;
        // the entry's byte offset in the slice's history (and in the parameter streams)
        const uint32_t so = hso + (uint32_t)k * (SLICE * TW);
        const double *nb = nbase + d.st.dx;
        const double nbx = nb[0], nby = nb[RS], nbz = nb[2 * RS];
        const int key = krow + *reinterpret_cast<const int *>(nb + 3 * RS);
        // the entry's key-table values: c (c * V in the mixed form), s0, sc
        const T *tk = GEN ? tab + key : lds_ptr<T>((uint32_t)key);
        constexpr uint32_t mark = (LAW == BBPD_LAW_LINEAR) ? FAILED_BIT : FLAG_BIT;
        // the stored value: the history is its magnitude (+0 for -0), the flag its sign bit; the law
        // replaces it, and the force reads the magnitude of what is stored after the law
        T hv = d.h;
        bool flag = MIXED ? (int)__float_as_uint((float)d.h) < 0 : __double_as_longlong((double)d.h) < 0;
        T s0 = rs0, sc = rsc, thr = rthr, c;
        if (PV == PV_TABLE) {
            // one 2-element LDS read: the record is 2 * sizeof(T) aligned
            const T2<T> kt = *reinterpret_cast<const T2<T> *>(tk);
            c = kt.x;
            thr = kt.y;
        } else {
            c = tk[0];
        }
        if (PV == PV_GENERAL) {
            // a streamed s0 / sc enters the threshold as the region's smallest value: a stretch at or
            // below it cannot change the law's state, and the entry's own value is read only when the
            // law runs (most entries, most steps, read no parameter stream)
            s0 = gs0;
            sc = gsc;
            thr = gthr;
        } else if (PV == PV_GENTAB) {
            s0 = P.s0mode == PM_TABLE ? tab[ts0 + key] : gs0;
            sc = P.scmode == PM_TABLE ? tab[tsc + key] : gsc;
            thr = thr_of(law_threshold<LAW, T>(s0, sc, rs1));
        }
        // the law at stretch s once its test has passed, and the store of a changed history (sign = flag)
        auto law_run = [&](T s) {
            if (PV == PV_TABLE) {
                const uint32_t ki = (uint32_t)(key - tabl) / KSCALE;
                s0 = bload<T>(rtab, (uint32_t)(ts0 + ki) * TW);
                sc = bload<T>(rtab, (uint32_t)(tsc + ki) * TW);
            }
            if (GEN && P.s0mode == PM_STREAM) s0 = bload<T>(r0, lb + so);
            if (GEN && P.scmode == PM_STREAM) sc = bload<T>(r1, lb + so);
            uint32_t w = flag ? mark : 0u;
            T H = MIXED ? (T)fabsf((float)d.h) : (T)fabs((double)d.h);
            apply_law<LAW, T>(s, s0, sc, rs1, lk, lalpha, ldenom, lsteel, lff, H, w);
            flag = (w & mark) != 0u;
            hv = flag ? -H : H;
            if (__builtin_bit_cast(typename std::conditional<MIXED, uint32_t, unsigned long long>::type, hv) !=
                __builtin_bit_cast(typename std::conditional<MIXED, uint32_t, unsigned long long>::type, d.h)) {
                bstore<T>(hv, rh, lb, so);
                // the first history word other than +0 at this offset of the slice
                if (!d.hb) atomicOr(hzp + (o >> 6), 1ull << (o & 63));
            }
        };
        if constexpr (!MIXED) {
            auto law = [&](double s) {
                if (law_active<LAW, T>(s, thr, flag)) law_run(s);
            };
            const double Xm0 = P.hx * (double)(kx + d.st.dx);
            // calculatedeformedlengthMex.c:37-43, row-oriented (x_m - x_n); l and 1 / l by the unscaled
            // sequences when every lane's squared length lies in their range (wave-uniform branch)
            const double dx = nbx - qn0, dy = nby - qn1, dz = nbz - qn2;
            const double l2 = dx * dx + dy * dy + dz * dz;
            const bool core = __builtin_amdgcn_ballot_w64(!(l2 >= CORE_LO && l2 <= CORE_HI)) == 0;
            double l, rl;
            if (core) {
                l = sqrt_core(l2);
                rl = rcp_core(l);
            } else {
                l = sqrt(l2);
                rl = 1.0 / l;
            }
            // buildhorizons.m:93-97: (ex^2 + ey^2) + ez^2, the y and z squares from the row
            const double ex = Xn0 - Xm0;
            double q0;
            if (GEN) {
                const double ey = Xn1 - P.hy * (double)(ky + d.st.dy()), ez = Xn2 - P.hz * (double)(kz + d.st.dz());
                q0 = (ex * ex + ey * ey) + ez * ez;
            } else {
                q0 = (ex * ex + eyy) + ezz;
            }
            double L0, s;
            if (fast64) {
                // sqrt(q0) by one correction step and RN(1 / L0) by two Newton steps from the
                // offset's nominal values (set-up checked both bit for bit for every value q0 takes
                // on this lattice), and (l - L0) / L0 by Markstein's correction with RN(1 / L0)
                L0 = fma(fma(-d.st.Ln, d.st.Ln, q0), d.st.hn, d.st.Ln);
                const double r1 = fma(fma(-L0, d.st.rn, 1.0), d.st.rn, d.st.rn);
                const double r = fma(fma(-L0, r1, 1.0), r1, r1);
                const double a = l - L0, y = a * r;
                s = fma(fma(-L0, y, a), r, y);
            } else {
                L0 = sqrt(q0);
                s = (l - L0) / L0;
            }
            law(s);
            double vcf; // calculatevolumecorrectionfactors.m:20-32
            if (L0 <= P.hmr) vcf = 1;
            else if ((L0 > P.hmr) && (L0 <= P.horizon)) {
                const double a = P.hpr - L0;
                if (fast64) {
                    const double y = a * P.rtwo_r;
                    vcf = fma(fma(-P.two_r, y, a), P.rtwo_r, y);
                } else {
                    vcf = a / P.two_r;
                }
            } else vcf = 0;
            const double fail = (LAW == BBPD_LAW_LINEAR && flag) ? 0.0 : 1.0;
            const double p = fail * c * (1 - fabs(hv)) * s * P.V * vcf;
            // correctly rounded quotients dx/l etc. (see k_bond)
            double qx = dx * rl, qy = dy * rl, qz = dz * rl;
            qx = fma(fma(-l, qx, dx), rl, qx);
            qy = fma(fma(-l, qy, dy), rl, qy);
            qz = fma(fma(-l, qz, dz), rl, qz);
            double gx = p * qx, gy = p * qy, gz = p * qz;
            // NaN -> 0 (calculatebondforcesMex.c:38) can only happen off the unscaled sequences: there
            // every lane's length is at least 2^-383, |dX / l| <= 1, and s, the history, c, V and vcf
            // are finite (a stencil offset is never zero, so L0 > 0)
            if (!core) {
                if (gx != gx) gx = 0.0;
                if (gy != gy) gy = 0.0;
                if (gz != gz) gz = 0.0;
            }
            f0 += gx;
            f1 += gy;
            f2 += gz;
        } else {
            const float et0 = (float)(nbx - qn0), et1 = (float)(nby - qn1), et2 = (float)(nbz - qn2);
            const float y0 = d.st.x + et0, y1 = d.st.y + et1, y2 = d.st.z + et2;
            // |y|^2 - |xi|^2 = eta . (2 xi + eta), without cancellation; with x = that / (2 |xi|^2),
            // |y| = |xi| sqrt(1 + 2x), the stretch is s = sqrt(1 + 2x) - 1 and g = s / (1 + s) =
            // s |xi| / |y| carries the force's 1 / |y|
            const float num = fmaf(et0, fmaf(2.0f, d.st.x, et0),
                                   fmaf(et1, fmaf(2.0f, d.st.y, et1), et2 * fmaf(2.0f, d.st.z, et2)));
            const float x = num * d.st.rq;
            // |x| <= 2^-10: series, g = x - 3/2 x^2 + 5/2 x^3 (the next term, -35/8 x^4, is below 2^-27
            // |g|), and, only where the law runs, s = x - x^2/2 + x^3/2 (next -5/8 x^4, below 2^-30
            // |s|).  The law's test compares x with its threshold mapped to x.  The choice is the bond's
            // own (x is the same from either end), so both ends store the same history.
            float g = x * fmaf(x, fmaf(x, 2.5f, -1.5f), 1.0f);
            // off the series: |y|^2 carries 1e-30 (below every representable bond's square by many
            // orders, so it changes no realistic length): 1/l stays finite for a bond of zero length,
            // whose force p * y is then 0, the reference's NaN -> 0 (calculatebondforcesMex.c:38)
            auto off_series = [&](float &rl) {
                const float l2 = fmaf(y0, y0, fmaf(y1, y1, fmaf(y2, y2, 1e-30f)));
                rl = __builtin_amdgcn_rsqf(l2);
                const float l = l2 * rl;
                return num * __builtin_amdgcn_rcpf(d.st.L0 * (l + d.st.L0)); // the stretch
            };
            if (!(fabsf(x) <= MIX_SERIES_X)) {
                float rl;
                const float so = off_series(rl);
                g = (so * rl) * d.st.L0;
            }
            // the stretch, where the law runs: its series, or the same sequence as above
            if (law_active<LAW, T>(x, thr, flag)) {
                float rl, sl;
                if ((__float_as_uint(x) & 0x7fffffffu) <= __float_as_uint(MIX_SERIES_X))
                    sl = x * fmaf(x, fmaf(x, 0.5f, -0.5f), 1.0f);
                else sl = off_series(rl);
                law_run(sl);
            }
            // kap (1 - H) s / l = (c vcf / |xi|) (1 - H) g; a failed bond of the brittle law carries
            // no force
            const float kap = c * d.st.vr;
            const float p = (LAW == BBPD_LAW_LINEAR && flag) ? 0.0f : fmaf(-kap, fabsf(hv), kap) * g;
            a.x = fmaf(p, y0, a.x);
            a.y = fmaf(p, y1, a.y);
            a.z = fmaf(p, y2, a.z);
        }
    };
    if (noff > 0 && wave_on) {
        RowRecs R;
        int dfetch = sload(rows, 1); // slot delta of the row in flight
        bool one = false;            // tiles: the row in flight loads only its last line
        if constexpr (TILE) {
            fetch_t(R, dfetch, false);
            stage_t(R, false);
            nbase = row_base();
        } else {
            fetch(R, dfetch);
            stage(R);
        }
        int r = 0, rend = sload(rows, 0);
        if (nrow > 1) {
            const int d1 = sload(rows, 3);
            if constexpr (TILE) {
                one = d1 - dfetch == tnx;
                fetch_t(R, d1, one);
            } else {
                fetch(R, d1);
            }
            dfetch = d1;
        }
        // the next row's offset count and the slot delta of the row after it, read a row ahead
        int cnext = sload(rows, 2), dnext = sload(rows, 5);
        // entering offset o: at a row's end, stage the next row and start loading the one after
        auto next_row = [&](int o) {
            if (o == rend) {
                r++;
                if constexpr (TILE) {
                    if (one) sh++;
                    stage_t(R, one);
                    nbase = row_base();
                } else {
                    stage(R);
                }
                rend += cnext;
                if (r + 1 < nrow) {
                    if constexpr (TILE) {
                        one = dnext - dfetch == tnx;
                        fetch_t(R, dnext, one);
                    } else {
                        fetch(R, dnext);
                    }
                    dfetch = dnext;
                }
                cnext = sload(rows, 2 * (r + 1));
                dnext = sload(rows, 2 * (r + 2) + 1);
            }
        };
        StEnt<T> A, B;
        load(A, 0, (uint32_t)hlo);
        // the lane masks are wavefront-uniform: a lane's bit is tested by running the entry under
        // the mask as the execution mask
        // the group's offsets k, k + 1 (o, o + 1) with loads one offset ahead
        auto pair = [&](int o, int k, uint32_t bits, float3 &a) {
            load(B, k + 1, bits);
            next_row(o);
            compute(A, o, k, __builtin_amdgcn_inverse_ballot_w64(A.mk), a);
            load(A, k + 2, bits);
            next_row(o + 1);
            compute(B, o + 1, k + 1, __builtin_amdgcn_inverse_ballot_w64(B.mk), a);
        };
        // groups of four offsets: the mixed form sums a group in FP32 and adds that to the FP64 row sum
        for (int o = 0; o < noff; o += 4) {
            const uint32_t bits = (uint32_t)hlo;
            float3 a = make_float3(0.0f, 0.0f, 0.0f);
            pair(o, 0, bits, a);
            if (o + 2 < noff) pair(o + 2, 2, bits, a);
            if constexpr (MIXED) {
                f0 += (double)a.x;
                f1 += (double)a.y;
                f2 += (double)a.z;
            }
            sp += 4;
            mp += 4;
            hso += 4 * SLICE * TW;
            hlo = (hlo >> 4) | (hhi << 60);
            hhi >>= 4;
        }
    }
    // the block reduction of finish_node (ADR) reuses the row buffers once every wavefront is past
    // its rows: no separate scratch, so less LDS per block and more blocks per CU
    if (P.adr) __syncthreads();
    finish_node<MIXED>(P, n, gb, real, cur, f0, f1, f2, qn0, qn1, qn2, Xn0, Xn1, Xn2, &s_row[0][0]);
}

// ADR: damping coefficient of each active region from its blocks' partials
// (calculatedampingcoefficient.m:90-112); one block per region, a fixed summation order
__global__ __launch_bounds__(BLOCK) void k_adr_cn(Params P)
{
    __shared__ double sh[3 * BLOCK];
    const int ri = P.active[blockIdx.x];
    const Region &RG = P.reg[ri];
    double c1 = 0, c2 = 0, c3 = 0;
    // the loads of eight steps are issued together; each thread still adds its partials in order
    const double2 *part = reinterpret_cast<const double2 *>(P.partials) + 2 * (size_t)RG.blk0;
#pragma unroll 8
    for (int q = threadIdx.x; q < RG.nblk; q += BLOCK) {
        const double2 p = part[2 * q], r = part[2 * q + 1];
        c1 = c1 + p.x;
        c2 = c2 + p.y;
        c3 = c3 + r.x;
    }
    bbpd::block_sum_fixed3<BLOCK>(c1, c2, c3, sh);
    const double s1 = c1, s2 = c2;
    if (threadIdx.x == 0) {
        double cn = 0;
        if (s2 != 0) cn = ((s1 / s2) > 0) ? 2 * sqrt(s1 / s2) : 0;
        if (cn > 2) cn = 1.9;
        RegionState &st = P.rst[ri];
        st.cn = cn;
        st.cn1 = s1;
        st.cn2 = s2;
        st.res2 = c3;
    }
}

// ADR: leapfrog update (timeintegrationADR.m:53-80), then x = X + u.  The hand-over F_prev = F
// (dynamicsolverdisplacementcontrolledADR.m:153-155) is the parity of the force sets (adr_force).
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_adr_update(Params P)
{
    const int gb = launch_block(P);
    if (gb < 0) return;
    const int n = gb * BLOCK + threadIdx.x;
    const int ri = P.blkreg[gb];
    if (n - P.reg[ri].base >= P.reg[ri].nN) return;
    const int cur = (int)(*P.tick & 1), nxt = cur ^ 1;
    const long long it = P.rst[ri].step + 1;
    const double cn = P.rst[ri].cn, DT = P.DT;
    const double minv = 1 / P.mass[n];
    const unsigned char cf = P.cons[n];
    double *vv[3] = {P.v0, P.v1, P.v2};
    const double *FF[3] = {adr_force(P, cur, 0), adr_force(P, cur, 1), adr_force(P, cur, 2)};
    const double *XX[3] = {P.X0, P.X1, P.X2};
    double *uu[3] = {P.u0, P.u1, P.u2};
    const NodeRec *pc = P.pos[cur] + n; // the position only (24 of the record's 32 bytes)
    const double pcv[3] = {pc->x, pc->y, pc->z};
    // every load before any store (the arrays are distinct, which the compiler cannot assume)
    double f[3], vo[3], uold[3], xr[3], out[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
        f[d] = FF[d][n];
        vo[d] = vv[d][n];
        uold[d] = MIXED ? pcv[d] : uu[d][n];
        xr[d] = MIXED ? 0.0 : XX[d][n];
    }
#pragma unroll
    for (int d = 0; d < 3; d++) {
        double nv;
        if (it == 1) nv = 0.5 * DT * minv * f[d];
        else nv = (((2 - cn * DT) * vo[d]) + (2 * DT * minv * f[d])) / (2 + cn * DT);
        double nu = uold[d] + (DT * nv);
        if (cf & (1u << d)) {
            nu = 0;
            nv = 0;
        }
        vv[d][n] = nv;
        if (MIXED) {
            out[d] = nu;
        } else {
            uu[d][n] = nu;
            out[d] = xr[d] + nu;
        }
    }
    NodeRec *po = P.pos[nxt] + n;
    po->x = out[0];
    po->y = out[1];
    po->z = out[2];
}

// ------------------------------------------------------------------------------------------------
// Step end: contact with every rigid body (calculatecontactforce.m:62-121), the load schedule
// (smoothstepdata.m:40-42 or a staircase of increments), the step's observables and step counters.
// One block per (region, body); a region without bodies still gets one block for its observables.
// ------------------------------------------------------------------------------------------------
// load-point displacement of step i: smoothstepdata(i, 1, N, 0, a) (the total displacement at step i);
// a staircase of `increment` every ipi iterations; or the equilibrium staircase, `increment` times
// the region's current increment, which the step end advances once the state is in equilibrium
__device__ inline double schedule_disp(const Region &R, const RegionState &S, long long i)
{
    if (R.schedule == 0) {
        double xi = ((double)i - 1) / ((double)R.nTimeSteps - 1);
        return 0 + (R.applied - 0) * (xi * xi * xi) * (10 - 15 * xi + 6 * (xi * xi));
    }
    if (R.schedule == 2) return R.increment * (double)S.inc;
    long long inc = (i + R.ipi - 1) / R.ipi;
    return R.increment * (double)inc;
}

template <bool MIXED>
__global__ __launch_bounds__(1024) void k_step_end(Params P)
{
    // the last block of a region records the region's observables and advances its step counter;
    // the last block of the launch advances the position-buffer parity
    constexpr int SE_CH = 4; // family chunks of blockDim.x members per group
    __shared__ double vals[SE_CH * 1024];
    __shared__ int wofs[SE_CH * 16 + 1]; // per (chunk, wavefront) contact offsets, then the count
    __shared__ int last_block, last_all;
    const int4 job = P.se[blockIdx.x];
    const int ri = job.x, bi = job.y;
    const Region &RG = P.reg[ri];
    RegionState &RS = P.rst[ri];
    const int nxt = (int)((*P.tick + 1) & 1);
    const long long it = RS.step + 1;
    NodeRec *pos = P.pos[nxt];
    double *uu[3] = {P.u0, P.u1, P.u2};
    double *vv[3] = {P.v0, P.v1, P.v2};
    const double *XX[3] = {P.X0, P.X1, P.X2};
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    double total = 0, rs = 0;
    // one body's contact (calculatecontactforce.m:62-121): its reaction in `total`, and for body 0
    // the static reaction in `rs`; gb is the body's index in P.bodies (the job carries it, so that
    // the body's record does not wait for the region's)
    auto body = [&](const int bi, const int gb) {
        const Body &B = P.bodies[gb];
        const double disp = schedule_disp(RG, RS, it);
        const double cx = B.cx, czm = B.cz + (B.moves ? disp : 0.0), R = B.R;
        const int nf = B.nFam;
        double part = 0; // body 0: this thread's share of the static reaction, members in order
        // the members in groups of SE_CH chunks of blockDim.x: every load of a group is issued
        // before any of its members is projected, so a family costs about one memory latency per
        // group rather than one chain of dependent loads per chunk
        for (int g0 = 0; g0 < nf; g0 += SE_CH * (int)blockDim.x) {
            int idx[SE_CH];
            NodeRec pr[SE_CH];
            double x0[SE_CH], x2[SE_CH], fz[SE_CH];
            bool hit[SE_CH];
            double d_z[SE_CH];
#pragma unroll
            for (int k = 0; k < SE_CH; k++) {
                const int t = g0 + k * (int)blockDim.x + (int)threadIdx.x;
                idx[k] = t < nf ? B.fam[t] : -1;
            }
#pragma unroll
            for (int k = 0; k < SE_CH; k++)
                if (idx[k] >= 0) {
                    pr[k] = pos[idx[k]];
                    if (bi == 0) fz[k] = P.adr ? adr_force(P, nxt ^ 1, 2)[idx[k]] : P.F2[idx[k]];
                    if (MIXED) {
                        x0[k] = XX[0][idx[k]];
                        x2[k] = XX[2][idx[k]];
                    }
                }
#pragma unroll
            for (int k = 0; k < SE_CH; k++) {
                hit[k] = false;
                d_z[k] = 0;
                if (idx[k] < 0) continue;
                const int i = idx[k];
                // calculatereactionforceFast.m:3-5 over body 0's family: F_z * cellVolume
                if (bi == 0) part = part + fz[k] * P.V;
                double pv[3] = {pr[k].x, pr[k].y, pr[k].z};
                double xc[3];
                xc[0] = MIXED ? x0[k] + pv[0] : pv[0];
                xc[2] = MIXED ? x2[k] + pv[2] : pv[2];
                const double ddx = xc[0] - cx, ddz = xc[2] - czm;
                const double dist = sqrt((ddx * ddx) + (ddz * ddz));
                if (dist < R) {
                    hit[k] = true;
                    xc[1] = MIXED ? XX[1][i] + pv[1] : pv[1];
                    xc[0] = cx + (ddx / dist) * R;
                    xc[2] = czm + (ddz / dist) * R;
#pragma unroll
                    for (int d = 0; d < 3; d++) {
                        const double uo = MIXED ? pv[d] : uu[d][i];
                        const double vo = vv[d][i];
                        const double nu = xc[d] - XX[d][i];
                        const double nv = (nu - uo) / P.DT;
                        if (MIXED) {
                            pv[d] = nu;
                        } else {
                            uu[d][i] = nu;
                            pv[d] = xc[d];
                        }
                        vv[d][i] = nv;
                        if (d == 2) d_z[k] = -1 * P.rho[i] * (nv - vo) / P.DT * P.V;
                    }
                    pos[i].x = pv[0];
                    pos[i].y = pv[1];
                    pos[i].z = pv[2];
                }
            }
            // compact the group's members in contact, in family order (chunk, wavefront, lane), then
            // sum them in that order (the order of MATLAB's sum over dpenfz; members not in contact
            // add nothing).  The SE_CH * nw <= 64 counts are scanned by one wavefront.
            unsigned long long mask[SE_CH];
#pragma unroll
            for (int k = 0; k < SE_CH; k++) {
                mask[k] = __ballot(hit[k]);
                if (lane == 0) wofs[k * nw + wv] = __popcll(mask[k]);
            }
            __syncthreads();
            if (wv == 0) {
                const int c = lane < SE_CH * nw ? wofs[lane] : 0;
                int incl = c;
#pragma unroll
                for (int d = 1; d < 64; d <<= 1) {
                    const int t = __shfl_up(incl, d);
                    if (lane >= d) incl += t;
                }
                if (lane < SE_CH * nw) wofs[lane] = incl - c; // exclusive offsets
                if (lane == 63) wofs[SE_CH * 16] = incl;     // the group's contacts
            }
            __syncthreads();
            const int ncon = wofs[SE_CH * 16];
#pragma unroll
            for (int k = 0; k < SE_CH; k++)
                if (hit[k]) vals[wofs[k * nw + wv] + __popcll(mask[k] & ((1ull << lane) - 1))] = d_z[k];
            __syncthreads();
            if (threadIdx.x == 0) {
#pragma unroll 8
                for (int q = 0; q < ncon; q++) total = total + vals[q];
            }
            __syncthreads();
        }
        // the static reaction: the threads' shares summed in a fixed tree
        if (bi == 0) rs = bbpd::block_sum_fixed<1024>(part, vals);
    };
    if (bi >= 0) {
        body(bi, job.z);
    } else if (bi == -2) {
        // overlapping families: this block applies the bodies in order, as the reference's calls
        // follow one another (dynamicsolverdisplacementcontrolled.m:108-112); a member shared with an
        // earlier body is seen as that body left it (the barrier orders the block's global writes)
        for (int b = 0; b < RG.nBodies; b++) {
            total = 0;
            body(b, RG.body0 + b);
            if (threadIdx.x == 0) RS.reaction[b] = total;
            if (b == 0 && threadIdx.x == 0) RS.reaction_static = rs;
            __syncthreads();
        }
    }
    // the block's two tickets, taken together: the region's (its last block records the region's
    // observables) and the launch's (its last block advances the parity, which no block reads after
    // its start)
    if (threadIdx.x == 0) {
        if (bi >= 0) RS.reaction[bi] = total;
        if (bi == 0) RS.reaction_static = rs;
        __threadfence();
        const unsigned need = (unsigned)(RG.nBodies > 0 && !RG.seq_contact ? RG.nBodies : 1);
        const unsigned ticket = atomicAdd(&RS.done, 1u);
        const unsigned tall = atomicAdd(P.tick_done, 1u);
        last_block = (ticket == need - 1);
        last_all = (tall == (unsigned)gridDim.x - 1);
    }
    __syncthreads();
    if (last_block && threadIdx.x == 0) {
        __threadfence();
        volatile RegionState *vrs = &RS;
        auto comp = [&](int slot, int d) -> double {
            if (MIXED) {
                const volatile double *pp = reinterpret_cast<const volatile double *>(pos + slot);
                return pp[d];
            }
            const volatile double *pu = uu[d];
            return pu[slot];
        };
        const double defl = RG.ref_slot >= 0 ? comp(RG.ref_slot, 2) : 0.0;
        const double cmod = (RG.cmod1 >= 0 && RG.cmod2 >= 0) ? comp(RG.cmod2, 0) - comp(RG.cmod1, 0) : 0.0;
        const double r0 = RG.nBodies > 0 ? vrs->reaction[0] : 0.0;
        const double rstat = vrs->reaction_static;
        // this step's displacement: the increment is read before this block advances it
        const double disp = RG.schedule == 2 ? RG.increment * (double)vrs->inc : schedule_disp(RG, RS, it);
        // equilibrium ratio of this step's forces (after calculateequilibriumstate.m:6-8): the
        // out-of-balance force over the free DOFs against the load, +infinity while the load is 0
        const double aload = fabs(rstat);
        const double ratio = aload > 0 ? (sqrt(vrs->res2) * P.V) / aload : INFINITY;
        bool inc_end = false;
        if (RG.schedule == 1) {
            inc_end = it % RG.ipi == 0;
        } else if (RG.schedule == 2) {
            const long long k = vrs->inc_iters + 1;
            inc_end = (k >= RG.ipi_min && ratio < RG.eq_tol) || k >= RG.ipi;
            vrs->inc_iters = inc_end ? 0 : k;
            if (inc_end) vrs->inc = vrs->inc + 1;
        }
        double *h = P.hist + (size_t)(RG.hist0 + (it - 1) % RG.hist_cap) * N_HIST;
        h[0] = r0;
        h[1] = rstat;
        h[2] = defl;
        h[3] = cmod;
        h[4] = disp;
        h[5] = ratio;
        h[6] = P.adr ? vrs->cn : 0.0;
        h[7] = inc_end ? 1.0 : 0.0;
        // stop criterion, judged on converged states: the ends of load increments
        if (inc_end && !vrs->finished) {
            const double load = aload;
            const double peak = vrs->peak;
            bool stop = false;
            if (RG.stop_frac > 0 && peak > 0 && load < RG.stop_frac * peak) stop = true;
            if (RG.stop_disp > 0 && fabs(disp) >= RG.stop_disp) stop = true;
            if (load > peak) vrs->peak = load;
            if (stop) vrs->finished = 1;
        }
        vrs->done = 0u;
        vrs->step = it;
    }
    if (last_all && threadIdx.x == 0) {
        *P.tick_done = 0u;
        *P.tick = *P.tick + 1;
    }
}

// observables over entries: softening flags, broken (H == 1 or failed), failed; each bond twice
template <bool MIXED>
__global__ __launch_bounds__(BLOCK) void k_count(Params P, int ri, unsigned long long *out)
{
    const Region &RG = P.reg[ri];
    const int n = (RG.blk0 + blockIdx.x) * BLOCK + threadIdx.x;
    const bool real = n - RG.base < RG.nN;
    unsigned long long cs = 0, cb = 0, cf = 0;
    if (real && P.stencil) {
        // stencil layout: history per (slot, offset), sign bit = flag (failed for the brittle law)
        const uint4 msk = P.smask[n];
        const uint32_t mw[4] = {msk.x, msk.y, msk.z, msk.w};
        const int noff = RG.noff;
        const long long ebase = RG.e0 + (long long)((n - RG.base) >> 6) * noff * SLICE + (n & (SLICE - 1));
        for (int o = 0; o < noff; o++) {
            if (!((mw[o >> 5] >> (o & 31)) & 1u)) continue;
            const long long e = ebase + (long long)o * SLICE;
            double h;
            if (MIXED) h = (double)reinterpret_cast<const float *>(P.hs)[e];
            else h = reinterpret_cast<const double *>(P.hs)[e];
            const bool sgn = std::signbit(h);
            const bool failed = P.law_linear && sgn;
            cs += (!P.law_linear && sgn) ? 1 : 0;
            cf += failed ? 1 : 0;
            cb += (failed || std::fabs(h) == 1.0) ? 1 : 0;
        }
    } else if (real) {
        const int len = P.rowlen[n];
        long long e = P.slice_off[n >> 6] + (n & (SLICE - 1));
        for (int kk = 0; kk < len; kk++, e += SLICE) {
            const uint4 r = P.rec[e];
            const uint32_t w = r.x;
            const bool one = MIXED ? (__uint_as_float(r.z) == 1.0f) : (__hiloint2double((int)r.w, (int)r.z) == 1.0);
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

// ------------------------------------------------------------------------------------------------
// Damage field (calculatedamage.m:34-45) and dissipated damage energy (calculatedamageenergy.m:35-62)
// of one region, on demand.  The reference updates a per-bond energy history every step as
// max(de(s)); de is non-decreasing in the stretch, so the history is de at the largest stretch
// reached.  Every softening law here is non-decreasing in the stretch on (s0, sc] and the stored
// factor H is its largest value so far, so that stretch is the law inverted at H (bisection).
// ------------------------------------------------------------------------------------------------
template <int LAW>
__device__ double law_factor(double s, double s0, double sc, double s1, double k, double alpha, double denom)
{
    if (LAW == BBPD_LAW_BILINEAR) return ((s - s0) / s) * (sc / (sc - s0));
    if (LAW == BBPD_LAW_TRILINEAR) {
        const double beta = 0.25, eta = s1 / s0;
        if (s <= s1) return (1 - ((eta - beta) / (eta - 1) * (s0 / s))) + ((1 - beta) / (eta - 1));
        return 1 - ((s0 * beta) / s) * ((sc - s) / (sc - s1));
    }
    const double t = (s - s0) / (sc - s0);
    const double numerator = 1 - exp(-k * t);
    const double residual = alpha * (1 - t);
    return 1 - (s0 / s) * ((1 - (numerator / denom)) + residual) / (1 + alpha);
}

// the largest stretch a bond has reached, from its factor H (s0 when it never softened)
template <int LAW>
__device__ double max_stretch(double H, double s0, double sc, double s1, double k, double alpha, double denom)
{
    if (!(H > 0)) return s0;
    if (H >= 1) return sc;
    double lo = s0, hi = sc;
    for (int it = 0; it < 64; it++) {
        const double mid = 0.5 * (lo + hi);
        if (law_factor<LAW>(mid, s0, sc, s1, k, alpha, denom) < H) lo = mid;
        else hi = mid;
    }
    return 0.5 * (lo + hi);
}

// calculatedamageenergy.m:47-57 at stretch s
__device__ double damage_energy(double c, double s0, double sc, double L0, double s, double k, double alpha, double V)
{
    const double denominator = 2 * k * (exp(k) - 1) * (1 + alpha);
    const double shape = (2 * k) - (2 * exp(k)) + (alpha * k) - (alpha * k * exp(k)) + 2;
    double numerator;
    if ((s0 < s) && (s <= sc)) numerator = c * s0 * L0 * (s0 - s) * shape;
    else if (s > sc) numerator = c * s0 * L0 * (s0 - sc) * shape;
    else return 0.0;
    return (numerator / denominator) * (V * V);
}

// Per node: the damage field (calculatedamage.m:34-45), the damage energy of the bonds at their lower
// slot (block partials), and, when sed is given, the strain energy density of
// calculatestrainenergy.m:10-24: each bond's micropotential 0.5 fail c s^2 L0 V vcf (left to right, s
// from the current positions in FP64) added to both its nodes in BONDLIST order, halved.
template <bool MIXED, int LAW>
__global__ __launch_bounds__(BLOCK) void k_damage(Params P, int ri, double *dmg, double *epart, double *sed)
{
    __shared__ double sh[BLOCK];
    const Region &RG = P.reg[ri];
    const int n = (RG.blk0 + blockIdx.x) * BLOCK + threadIdx.x;
    const bool real = n - RG.base < RG.nN;
    const int cur = (int)(*P.tick & 1);
    const NodeRec *pos = P.pos[cur];
    double e_acc = 0, mu_acc = 0;
    int cnt = 0, unbroken = 0;
    if (real) {
        const NodeRec me = pos[n];
        double Xn[3];
        if (P.Xrec) {
            Xn[0] = P.X0[n], Xn[1] = P.X1[n], Xn[2] = P.X2[n];
        } else {
            Xn[0] = P.hx * (double)(int)(me.q & LAT_MASK);
            Xn[1] = P.hy * (double)(int)((me.q >> LAT_BITS) & LAT_MASK);
            Xn[2] = P.hz * (double)(int)((me.q >> (2 * LAT_BITS)) & LAT_MASK);
        }
        // deformed coordinates: the FP64 form stores them, the mixed form stores u (x = X + u)
        const double xn0 = MIXED ? Xn[0] + me.x : me.x, xn1 = MIXED ? Xn[1] + me.y : me.y,
                     xn2 = MIXED ? Xn[2] + me.z : me.z;
        // gL0 / gvcf: the given undeformed length and volume correction factor (streamed, FP64 record
        // layout), or negative to recompute them from the coordinates
        auto bond = [&](double H, bool failed, int m, double c, double s0, double sc, double Xm0, double Xm1,
                        double Xm2, double gL0 = -1.0, double gvcf = -1.0) {
            const bool broken = failed || H == 1.0;
            cnt++;
            unbroken += broken ? 0 : 1;
            if (sed) {
                // buildhorizons.m:93-97, calculatedeformedlengthMex.c:40-43 and
                // calculatevolumecorrectionfactors.m:20-32, then the micropotential of
                // calculatestrainenergy.m:16 (BFMULTIPLIER = 1 for concrete, buildbonddata.m:56-57)
                const double ex = Xn[0] - Xm0, ey = Xn[1] - Xm1, ez = Xn[2] - Xm2;
                const double L0 = gL0 >= 0 ? gL0 : sqrt((ex * ex + ey * ey) + ez * ez);
                const NodeRec pm = pos[m];
                const double dx = (MIXED ? Xm0 + pm.x : pm.x) - xn0, dy = (MIXED ? Xm1 + pm.y : pm.y) - xn1,
                             dz = (MIXED ? Xm2 + pm.z : pm.z) - xn2;
                const double l = sqrt(dx * dx + dy * dy + dz * dz);
                const double st = (l - L0) / L0;
                double vcf;
                if (gvcf >= 0) vcf = gvcf;
                else if (L0 <= P.hmr) vcf = 1;
                else if ((L0 > P.hmr) && (L0 <= P.horizon)) vcf = (P.hpr - L0) / P.two_r;
                else vcf = 0;
                const double fl = (LAW == BBPD_LAW_LINEAR && failed) ? 0.0 : 1.0;
                mu_acc = mu_acc + 0.5 * fl * 1.0 * c * (st * st) * L0 * P.V * vcf;
            }
            if (m > n && LAW != BBPD_LAW_LINEAR) { // each bond once, at its lower slot
                const double ex = Xn[0] - Xm0, ey = Xn[1] - Xm1, ez = Xn[2] - Xm2;
                const double L0 = gL0 >= 0 ? gL0 : sqrt((ex * ex + ey * ey) + ez * ez);
                const double smax = max_stretch<LAW>(H, s0, sc, RG.s1, P.k, P.alpha, P.denom_exp);
                e_acc += damage_energy(c, s0, sc, L0, smax, P.k, P.alpha, P.V);
            }
        };
        if (P.stencil) {
            const uint4 msk = P.smask[n];
            const uint32_t mw[4] = {msk.x, msk.y, msk.z, msk.w};
            const long long ebase = RG.e0 + (long long)((n - RG.base) >> 6) * RG.noff * SLICE + (n & (SLICE - 1));
            const int gme = (int)(me.q >> G_SHIFT);
            for (int o = 0; o < RG.noff; o++) {
                if (!((mw[o >> 5] >> (o & 31)) & 1u)) continue;
                const int4 so = P.soff[RG.off0 + o];
                const long long e = ebase + (long long)o * SLICE;
                const int m = n + so.x;
                const double h = MIXED ? (double)reinterpret_cast<const float *>(P.hs)[e]
                                       : reinterpret_cast<const double *>(P.hs)[e];
                const bool sgn = signbit(h);
                const int key = gme * RG.kg + (int)(pos[m].q >> G_SHIFT);
                const double c = P.tab[RG.tab0 + key];
                const double s0 = P.s0mode == PM_TABLE ? P.tab[RG.tab0 + RG.ts0 + key]
                                  : (P.s0mode == PM_STREAM ? (MIXED ? (double)((const float *)P.s0str)[e]
                                                                    : ((const double *)P.s0str)[e])
                                                           : RG.s0);
                const double sc = P.scmode == PM_TABLE ? P.tab[RG.tab0 + RG.tsc + key]
                                  : (P.scmode == PM_STREAM ? (MIXED ? (double)((const float *)P.scstr)[e]
                                                                    : ((const double *)P.scstr)[e])
                                                           : RG.sc);
                bond(fabs(h), P.law_linear && sgn, m, c, s0, sc, P.hx * (double)((int)(me.q & LAT_MASK) + so.y),
                     P.hy * (double)((int)((me.q >> LAT_BITS) & LAT_MASK) + so.z),
                     P.hz * (double)((int)((me.q >> (2 * LAT_BITS)) & LAT_MASK) + so.w));
            }
        } else {
            const int len = P.rowlen[n];
            long long e = P.slice_off[n >> 6] + (n & (SLICE - 1));
            for (int kk = 0; kk < len; kk++, e += SLICE) {
                const uint4 r = P.rec[e];
                const uint32_t w = r.x;
                const int m = (int)(w & NODE_MASK);
                double H, c, s0, sc;
                if (MIXED) {
                    H = (double)__uint_as_float(r.z);
                    s0 = P.s0mode == PM_INLINE ? (double)__uint_as_float(r.y)
                         : (P.s0mode == PM_STREAM ? (double)((const float *)P.s0str)[e] : RG.s0);
                    sc = P.scmode == PM_INLINE ? (double)__uint_as_float(r.y)
                         : (P.scmode == PM_STREAM ? (double)((const float *)P.scstr)[e] : RG.sc);
                    c = 0; // set below from c * V * vcf
                } else {
                    H = __hiloint2double((int)r.w, (int)r.z);
                    c = P.cmode == PM_TABLE ? P.tab[r.y & 0xFFFFu] : P.cstr[e];
                    s0 = P.s0mode == PM_TABLE ? P.tab[P.ntabc + (r.y >> 16)]
                         : (P.s0mode == PM_STREAM ? ((const double *)P.s0str)[e] : RG.s0);
                    sc = P.scmode == PM_TABLE ? P.tab[P.ntabc + (r.y >> 16)]
                         : (P.scmode == PM_STREAM ? ((const double *)P.scstr)[e] : RG.sc);
                }
                double Xm[3];
                if (P.Xrec) {
                    Xm[0] = P.X0[m], Xm[1] = P.X1[m], Xm[2] = P.X2[m];
                } else {
                    const uint64_t q = pos[m].q;
                    Xm[0] = P.hx * (double)(int)(q & LAT_MASK);
                    Xm[1] = P.hy * (double)(int)((q >> LAT_BITS) & LAT_MASK);
                    Xm[2] = P.hz * (double)(int)((q >> (2 * LAT_BITS)) & LAT_MASK);
                }
                if (MIXED) {
                    // the mixed record holds c * V * vcf in FP32
                    const double ex = Xn[0] - Xm[0], ey = Xn[1] - Xm[1], ez = Xn[2] - Xm[2];
                    const double L0 = sqrt((ex * ex + ey * ey) + ez * ez);
                    double vcf;
                    if (L0 <= P.hmr) vcf = 1;
                    else if ((L0 > P.hmr) && (L0 <= P.horizon)) vcf = (P.hpr - L0) / P.two_r;
                    else vcf = 0;
                    c = vcf > 0 ? (double)__uint_as_float(r.w) / (P.V * vcf) : 0.0;
                }
                if (!MIXED && P.L0str) bond(H, (w & FAILED_BIT) != 0, m, c, s0, sc, Xm[0], Xm[1], Xm[2], P.L0str[e], P.vcfstr[e]);
                else bond(H, (w & FAILED_BIT) != 0, m, c, s0, sc, Xm[0], Xm[1], Xm[2]);
            }
        }
        if (dmg) dmg[n - RG.base] = cnt > 0 ? 1 - ((double)unbroken / (double)cnt) : 0.0;
        if (sed) sed[n - RG.base] = 0.5 * mu_acc;
    }
    const double sum = bbpd::block_sum_fixed<BLOCK>(e_acc, sh);
    if (threadIdx.x == 0) epart[blockIdx.x] = sum;
}



void launch_count(bbpd_solver *s, int ri)
{
    const unsigned g = (unsigned)s->reg[ri].nblk;
    if (s->precision == BBPD_PRECISION_MIXED)
        hipLaunchKernelGGL(k_count<true>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, ri, s->counts);
    else
        hipLaunchKernelGGL(k_count<false>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, ri, s->counts);
}

using DamageFn = void (*)(Params, int, double *, double *, double *);

template <bool MIXED>
DamageFn pick_damage(int law)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return k_damage<MIXED, BBPD_LAW_LINEAR>;
    case BBPD_LAW_BILINEAR: return k_damage<MIXED, BBPD_LAW_BILINEAR>;
    case BBPD_LAW_TRILINEAR: return k_damage<MIXED, BBPD_LAW_TRILINEAR>;
    default: return k_damage<MIXED, BBPD_LAW_EXPONENTIAL>;
    }
}

void launch_damage(bbpd_solver *s, int ri, double *dmg, double *epart, double *sed)
{
    DamageFn f = s->precision == BBPD_PRECISION_MIXED ? pick_damage<true>(s->law) : pick_damage<false>(s->law);
    hipLaunchKernelGGL(f, dim3(s->reg[ri].nblk), dim3(BLOCK), 0, s->stream, s->P, ri, dmg, epart, sed);
}

template <bool MIXED, int LAW>
KernelFn pick_lat(bool lat, bool given)
{
    // given lengths and factors (Params::L0str): the FP64 form, off-lattice inputs only
    if constexpr (!MIXED)
        if (given) return k_bond<false, LAW, false, true>;
    return lat ? k_bond<MIXED, LAW, true> : k_bond<MIXED, LAW, false>;
}

template <bool MIXED>
KernelFn pick_law(int law, bool lat, bool given)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return pick_lat<MIXED, BBPD_LAW_LINEAR>(lat, given);
    case BBPD_LAW_BILINEAR: return pick_lat<MIXED, BBPD_LAW_BILINEAR>(lat, given);
    case BBPD_LAW_TRILINEAR: return pick_lat<MIXED, BBPD_LAW_TRILINEAR>(lat, given);
    default: return pick_lat<MIXED, BBPD_LAW_EXPONENTIAL>(lat, given);
    }
}

// launch variants of the stencil bond kernel: (minimum waves per SIMD, block order); orders: XCD
// remap, launch order, and launch order over the stripe list (variants 4 and 5)

template <bool MIXED, int LAW, int PV, bool TILE>
KernelFn pick_occ(int v)
{
    switch (v) {
    case 1: return k_bond_st<MIXED, LAW, 1, false, PV, TILE>;
    case 2: return k_bond_st<MIXED, LAW, 5, true, PV, TILE>;
    case 3: return k_bond_st<MIXED, LAW, 5, false, PV, TILE>;
    case 4: return k_bond_st<MIXED, LAW, 1, false, PV, TILE>;
    case 5: return k_bond_st<MIXED, LAW, 5, false, PV, TILE>;
    default: return k_bond_st<MIXED, LAW, 1, true, PV, TILE>;
    }
}

// The launch list in XCD stripes: launch block b runs on XCD b % 8 (the hardware's placement), and
// the b-th launch takes entry j % S of stripe b % 8 of group j / S (j = b / 8, S = ST_STRIPE) of the
// list, so each XCD walks runs of S consecutive list entries (neighbour rows shared in its L2) while
// a localised process zone still spreads over all XCDs.  -1 = idle.
std::vector<int> stripe_list(const std::vector<int> &list)
{
    const int L = (int)list.size(), M = 8 * ST_STRIPE;
    std::vector<int> r((size_t)((L + M - 1) / M) * M, -1);
    for (int b = 0; b < (int)r.size(); b++) {
        const int j = b >> 3, g = (j / ST_STRIPE) * M + (b & 7) * ST_STRIPE + j % ST_STRIPE;
        r[b] = g < L ? list[g] : -1;
    }
    return r;
}

// BBPD_LAUNCH_VARIANT=v in the environment fixes the stencil bond kernel's launch variant at set-up
// (the autotuner may change it later); -1 = unset
int env_launch_variant()
{
    const char *e = std::getenv("BBPD_LAUNCH_VARIANT");
    if (!e || !*e) return -1;
    const int v = std::atoi(e);
    return v >= 0 && v < N_OCC ? v : -1;
}

// Launch the solver's bond kernel `kb` of launch variant v
void launch_bond(bbpd_solver *s, KernelFn kb, int v)
{
    if (v >= 4) {
        Params Pb = s->P;
        Pb.grid = s->d_grid_st;
        hipLaunchKernelGGL(kb, dim3(s->grid_n_st), dim3(BLOCK), s->lds_bytes, s->stream, Pb);
    } else {
        hipLaunchKernelGGL(kb, dim3(s->grid_n), dim3(BLOCK), s->lds_bytes, s->stream, s->P);
    }
}

template <bool MIXED, int PV, bool TILE>
KernelFn pick_law_st(int law, int v)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return pick_occ<MIXED, BBPD_LAW_LINEAR, PV, TILE>(v);
    case BBPD_LAW_BILINEAR: return pick_occ<MIXED, BBPD_LAW_BILINEAR, PV, TILE>(v);
    case BBPD_LAW_TRILINEAR: return pick_occ<MIXED, BBPD_LAW_TRILINEAR, PV, TILE>(v);
    default: return pick_occ<MIXED, BBPD_LAW_EXPONENTIAL, PV, TILE>(v);
    }
}

template <bool MIXED, bool TILE>
KernelFn pick_pv_st(int pv, int law, int v)
{
    switch (pv) {
    case PV_SCALAR: return pick_law_st<MIXED, PV_SCALAR, TILE>(law, v);
    case PV_TABLE: return pick_law_st<MIXED, PV_TABLE, TILE>(law, v);
    case PV_GENERAL: return pick_law_st<MIXED, PV_GENERAL, TILE>(law, v);
    default: return pick_law_st<MIXED, PV_GENTAB, TILE>(law, v);
    }
}

KernelFn pick_bond_kernel(const bbpd_solver *s, int variant)
{
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int v = variant >= 0 ? variant : s->occ;
    if (s->stencil) {
        const int m0 = s->P.s0mode, mc = s->P.scmode;
        const int pv = (m0 == PM_SCALAR && mc == PM_SCALAR) ? PV_SCALAR
                       : (m0 == PM_TABLE && mc == PM_TABLE) ? PV_TABLE
                       : (m0 == PM_TABLE || mc == PM_TABLE) ? PV_GENTAB : PV_GENERAL;
        // tiles are built for the mixed form only (the FP64 form keeps the slices' reduction order,
        // which the oracle's device-order sums follow)
        if (mixed) return s->tiles ? pick_pv_st<true, true>(pv, s->law, v) : pick_pv_st<true, false>(pv, s->law, v);
        return pick_pv_st<false, false>(pv, s->law, v);
    }
    return mixed ? pick_law<true>(s->law, s->lattice, false) : pick_law<false>(s->law, s->lattice, s->P.L0str != nullptr);
}

// Let every launch variant of the solver's bond kernel take its key tables in dynamic LDS beside
// its static LDS (the stencil kernel's row buffers), past the default per-block limit if needed.
int allow_lds(bbpd_solver *s)
{
    if (s->lds_bytes == 0) return BBPD_OK;
    for (int v = 0; v < (s->stencil ? N_OCC : 1); v++)
        BBPD_HIP(hipFuncSetAttribute(reinterpret_cast<const void *>(pick_bond_kernel(s, v)),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)s->lds_bytes));
    return BBPD_OK;
}

int launch_step(bbpd_solver *s, bool allow_timing)
{
    if (s->nactive == 0) return BBPD_OK; // nothing to step
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
    launch_bond(s, kb, s->stencil ? s->occ : 0);
    BBPD_HIP(hipGetLastError());
    if (timed) {
        BBPD_HIP(hipEventRecord(s->ev[s->ev_used + 1], s->stream));
        s->ev_used += 2;
    }
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    if (s->integrator == BBPD_INTEGRATOR_ADR) {
        hipLaunchKernelGGL(k_adr_cn, dim3(s->nactive), dim3(BLOCK), 0, s->stream, s->P);
        BBPD_HIP(hipGetLastError());
        if (mixed) hipLaunchKernelGGL(k_adr_update<true>, dim3(s->grid_n), dim3(BLOCK), 0, s->stream, s->P);
        else hipLaunchKernelGGL(k_adr_update<false>, dim3(s->grid_n), dim3(BLOCK), 0, s->stream, s->P);
        BBPD_HIP(hipGetLastError());
    }
    if (mixed) hipLaunchKernelGGL(k_step_end<true>, dim3(s->nse), dim3(1024), 0, s->stream, s->P);
    else hipLaunchKernelGGL(k_step_end<false>, dim3(s->nse), dim3(1024), 0, s->stream, s->P);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

} // namespace bbpd_sv
