// bbpd_solver.hip -- layer 2: the fused, device-resident explicit step.
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
//     reference formulas (buildhorizons.m:93-97, calculatevolumecorrectionfactors.m:16-34); set-up
//     checks bit for bit that the given arrays equal them.
//   * A step reads position buffer step&1 and writes the other, so the Euler-Cromer update of a node
//     runs in the kernel that gathers its neighbours.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>
#include <type_traits>
#include <unordered_map>
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
constexpr int N_HIST = 8;      // reaction, static reaction, deflection, CMOD, load-point displacement,
                               // equilibrium ratio, damping coefficient, 1 at a load-increment end
constexpr int TAB_MAX = 2048;  // entries of one LDS parameter table
constexpr int LAT_BITS = 18;   // bits per packed lattice coordinate
constexpr uint64_t LAT_MASK = (1ull << LAT_BITS) - 1;
constexpr int G_SHIFT = 56;    // node record: the node's bond count (nFAMILYMEMBERS) in the top byte
constexpr int ST_MAX = 128;    // offsets of the stencil layout (bits of the per-node mask)
constexpr int ST_GMAX = 45;    // distinct family sizes of the stencil layout (key table <= 45^2)
constexpr int ST_RMAX = 8;     // largest |dx| of a stencil offset
constexpr int ST_ROW = 80;     // sites of a staged stencil row: 64 + 2 * ST_RMAX
// node-flag bits 0-2: DOF d constrained; bit 3: member of a contact family (its DOFs are driven by
// the rigid-body projection, so the ADR Rayleigh quotient leaves the node out)
constexpr unsigned char CONTACT_BIT = 1u << 3;

typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x8 __attribute__((ext_vector_type(8)));
template <typename T>
using T2 = T __attribute__((ext_vector_type(2)));

// how a per-bond quantity reaches the kernel
enum ParamMode { PM_SCALAR = 0, PM_TABLE = 1, PM_STREAM = 2, PM_INLINE = 3 };
// stencil bond kernel variants by where s0 and sc come from
enum ParamVariant { PV_SCALAR = 0, PV_TABLE = 1, PV_GENERAL = 2 };

// A region is one problem instance: a block-aligned range of slots with its own stencil, tables,
// law scalars, contact bodies, load schedule, history and step counter.  A single problem is one
// region; a batch packs many into one launch of each kernel.
struct Region {
    int base, nN, nSlots;          // slots [base, base + nSlots); the first nN hold nodes
    int blk0, nblk;                // global blocks [blk0, blk0 + nblk)
    long long e0;                  // stencil layout: first entry of the region
    int noff, off0;                // stencil layout: offsets off0 .. off0 + noff - 1
    int nrow, row0, rx;            // stencil layout: rows row0 .. row0 + nrow - 1, largest |dx|
    int sst0;                      // stencil layout: first per-offset record
    int kg, tab0, ntab, ts0, tsc;  // stencil layout: key tables (ts0 / tsc relative to tab0)
    int fast64;                    // stencil layout, FP64 form: undeformed lengths, their reciprocals
                                   // and the volume correction factors from the per-offset nominal
                                   // values (checked bit for bit at set-up)
    double s0, s1, sc;             // scalar law parameters
    double s0lo, sclo;             // stencil layout: the smallest streamed s0 / sc of the region
    int body0, nBodies;            // contact bodies body0 .. body0 + nBodies - 1
    int seq_contact;               // the families overlap: one step-end block applies them in order
    int schedule, nTimeSteps, ipi; // load schedule (ipi: iterations per increment, at most, for the
    double applied, increment;     // equilibrium staircase)
    double eq_tol;                 // equilibrium staircase: advance below this equilibrium ratio ...
    int ipi_min;                   // ... after at least this many iterations at the increment
    int ref_slot, cmod1, cmod2;    // observation slots, -1 = none
    long long hist0;               // history ring rows hist0 .. hist0 + hist_cap - 1
    int hist_cap;
    double stop_frac, stop_disp;   // finish when the load falls below stop_frac of its peak after
                                   // the peak, or the load point passes stop_disp (0 = never)
};

struct RegionState {
    long long step;     // steps completed
    unsigned int done;  // step-end blocks of this region finished this step
    double cn, cn1, cn2;
    double res2;        // sum of F_d^2 over the free DOFs (the equilibrium ratio's numerator squared)
    long long inc;      // equilibrium staircase: the current load increment (from 1) ...
    long long inc_iters;// ... and the iterations done at it
    double reaction[BBPD_MAX_BODIES];
    double reaction_static;
    double peak;        // largest static reaction at an increment end
    int finished;       // the stop criterion was met (the region keeps stepping until retired)
};

struct Body {
    const int *fam;     // member slots (absolute)
    int nFam;
    double cx, cz, R;
    int moves;          // follows the region's load schedule
    int region, index;  // owner and index within the region (body 0 gives the reaction)
};

struct __align__(32) NodeRec {
    double x, y, z; // position: deformed coordinates (FP64 form) or displacement (mixed form)
    uint64_t q;     // packed lattice coordinates kx | ky << 18 | kz << 36, bond count << 56
};

// stencil layout, per offset (one 32-byte scalar load)
struct __align__(32) StOff {
    union {
        struct {
            float x, y, z, L0; // mixed form: xi = h * d rounded to FP32, and |xi|
            float vr;          // mixed form: vcf / |xi|, vcf the volume correction factor of |h * d|
            float rq;          // mixed form: 1 / (2 |xi|^2)
        };
        struct {
            double Ln, hn, rn; // FP64 form: a nominal length L of the offset, RN(0.5 / L), RN(1 / L)
        };
    };
    int dx;            // the lattice offset d: dx, and dy | dz << 16 (16-bit fields)
    int dyz;
    __host__ __device__ int dy() const { return (int)(int16_t)(dyz & 0xFFFF); }
    __host__ __device__ int dz() const { return dyz >> 16; }
};
constexpr int ST_PAD = 2; // per-offset entries past the last (copies of it): loads run two offsets ahead

struct Params {
    // nodes
    int nSlots;
    const double *X0, *X1, *X2;
    NodeRec *pos[2];
    const NodeRec *Xrec; // undeformed coordinates as records (off-lattice inputs)
    double hx, hy, hz;   // lattice spacing per axis (lattice inputs)
    double *u0, *u1, *u2; // displacement (FP64 form only)
    double *v0, *v1, *v2;
    double *F0, *F1, *F2;
    double *Fp0, *Fp1, *Fp2; // ADR: the second force set (see adr_force)
    const double *rho, *mass;
    const unsigned char *cons; // bit d set = DOF d constrained; CONTACT_BIT = contact-family member
    // entries
    const int *rowlen;
    const long long *slice_off;
    uint4 *rec;
    const double *cstr;          // FP64: BONDSTIFFNESS per entry when not tabled
    const void *s0str, *scstr;   // per-entry s0 / sc when streamed (FP64 double, mixed float)
    const double *tab;           // FP64 tables: c values, then parameter values
    int ntabc, ntabp;
    int cmode, s0mode, scmode;
    int adr;
    int law_linear; // the brittle latch: the stencil layout's sign bit is the failed flag
    // scalars shared by every region
    double V, horizon, hmr, hpr, two_r; // hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2*RADIJ
    double rtwo_r;                      // RN(1 / two_r)
    double k, alpha, denom_exp, sc_steel, ff;
    double DT, D;
    double *partials;     // per global block, 4 doubles: the ADR Rayleigh partial sums cn1, cn2, the
                          // squared out-of-balance force over the free DOFs, 0
    double *hist;         // history rows (N_HIST doubles each), regions own ranges of rows
    const int *active;    // active regions (one k_adr_cn block each)
    // regions
    const Region *reg;
    RegionState *rst;
    const int *blkreg;    // global block -> region
    const int *grid;      // launch block (after the XCD remap) -> global block, -1 = idle
    const Body *bodies;
    const int2 *se;       // step-end blocks: (region, body index or -1)
    int nse;
    long long *tick;      // launches completed: the position-buffer parity
    unsigned int *tick_done; // step-end blocks finished this launch
    // stencil layout (full lattice box, rows in lattice-offset order): neighbours are implicit
    int stencil;
    const int4 *soff;     // per offset: slot delta, dx, dy, dz
    const StOff *sst;     // per offset, ST_PAD entries past each region's last
    const int2 *srow;     // per stencil row (offsets sharing dy, dz): offset count, slot delta at dx = 0;
                          // ST_PAD entries past each region's last
    const uint4 *smask;   // per slot: bit o set when the node has the bond at offset o
    const unsigned long long *smaskT; // per (slice, offset) in entry order: bit lane set when the slot
                                      // (slice, lane) has the bond; ST_PAD entries past the last
    void *hs;             // per entry (slice, offset, lane): bondSofteningFactor, sign bit = flag
    unsigned long long *snz; // per slice, two words: bit o set when an entry of offset o may hold a
                             // history word other than +0 (the others all hold +0)
    const float *tabf;    // mixed-form key tables (c * V, then parameter tables)
};

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

// e^x: in FP64 the library exponential; in FP32 (the mixed form) 2^(x log2 e) by the hardware
// exponential.  For the law's arguments (x in [-k, 0]) its absolute error stays at the FP32 level.
template <typename T>
__device__ __forceinline__ T lexp(T x)
{
    if constexpr (std::is_same<T, float>::value)
        return __builtin_amdgcn_exp2f(x * 1.44269504088896341f);
    else
        return exp(x);
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
    NodeRec nb;          // neighbour record
    double b0, b1, b2;   // neighbour undeformed coordinates (off-lattice inputs)
};

template <bool MIXED, int LAW, bool LAT>
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
            const double ex = Xn0 - Xm0, ey = Xn1 - Xm1, ez = Xn2 - Xm2;
            const double L0 = sqrt((ex * ex + ey * ey) + ez * ez);
            const double s = (l - L0) / L0;
            apply_law<LAW, double>(s, s0, sc, rs1, P.k, P.alpha, P.denom_exp, P.sc_steel, P.ff, H, w);
            // calculatevolumecorrectionfactors.m:20-32
            double vcf;
            if (L0 <= P.hmr) vcf = 1;
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

// Mixed form, stencil layout: the stretch by its Taylor series in x = (|y|^2 - |xi|^2) / (2 |xi|^2)
// for |x| at most this, else by a reciprocal square root and a reciprocal.  Both round to within a
// few ulps of the exact stretch.
constexpr float MIX_SERIES_X = 0x1p-9f;

// Launch variants, picked per solver by an autotuner: OCC is the minimum waves per SIMD the compiler
// must allow (1 = its own choice), which bounds the registers per wave, trading instruction-level
// for wave-level parallelism; REMAP keeps consecutive blocks on one XCD (its L2) or not.  PV says
// where s0 and sc come from: PV_SCALAR the region's scalars, PV_TABLE both from key tables,
// PV_GENERAL each from its scalar, key table or per-entry stream as the solver's modes say.
template <bool MIXED, int LAW, int OCC, bool REMAP, int PV>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(OCC))) void k_bond_st(Params P)
{
    using T = typename std::conditional<MIXED, float, double>::type;
    constexpr uint32_t TW = sizeof(T);
    // per wavefront, a row buffer: x, y, z and the key word (low word) of ST_ROW sites.  One buffer
    // suffices: a wavefront's LDS accesses complete in program order, so staging the next row
    // (after the last read of this one) cannot overtake a read.
    __shared__ double s_row[BLOCK / SLICE][4 * ST_ROW];
    extern __shared__ double tabs[]; // key tables (FP64, or FP32 in the mixed form)
    const int gb = launch_block(P, REMAP);
    if (gb < 0) return;
    const int n = gb * BLOCK + threadIdx.x;
    const Region &RG = P.reg[P.blkreg[gb]];
    const int base = RG.base, noff = RG.noff, off0 = RG.off0, kg = RG.kg, tab0 = RG.tab0, ntab = RG.ntab;
    const int ts0 = RG.ts0, tsc = RG.tsc;
    const bool fast64 = !MIXED && RG.fast64;
    const bool real = n - base < RG.nN;
    const T rs0 = (T)RG.s0, rs1 = (T)RG.s1, rsc = (T)RG.sc, rthr = law_threshold<LAW, T>(rs0, rsc, rs1);
    const T rs0lo = (T)RG.s0lo, rsclo = (T)RG.sclo;
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
            tdst[2 * i + 1] = law_threshold<LAW, T>(tsrc[tab0 + ts0 + i], tsrc[tab0 + tsc + i], (T)RG.s1);
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
    constexpr int KSCALE = PV == PV_GENERAL ? 1 : (PV == PV_TABLE ? 2 : 1) * (int)sizeof(T);
    const int tabl = PV == PV_GENERAL ? 0 : (int)lds_addr(tabs);

    const NodeRec me = pos[n];
    const double qn0 = me.x, qn1 = me.y, qn2 = me.z;
    const int kx = (int)(me.q & LAT_MASK), ky = (int)((me.q >> LAT_BITS) & LAT_MASK),
              kz = (int)((me.q >> (2 * LAT_BITS)) & LAT_MASK);
    const int gme = (int)(me.q >> G_SHIFT); // index of the node's family size
    const int krow = gme * kg * KSCALE + tabl; // the node's row of the key table (index or LDS address)
    // X == h * k exactly (checked at set-up)
    const double Xn0 = P.hx * (double)kx, Xn1 = P.hy * (double)ky, Xn2 = P.hz * (double)kz;
    // this wavefront's slice of the entry arrays: entry (o, lane) at byte (o * SLICE + lane) * TW
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / SLICE));
    const long long eslice = RG.e0 + ((long long)((gb * BLOCK - base) / SLICE) + wv) * noff * SLICE;
    const uint32_t sbytes = (uint32_t)noff * SLICE * TW;
    const brsrc rh = make_rsrc((const T *)P.hs + eslice, sbytes);
    const brsrc r0 = PV == PV_GENERAL && P.s0mode == PM_STREAM ? make_rsrc((const T *)P.s0str + eslice, sbytes) : rh;
    const brsrc r1 = PV == PV_GENERAL && P.scmode == PM_STREAM ? make_rsrc((const T *)P.scstr + eslice, sbytes) : rh;
    const uint32_t lb = (threadIdx.x % SLICE) * TW;
    // Neighbour records, staged per stencil row (the offsets that share dy and dz, consecutive in
    // offset order).  For the slice starting at slot n0, row r covers the slots
    // n0 + delta_r - rx .. n0 + delta_r + 63 + rx; the wavefront loads them once into its LDS row
    // buffer, and the offset (dx, dy, dz) of lane i reads site i + rx + dx there.  The next row's
    // loads are in flight while a row is computed.  Slot offsets are 32-bit: a site before slot 0
    // wraps to an offset past every valid one (set-up keeps slots below 2^26), and a read outside
    // the slot array returns zeros (such a site is never a bond).
    const brsrc rp = make_rsrc(pos, (uint32_t)P.nSlots * (uint32_t)sizeof(NodeRec));
    const int lane = (int)(threadIdx.x % SLICE);
    const int n0 = gb * BLOCK + wv * SLICE;
    const int rx = RG.rx, nrow = RG.nrow;
    const int *rows = reinterpret_cast<const int *>(P.srow + RG.row0); // (count, delta) per row
    auto fetch = [&](RowRecs &R, int dl) {
        const uint32_t sb = (uint32_t)(n0 + dl - rx);
        R.a = bload_rec(rp, (sb + (uint32_t)lane) * (uint32_t)sizeof(NodeRec));
        if (lane < 2 * rx) R.b = bload_rec(rp, (sb + SLICE + (uint32_t)lane) * (uint32_t)sizeof(NodeRec));
    };
    auto put = [&](double *w, int site, const NodeRec &q) {
        w[site] = q.x;
        w[ST_ROW + site] = q.y;
        w[2 * ST_ROW + site] = q.z;
        *reinterpret_cast<int *>(w + 3 * ST_ROW + site) = (int)(q.q >> G_SHIFT) * KSCALE;
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
    double f0 = 0, f1 = 0, f2 = 0;
    const StOff *sst = P.sst + RG.sst0;
    const unsigned long long *mks = P.smaskT + eslice / SLICE;
    // the slice's history words (noff <= 128): an offset whose bit is clear holds +0 in every entry,
    // which is taken without a load.  This kernel sets bits of them (atomicOr below), so they are read
    // once, here, with an ordinary global load (every lane the same address) made wavefront-uniform
    // with readfirstlane, not through the constant address space; the bits set later are for offsets
    // this wavefront has already passed.
    unsigned long long *hzp = P.snz + 2 * (n0 / SLICE);
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
        if (!MIXED && PV != PV_GENERAL && d.st.dyz != cyz) {
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
        const double *nb = &s_row[wv][0] + (lane + rx + d.st.dx);
        const double nbx = nb[0], nby = nb[ST_ROW], nbz = nb[2 * ST_ROW];
        const int key = krow + *reinterpret_cast<const int *>(nb + 3 * ST_ROW);
        // the entry's key-table values: c (c * V in the mixed form), s0, sc
        const T *tk = PV == PV_GENERAL ? tab + key : lds_ptr<T>((uint32_t)key);
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
            s0 = P.s0mode == PM_TABLE ? tab[ts0 + key] : (P.s0mode == PM_STREAM ? rs0lo : rs0);
            sc = P.scmode == PM_TABLE ? tab[tsc + key] : (P.scmode == PM_STREAM ? rsclo : rsc);
            thr = law_threshold<LAW, T>(s0, sc, rs1);
        }
        // the law, and the store of a changed history (sign = flag)
        auto law = [&](T s) {
            if (!law_active<LAW, T>(s, thr, flag)) return;
            if (PV == PV_TABLE) {
                const uint32_t ki = (uint32_t)(key - tabl) / KSCALE;
                s0 = bload<T>(rtab, (uint32_t)(ts0 + ki) * TW);
                sc = bload<T>(rtab, (uint32_t)(tsc + ki) * TW);
            }
            if (PV == PV_GENERAL && P.s0mode == PM_STREAM) s0 = bload<T>(r0, lb + so);
            if (PV == PV_GENERAL && P.scmode == PM_STREAM) sc = bload<T>(r1, lb + so);
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
            if (PV == PV_GENERAL) {
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
            if (gx != gx) gx = 0.0;
            if (gy != gy) gy = 0.0;
            if (gz != gz) gz = 0.0;
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
            // |x| <= 2^-9: Taylor series, s = x - x^2/2 + x^3/2 (the next term, -5 x^4 / 8, is below
            // 2^-27 |s|) and g = s - s^2 + s^3 (next s^4, below 2^-26 |g|).  The choice is the
            // bond's own (x is the same from either end), so both ends store the same history.
            float s = x * fmaf(x, fmaf(x, 0.5f, -0.5f), 1.0f);
            float g = fmaf(s, fmaf(s, s, -s), s);
            if (!(fabsf(x) <= MIX_SERIES_X)) {
                // |y|^2 carries 1e-30 (below every representable bond's square by many orders, so it
                // changes no realistic length): 1/l stays finite for a bond of zero length, whose
                // force p * y is then 0, the reference's NaN -> 0 (calculatebondforcesMex.c:38)
                const float l2 = fmaf(y0, y0, fmaf(y1, y1, fmaf(y2, y2, 1e-30f)));
                const float rl = __builtin_amdgcn_rsqf(l2);
                const float l = l2 * rl;
                s = num * __builtin_amdgcn_rcpf(d.st.L0 * (l + d.st.L0));
                g = (s * rl) * d.st.L0;
            }
            law(s);
            // kap (1 - H) s / l = (c vcf / |xi|) (1 - H) g; a failed bond of the brittle law carries
            // no force
            const float kap = c * d.st.vr;
            const float p = (LAW == BBPD_LAW_LINEAR && flag) ? 0.0f : fmaf(-kap, fabsf(hv), kap) * g;
            a.x = fmaf(p, y0, a.x);
            a.y = fmaf(p, y1, a.y);
            a.z = fmaf(p, y2, a.z);
        }
    };
    if (noff > 0) {
        RowRecs R;
        fetch(R, sload(rows, 1));
        stage(R);
        int r = 0, rend = sload(rows, 0);
        if (nrow > 1) fetch(R, sload(rows, 3));
        // the next row's offset count and the slot delta of the row after it, read a row ahead
        int cnext = sload(rows, 2), dnext = sload(rows, 5);
        // entering offset o: at a row's end, stage the next row and start loading the one after
        auto next_row = [&](int o) {
            if (o == rend) {
                r++;
                stage(R);
                rend += cnext;
                if (r + 1 < nrow) fetch(R, dnext);
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
    const int2 job = P.se[blockIdx.x];
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
    // the static reaction in `rs`
    auto body = [&](const int bi) {
        const Body &B = P.bodies[RG.body0 + bi];
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
        body(bi);
    } else if (bi == -2) {
        // overlapping families: this block applies the bodies in order, as the reference's calls
        // follow one another (dynamicsolverdisplacementcontrolled.m:108-112); a member shared with an
        // earlier body is seen as that body left it (the barrier orders the block's global writes)
        for (int b = 0; b < RG.nBodies; b++) {
            total = 0;
            body(b);
            if (threadIdx.x == 0) RS.reaction[b] = total;
            if (b == 0 && threadIdx.x == 0) RS.reaction_static = rs;
            __syncthreads();
        }
    }
    if (threadIdx.x == 0) {
        if (bi >= 0) RS.reaction[bi] = total;
        if (bi == 0) RS.reaction_static = rs;
        __threadfence();
        const unsigned need = (unsigned)(RG.nBodies > 0 && !RG.seq_contact ? RG.nBodies : 1);
        const unsigned ticket = atomicAdd(&RS.done, 1u);
        last_block = (ticket == need - 1);
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
    if (threadIdx.x == 0) {
        __threadfence();
        const unsigned ticket = atomicAdd(P.tick_done, 1u);
        last_all = (ticket == (unsigned)gridDim.x - 1);
    }
    __syncthreads();
    if (last_all && threadIdx.x == 0) {
        __threadfence();
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
        auto bond = [&](double H, bool failed, int m, double c, double s0, double sc, double Xm0, double Xm1,
                        double Xm2) {
            const bool broken = failed || H == 1.0;
            cnt++;
            unbroken += broken ? 0 : 1;
            if (sed) {
                // buildhorizons.m:93-97, calculatedeformedlengthMex.c:40-43 and
                // calculatevolumecorrectionfactors.m:20-32, then the micropotential of
                // calculatestrainenergy.m:16 (BFMULTIPLIER = 1 for concrete, buildbonddata.m:56-57)
                const double ex = Xn[0] - Xm0, ey = Xn[1] - Xm1, ez = Xn[2] - Xm2;
                const double L0 = sqrt((ex * ex + ey * ey) + ez * ez);
                const NodeRec pm = pos[m];
                const double dx = (MIXED ? Xm0 + pm.x : pm.x) - xn0, dy = (MIXED ? Xm1 + pm.y : pm.y) - xn1,
                             dz = (MIXED ? Xm2 + pm.z : pm.z) - xn2;
                const double l = sqrt(dx * dx + dy * dy + dz * dz);
                const double st = (l - L0) / L0;
                double vcf;
                if (L0 <= P.hmr) vcf = 1;
                else if ((L0 > P.hmr) && (L0 <= P.horizon)) vcf = (P.hpr - L0) / P.two_r;
                else vcf = 0;
                const double fl = (LAW == BBPD_LAW_LINEAR && failed) ? 0.0 : 1.0;
                mu_acc = mu_acc + 0.5 * fl * 1.0 * c * (st * st) * L0 * P.V * vcf;
            }
            if (m > n && LAW != BBPD_LAW_LINEAR) { // each bond once, at its lower slot
                const double ex = Xn[0] - Xm0, ey = Xn[1] - Xm1, ez = Xn[2] - Xm2;
                const double L0 = sqrt((ex * ex + ey * ey) + ez * ez);
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
                bond(H, (w & FAILED_BIT) != 0, m, c, s0, sc, Xm[0], Xm[1], Xm[2]);
            }
        }
        if (dmg) dmg[n - RG.base] = cnt > 0 ? 1 - ((double)unbroken / (double)cnt) : 0.0;
        if (sed) sed[n - RG.base] = 0.5 * mu_acc;
    }
    const double sum = bbpd::block_sum_fixed<BLOCK>(e_acc, sh);
    if (threadIdx.x == 0) epart[blockIdx.x] = sum;
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

double bits_to_double(uint32_t lo, uint32_t hi)
{
    uint64_t v = ((uint64_t)hi << 32) | lo;
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

void double_to_bits(double d, uint32_t &lo, uint32_t &hi)
{
    uint64_t v;
    std::memcpy(&v, &d, 8);
    lo = (uint32_t)v;
    hi = (uint32_t)(v >> 32);
}

} // namespace

struct bbpd_solver {
    int device = 0;
    hipStream_t stream = nullptr;
    int precision = 0, law = 0, integrator = 0;
    bool lattice = false;
    bool stencil = false;   // stencil layout (implicit neighbours) rather than entry records
    std::string layout;     // which layout, and why not the stencil one
    int nN = 0, nB = 0, nSlots = 0, nSlices = 0;
    long long nEntries = 0, nPadded = 0;
    std::vector<int> slot_of;                // original node (0-based) -> slot
    std::vector<long long> entry_i, entry_j; // bond -> entry index at its two ends
    std::vector<void *> owned;
    Params P{};
    size_t lds_bytes = 0;
    unsigned long long *counts = nullptr;
    size_t device_bytes = 0;
    // regions (a single problem is region 0)
    std::vector<Region> reg;            // host copies of the descriptors
    std::vector<char> active;           // region in use
    std::vector<long long> hist_read;   // per region: steps already returned by the history call
    Region *d_reg = nullptr;
    RegionState *d_rst = nullptr;
    int *d_blkreg = nullptr, *d_grid = nullptr, *d_active = nullptr;
    int *d_grid_st = nullptr; // the launch list in XCD stripes (launch variants 4 and 5)
    int grid_n_st = 0;
    int2 *d_se = nullptr;
    Body *d_bodies = nullptr;
    long long *d_tick = nullptr;
    unsigned *d_tick_done = nullptr;
    double *d_hist = nullptr;
    int grid_n = 0;   // launch blocks of the per-slot kernels (a multiple of 8)
    int nse = 0;      // step-end blocks
    int nactive = 0;  // active regions
    int version = 0;  // bumped when the set of active regions changes (captured graphs go stale)
    // batches (bbpd_batch_*): regions admitted and retired at run time, stencil layout
    bool batch = false;
    int maxRegions = 0, histRows = 0;
    std::vector<char> freeblk;          // per global block: 1 = free
    struct RegionHost {
        int nN = 0, nB = 0;
        std::vector<int> slot_of;             // node -> slot
        std::vector<long long> entry_i, entry_j; // bond -> entry at its two ends
    };
    std::vector<RegionHost> rh;
    bbpd_problem shape{};               // scalars every admitted problem shares
    double DX = 0;
    int *d_fampool = nullptr;
    // timing
    bool timing = false;
    bool tuned = false; // every launch variant has run once (its code is loaded)
    std::vector<hipEvent_t> ev;
    int ev_used = 0;
    double t_acc_ms = 0;
    long long t_launches = 0;
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    int graph_steps = 0;
    int graph_version = -1;
    int occ = 0; // launch variant of the stencil bond kernel (pick_occ)
};

namespace {

using KernelFn = void (*)(Params);

template <bool MIXED, int LAW>
KernelFn pick_lat(bool lat)
{
    return lat ? k_bond<MIXED, LAW, true> : k_bond<MIXED, LAW, false>;
}

template <bool MIXED>
KernelFn pick_law(int law, bool lat)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return pick_lat<MIXED, BBPD_LAW_LINEAR>(lat);
    case BBPD_LAW_BILINEAR: return pick_lat<MIXED, BBPD_LAW_BILINEAR>(lat);
    case BBPD_LAW_TRILINEAR: return pick_lat<MIXED, BBPD_LAW_TRILINEAR>(lat);
    default: return pick_lat<MIXED, BBPD_LAW_EXPONENTIAL>(lat);
    }
}

// launch variants of the stencil bond kernel: (minimum waves per SIMD, block order); orders: XCD
// remap, launch order, and launch order over the stripe list (variants 4 and 5)
constexpr int N_OCC = BBPD_AUTOTUNE_VARIANTS;
constexpr int ST_STRIPE = 8; // consecutive global blocks per XCD in the stripe list

template <bool MIXED, int LAW, int PV>
KernelFn pick_occ(int v)
{
    switch (v) {
    case 1: return k_bond_st<MIXED, LAW, 1, false, PV>;
    case 2: return k_bond_st<MIXED, LAW, 5, true, PV>;
    case 3: return k_bond_st<MIXED, LAW, 5, false, PV>;
    case 4: return k_bond_st<MIXED, LAW, 1, false, PV>;
    case 5: return k_bond_st<MIXED, LAW, 5, false, PV>;
    default: return k_bond_st<MIXED, LAW, 1, true, PV>;
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

template <bool MIXED, int PV>
KernelFn pick_law_st(int law, int v)
{
    switch (law) {
    case BBPD_LAW_LINEAR: return pick_occ<MIXED, BBPD_LAW_LINEAR, PV>(v);
    case BBPD_LAW_BILINEAR: return pick_occ<MIXED, BBPD_LAW_BILINEAR, PV>(v);
    case BBPD_LAW_TRILINEAR: return pick_occ<MIXED, BBPD_LAW_TRILINEAR, PV>(v);
    default: return pick_occ<MIXED, BBPD_LAW_EXPONENTIAL, PV>(v);
    }
}

template <bool MIXED>
KernelFn pick_pv_st(int pv, int law, int v)
{
    switch (pv) {
    case PV_SCALAR: return pick_law_st<MIXED, PV_SCALAR>(law, v);
    case PV_TABLE: return pick_law_st<MIXED, PV_TABLE>(law, v);
    default: return pick_law_st<MIXED, PV_GENERAL>(law, v);
    }
}

KernelFn pick_bond_kernel(const bbpd_solver *s, int variant = -1)
{
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int v = variant >= 0 ? variant : s->occ;
    if (s->stencil) {
        const int m0 = s->P.s0mode, mc = s->P.scmode;
        const int pv = (m0 == PM_SCALAR && mc == PM_SCALAR) ? PV_SCALAR
                       : (m0 == PM_TABLE && mc == PM_TABLE) ? PV_TABLE : PV_GENERAL;
        return mixed ? pick_pv_st<true>(pv, s->law, v) : pick_pv_st<false>(pv, s->law, v);
    }
    return mixed ? pick_law<true>(s->law, s->lattice) : pick_law<false>(s->law, s->lattice);
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

bool bits_equal(double a, double b)
{
    uint64_t x, y;
    std::memcpy(&x, &a, 8);
    std::memcpy(&y, &b, 8);
    return x == y;
}

// distinct values of a per-bond array, as a table and per-bond indices; false if more than cap
bool make_table(const double *a, int n, int cap, std::vector<double> &tab, std::vector<uint32_t> &idx)
{
    std::unordered_map<uint64_t, uint32_t> m;
    tab.clear();
    idx.assign(n, 0);
    for (int k = 0; k < n; k++) {
        uint64_t key;
        std::memcpy(&key, &a[k], 8);
        auto it = m.find(key);
        if (it == m.end()) {
            if ((int)tab.size() >= cap) return false;
            it = m.emplace(key, (uint32_t)tab.size()).first;
            tab.push_back(a[k]);
        }
        idx[k] = it->second;
    }
    return true;
}

} // namespace

namespace {

// Checks of a problem's bond arrays: ids, UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS against their
// formulas (both are recomputed on the device: buildhorizons.m:93-97,
// calculatevolumecorrectionfactors.m:20-32), fail and BONDTYPE values.  Fills the 0-based ends.
int check_bonds(const bbpd_problem *pr, std::vector<int> &bi, std::vector<int> &bj)
{
    const int nN = pr->nNodes, nB = pr->nBonds;
    const double *X = pr->undeformedCoordinates, *BL = pr->BONDLIST;
    const double horizon = pr->horizon, RADIJ = pr->RADIJ;
    const double hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2 * RADIJ;
    bi.assign(nB, 0);
    bj.assign(nB, 0);
    for (int k = 0; k < nB; k++) {
        double di = BL[k], dj = BL[(size_t)nB + k];
        int i = (int)di, j = (int)dj;
        if (di != (double)i || dj != (double)j || i < 1 || j < 1 || i > nN || j > nN || i == j) {
            bbpd::set_error("BONDLIST row " + std::to_string(k + 1) + " is not a pair of distinct ids in 1..nNodes");
            return BBPD_ERR_ARG;
        }
        bi[k] = i - 1;
        bj[k] = j - 1;
        double ex = X[i - 1] - X[j - 1], ey = X[i - 1 + nN] - X[j - 1 + nN],
               ez = X[i - 1 + 2 * (size_t)nN] - X[j - 1 + 2 * (size_t)nN];
        double L0 = std::sqrt((ex * ex + ey * ey) + ez * ez);
        if (!bits_equal(L0, pr->UNDEFORMEDLENGTH[k])) {
            bbpd::set_error("UNDEFORMEDLENGTH(" + std::to_string(k + 1) +
                            ") differs from the length recomputed from undeformedCoordinates (buildhorizons.m:93-97)");
            return BBPD_ERR_UNSUPPORTED;
        }
        double vcf;
        if (L0 <= hmr) vcf = 1;
        else if ((L0 > hmr) && (L0 <= horizon)) vcf = (hpr - L0) / two_r;
        else vcf = 0;
        if (!bits_equal(vcf, pr->VOLUMECORRECTIONFACTORS[k])) {
            bbpd::set_error("VOLUMECORRECTIONFACTORS(" + std::to_string(k + 1) +
                            ") differs from calculatevolumecorrectionfactors(L0, horizon, RADIJ)");
            return BBPD_ERR_UNSUPPORTED;
        }
        if (pr->fail && !(pr->fail[k] == 0 || pr->fail[k] == 1)) {
            bbpd::set_error("fail must hold 0 or 1");
            return BBPD_ERR_ARG;
        }
        if (pr->BONDTYPE && !(pr->BONDTYPE[k] == 0 || pr->BONDTYPE[k] == 1 || pr->BONDTYPE[k] == 2)) {
            bbpd::set_error("BONDTYPE must hold 0, 1 or 2");
            return BBPD_ERR_ARG;
        }
    }
    return BBPD_OK;
}

// Lattice detection: X == h * k exactly with small non-negative integers k, per axis.
bool detect_lattice(const bbpd_problem *pr, double hlat[3], double xmin[3], double xmax[3], double &hq)
{
    const int nN = pr->nNodes, nB = pr->nBonds;
    const double *X = pr->undeformedCoordinates;
    for (int d = 0; d < 3; d++) {
        xmin[d] = 1e300;
        xmax[d] = -1e300;
        for (int n = 0; n < nN; n++) {
            xmin[d] = std::min(xmin[d], X[n + (size_t)nN * d]);
            xmax[d] = std::max(xmax[d], X[n + (size_t)nN * d]);
        }
    }
    hq = 1e300;
    for (int k = 0; k < nB; k++) hq = std::min(hq, pr->UNDEFORMEDLENGTH[k]);
    if (!(hq > 0) || hq > 1e299) hq = 1;
    bool lattice = true;
    for (int d = 0; d < 3; d++) hlat[d] = 0;
    for (int d = 0; d < 3 && lattice; d++) {
        // the smallest coordinate is h * k0 with k0 = round(xmin / shortest bond)
        double k0 = std::floor(xmin[d] / hq + 0.5);
        double h = (k0 >= 1) ? xmin[d] / k0 : hq;
        hlat[d] = h;
        for (int n = 0; n < nN && lattice; n++) {
            double xv = X[n + (size_t)nN * d];
            double kk = std::floor(xv / h + 0.5);
            if (!(kk >= 0 && kk < (double)(1 << LAT_BITS)) || !bits_equal(h * (double)(int)kk, xv)) lattice = false;
        }
    }
    return lattice;
}

// Stencil layout plan.  Applies to a lattice input (X == h * k) that fills a box, whose bonds all
// have BONDTYPE 0 (fail 1 unless the law is the brittle latch), whose rows list their bonds in
// ascending lattice-offset order (z, then y, then x: what a BONDLIST sorted by node id on an
// x-fastest lattice gives), with at most ST_MAX distinct offsets, and whose BONDSTIFFNESS is a
// function of the pair (nFAMILY_i, nFAMILY_j).  Slots are then the x-fastest lattice order.
struct StencilPlan {
    bool ok = false;
    std::string why;              // reason the layout does not apply
    int noff = 0;                 // offsets, padded to even
    std::vector<int4> off;        // slot delta, dx, dy, dz
    std::vector<int2> rows;       // offset count and slot delta of dx = 0, per run of equal (dy, dz);
                                  // nrow rows, then ST_PAD padding entries
    int nrow = 0;
    int rx = 0;                   // largest |dx|
    std::vector<int> slot;        // node -> slot
    std::vector<int> oi, oj;      // per bond: offset index at the i end and at the j end
    std::vector<int> g;           // per node: bond count
    std::vector<int> gidx;        // per node: index of its bond count among the distinct counts
    int G = 0;                    // distinct bond counts (key = gidx_n * G + gidx_m)
    std::vector<int> k;           // per node: lattice coordinates (absolute), 3 per node
    int nkey = 0;
    std::vector<double> tc, ts0, tsc; // key tables (empty when not tabled)
    int s0mode = PM_SCALAR, scmode = PM_SCALAR;
    int kmin[3] = {0, 0, 0}, kmax[3] = {0, 0, 0}; // lattice coordinate ranges
};

// Lane masks of the stencil layout: per slice of 64 slots and per offset, bit lane set when the slot
// (slice, lane) has the bond at that offset; from the per-slot masks `m` of `nslots` slots.
std::vector<unsigned long long> lane_masks(const std::vector<uint4> &m, long long nslots, int noff)
{
    std::vector<unsigned long long> r((size_t)(nslots / SLICE) * noff, 0ull);
    for (long long q = 0; q < nslots; q++) {
        const uint32_t *w = reinterpret_cast<const uint32_t *>(&m[q]);
        for (int o = 0; o < noff; o++)
            if ((w[o >> 5] >> (o & 31)) & 1u) r[(size_t)(q / SLICE) * noff + o] |= 1ull << (q % SLICE);
    }
    return r;
}

// History words of the stencil layout: per slice of 64 slots, two words whose bit o is set when an
// entry of offset o holds a history word other than +0 (a set flag included); from the history `h`
// of `nsl` slices in entry order (slice, offset, lane).  The bond kernel takes +0 for an offset whose
// bit is clear without loading it, and sets the bit when it stores a history word there.
template <typename V>
std::vector<unsigned long long> history_words(const V *h, long long nsl, int noff)
{
    std::vector<unsigned long long> w((size_t)nsl * 2, 0ull);
    for (long long q = 0; q < nsl; q++)
        for (int o = 0; o < noff; o++) {
            const V *e = h + (q * noff + o) * SLICE;
            bool nz = false;
            for (int l = 0; l < SLICE; l++) nz |= e[l] != 0 || std::signbit(e[l]);
            if (nz) w[(size_t)q * 2 + (o >> 6)] |= 1ull << (o & 63);
        }
    return w;
}

// Per-offset records of the stencil layout, then ST_PAD copies of the last.
// Mixed form: xi = h * d rounded to FP32 per component (an offset and its negation give exactly
// negated vectors), |xi| and |xi|^2 from those, and the volume correction factor of |h * d|
// (calculatevolumecorrectionfactors.m:20-32).
// FP64 form: a nominal length L of the offset with RN(0.5 / L) and RN(1 / L).  The bond kernel
// rebuilds X = h * k per entry, so the undeformed length sqrt(q) (buildhorizons.m:93-97) of one
// offset takes values up to some hundred ulps apart over the lattice; it takes sqrt(q) from the
// nominal values by one correction step, RN(1 / L0) by two Newton steps, and vcf by Markstein's
// correction with RN(1 / 2r).
// `*fast64` says whether those steps give sqrt(q), RN(1 / L0) and the quotient (hpr - L0) / 2r bit
// for bit for every value q takes: every combination of the per-axis differences
// h k - h (k + d) over the coordinate ranges kmin .. kmax.
std::vector<StOff> offset_records(const std::vector<int4> &off, const double *h, double hmr, double hpr,
                                  double horizon, double two_r, bool mixed, const int *kmin, const int *kmax,
                                  int *fast64)
{
    std::vector<StOff> r(off.size());
    // per axis and offset component: the values h k - h (k + d) takes
    auto diffs = [&](int ax, int dd) {
        std::vector<double> v;
        for (int k = kmin[ax]; k <= kmax[ax]; k++)
            if (k + dd >= kmin[ax] && k + dd <= kmax[ax]) v.push_back(h[ax] * (double)k - h[ax] * (double)(k + dd));
        std::sort(v.begin(), v.end());
        v.erase(std::unique(v.begin(), v.end()), v.end());
        return v;
    };
    const double rtwo_r = 1.0 / two_r;
    bool ok = !mixed;
    for (size_t o = 0; o < off.size(); o++) {
        StOff &R = r[o];
        R.dx = off[o].y;
        R.dyz = (off[o].z & 0xFFFF) | (int)((unsigned)off[o].w << 16);
        const double e0 = h[0] * off[o].y, e1 = h[1] * off[o].z, e2 = h[2] * off[o].w;
        if (mixed) {
            const float x0 = (float)e0, x1 = (float)e1, x2 = (float)e2;
            const double L0 = std::sqrt((e0 * e0 + e1 * e1) + e2 * e2);
            double vcf;
            if (L0 <= hmr) vcf = 1;
            else if ((L0 > hmr) && (L0 <= horizon)) vcf = (hpr - L0) / two_r;
            else vcf = 0;
            // |xi|^2 of the FP32 vector, in FP64
            const double q = ((double)x0 * x0 + (double)x1 * x1) + (double)x2 * x2;
            const float L0f = (float)std::sqrt(q);
            R.x = x0, R.y = x1, R.z = x2, R.L0 = L0f;
            R.vr = q > 0 ? (float)(vcf / std::sqrt(q)) : 0.0f;
            R.rq = q > 0 ? (float)(0.5 / q) : 0.0f;
            continue;
        }
        const std::vector<double> ex = diffs(0, off[o].y), ey = diffs(1, off[o].z), ez = diffs(2, off[o].w);
        if (ex.empty() || ey.empty() || ez.empty()) {
            // no bond has this offset on this lattice: any finite values
            R.Ln = 1.0, R.hn = 0.5, R.rn = 1.0;
            continue;
        }
        const double q = (ex[0] * ex[0] + ey[0] * ey[0]) + ez[0] * ez[0];
        R.Ln = std::sqrt(q);
        R.hn = 0.5 / R.Ln;
        R.rn = 1.0 / R.Ln;
        for (double a : ex)
            for (double b : ey)
                for (double c : ez) {
                    const double qq = (a * a + b * b) + c * c;
                    const double L0 = std::sqrt(qq);
                    const double g = std::fma(std::fma(-R.Ln, R.Ln, qq), R.hn, R.Ln);
                    const double r1 = std::fma(std::fma(-L0, R.rn, 1.0), R.rn, R.rn);
                    const double rr = std::fma(std::fma(-L0, r1, 1.0), r1, r1);
                    ok = ok && bits_equal(g, L0) && bits_equal(rr, 1.0 / L0);
                    if ((L0 > hmr) && (L0 <= horizon)) {
                        const double y = (hpr - L0) * rtwo_r;
                        ok = ok && bits_equal(std::fma(std::fma(-two_r, y, hpr - L0), rtwo_r, y), (hpr - L0) / two_r);
                    }
                }
    }
    for (int k = 0; k < ST_PAD && !r.empty(); k++) r.push_back(r.back());
    if (fast64) *fast64 = ok ? 1 : 0;
    return r;
}

bool key_table(const double *a, long long nB, const std::vector<int> &key, int nkey, std::vector<double> &tab)
{
    tab.assign(nkey, 0.0);
    std::vector<char> set(nkey, 0);
    for (long long b = 0; b < nB; b++) {
        const int q = key[b];
        if (!set[q]) {
            set[q] = 1;
            tab[q] = a[b];
        } else if (!bits_equal(tab[q], a[b])) {
            return false;
        }
    }
    return true;
}

void plan_stencil(const bbpd_problem *pr, const double *X, const std::vector<int> &bi, const std::vector<int> &bj,
                  const double *hlat, StencilPlan &sp)
{
    const int nN = pr->nNodes, nB = pr->nBonds;
    sp.ok = false;
    sp.k.assign(3 * (size_t)nN, 0);
    int *kmin = sp.kmin, *kmax = sp.kmax;
    for (int d = 0; d < 3; d++) kmin[d] = 1 << 30, kmax[d] = -1;
    for (int n = 0; n < nN; n++)
        for (int d = 0; d < 3; d++) {
            const int kk = (int)std::floor(X[n + (size_t)nN * d] / hlat[d] + 0.5);
            sp.k[3 * (size_t)n + d] = kk;
            kmin[d] = std::min(kmin[d], kk);
            kmax[d] = std::max(kmax[d], kk);
        }
    long long dims[3];
    for (int d = 0; d < 3; d++) dims[d] = kmax[d] - kmin[d] + 1;
    if (dims[0] * dims[1] * dims[2] != (long long)nN) {
        sp.why = "the nodes do not fill a lattice box";
        return;
    }
    if ((long long)nN + SLOT_ALIGN > (1LL << 26)) {
        sp.why = "more than 2^26 nodes (node-record offsets are 32-bit)";
        return;
    }
    sp.slot.assign(nN, 0);
    std::vector<char> seen(nN, 0);
    for (int n = 0; n < nN; n++) {
        const long long q = (sp.k[3 * (size_t)n] - kmin[0]) +
                            dims[0] * ((sp.k[3 * (size_t)n + 1] - kmin[1]) + dims[1] * (sp.k[3 * (size_t)n + 2] - kmin[2]));
        if (seen[q]) {
            sp.why = "two nodes share a lattice site";
            return;
        }
        seen[q] = 1;
        sp.slot[n] = (int)q;
    }
    for (int b = 0; b < nB; b++) {
        if (pr->BONDTYPE && pr->BONDTYPE[b] != 0) {
            sp.why = "BONDTYPE other than 0";
            return;
        }
        if (pr->law != BBPD_LAW_LINEAR && pr->fail && pr->fail[b] != 1) {
            sp.why = "fail other than 1 with a softening law";
            return;
        }
    }
    // offsets: the union over all bonds of k_j - k_i and its negation
    int R = 0;
    for (int b = 0; b < nB; b++)
        for (int d = 0; d < 3; d++) R = std::max(R, std::abs(sp.k[3 * (size_t)bj[b] + d] - sp.k[3 * (size_t)bi[b] + d]));
    if (R > ST_RMAX) {
        sp.why = "bond offsets longer than " + std::to_string(ST_RMAX) + " lattice spacings";
        return;
    }
    const int Wd = 2 * R + 1;
    auto cell = [&](int dx, int dy, int dz) { return (dx + R) + Wd * ((dy + R) + Wd * (dz + R)); };
    std::vector<int> idx((size_t)Wd * Wd * Wd, -1);
    for (int b = 0; b < nB; b++) {
        int dd[3];
        for (int d = 0; d < 3; d++) dd[d] = sp.k[3 * (size_t)bj[b] + d] - sp.k[3 * (size_t)bi[b] + d];
        idx[cell(dd[0], dd[1], dd[2])] = 0;
        idx[cell(-dd[0], -dd[1], -dd[2])] = 0;
    }
    // ascending z, then y, then x: ascending slot delta
    int no = 0;
    for (int dz = -R; dz <= R; dz++)
        for (int dy = -R; dy <= R; dy++)
            for (int dx = -R; dx <= R; dx++)
                if (idx[cell(dx, dy, dz)] == 0) {
                    idx[cell(dx, dy, dz)] = no++;
                    sp.off.push_back(make_int4((int)(dx + dims[0] * (dy + dims[1] * dz)), dx, dy, dz));
                }
    if (no > ST_MAX) {
        sp.why = "more than 128 distinct bond offsets";
        return;
    }
    // padding offset, never a bond: a copy of the last, so it extends the last row
    if (no & 1) sp.off.push_back(sp.off.back());
    sp.noff = (int)sp.off.size();
    // rows: runs of offsets with equal (dy, dz), with the slot delta of their dx = 0 site
    sp.rows.clear();
    sp.rx = 0;
    for (int o = 0; o < sp.noff;) {
        int e = o;
        while (e < sp.noff && sp.off[e].z == sp.off[o].z && sp.off[e].w == sp.off[o].w) e++;
        sp.rows.push_back(make_int2(e - o, sp.off[o].x - sp.off[o].y));
        o = e;
    }
    for (const int4 &q : sp.off) sp.rx = std::max(sp.rx, std::abs(q.y));
    sp.nrow = (int)sp.rows.size();
    for (int k = 0; k < ST_PAD; k++) sp.rows.push_back(make_int2(0, 0)); // read ahead, never a row
    // rows in ascending offset order, each offset at most once per node
    sp.oi.assign(nB, 0);
    sp.oj.assign(nB, 0);
    std::vector<int> last(nN, -1);
    sp.g.assign(nN, 0);
    for (int b = 0; b < nB; b++) {
        const int i = bi[b], j = bj[b];
        int dd[3];
        for (int d = 0; d < 3; d++) dd[d] = sp.k[3 * (size_t)j + d] - sp.k[3 * (size_t)i + d];
        const int a = idx[cell(dd[0], dd[1], dd[2])], c = idx[cell(-dd[0], -dd[1], -dd[2])];
        if (a <= last[i] || c <= last[j]) {
            sp.why = "rows not in ascending lattice-offset order (BONDLIST not sorted by node id)";
            return;
        }
        last[i] = a;
        last[j] = c;
        sp.oi[b] = a;
        sp.oj[b] = c;
        sp.g[i]++;
        sp.g[j]++;
    }
    // BONDSTIFFNESS (and s0 / sc where possible) as tables keyed by the pair of family sizes
    // (nFAMILY_i, nFAMILY_j): each node carries the index of its size among the distinct sizes
    std::vector<int> sizes(sp.g.begin(), sp.g.end());
    std::sort(sizes.begin(), sizes.end());
    sizes.erase(std::unique(sizes.begin(), sizes.end()), sizes.end());
    if ((int)sizes.size() > ST_GMAX) {
        sp.why = "more than " + std::to_string(ST_GMAX) + " distinct family sizes";
        return;
    }
    sp.G = (int)sizes.size();
    sp.gidx.assign(nN, 0);
    for (int n = 0; n < nN; n++) sp.gidx[n] = (int)(std::lower_bound(sizes.begin(), sizes.end(), sp.g[n]) - sizes.begin());
    sp.nkey = sp.G * sp.G;
    // both orders of the pair index the same value (the reference's vi + vj is commutative)
    std::vector<int> key(2 * (size_t)nB);
    for (int b = 0; b < nB; b++) {
        key[2 * (size_t)b] = sp.gidx[bi[b]] * sp.G + sp.gidx[bj[b]];
        key[2 * (size_t)b + 1] = sp.gidx[bj[b]] * sp.G + sp.gidx[bi[b]];
    }
    auto pair_table = [&](const double *a, std::vector<double> &tab) {
        std::vector<double> both(2 * (size_t)nB);
        for (int b = 0; b < nB; b++) both[2 * (size_t)b] = both[2 * (size_t)b + 1] = a[b];
        return key_table(both.data(), 2 * nB, key, sp.nkey, tab);
    };
    if (!pair_table(pr->BONDSTIFFNESS, sp.tc)) {
        sp.why = "BONDSTIFFNESS is not a function of (nFAMILY_i, nFAMILY_j)";
        return;
    }
    sp.s0mode = PM_SCALAR;
    sp.scmode = PM_SCALAR;
    if (pr->s0) sp.s0mode = pair_table(pr->s0, sp.ts0) ? PM_TABLE : PM_STREAM;
    if (pr->sc) sp.scmode = pair_table(pr->sc, sp.tsc) ? PM_TABLE : PM_STREAM;
    if (sp.s0mode != PM_TABLE) sp.ts0.clear();
    if (sp.scmode != PM_TABLE) sp.tsc.clear();
    // one of s0 / sc tabled and the other scalar: table the scalar too, so the bond kernel reads
    // both from tables with no per-entry choice
    if (sp.s0mode == PM_TABLE && sp.scmode == PM_SCALAR) {
        sp.tsc.assign(sp.nkey, pr->sc_scalar);
        sp.scmode = PM_TABLE;
    } else if (sp.scmode == PM_TABLE && sp.s0mode == PM_SCALAR) {
        sp.ts0.assign(sp.nkey, pr->s0_scalar);
        sp.s0mode = PM_TABLE;
    }
    sp.ok = true;
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
    BBPD_REQUIRE(pr->schedule >= 0 && pr->schedule <= 2, "unknown load schedule");
    BBPD_REQUIRE(pr->schedule == 0 ? pr->nTimeSteps > 1 : pr->itersPerIncrement > 0, "bad load schedule");
    BBPD_REQUIRE(pr->schedule != 2 || (pr->integrator == BBPD_INTEGRATOR_ADR && pr->eqTol >= 0 &&
                                       pr->itersPerIncrementMin >= 0),
                 "the equilibrium staircase needs ADR, eqTol >= 0 and itersPerIncrementMin >= 0");
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

    // ---- bond arrays and lattice
    const double horizon = pr->horizon, RADIJ = pr->RADIJ;
    const double hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2 * RADIJ;
    std::vector<int> bi, bj;
    {
        const int rc0 = check_bonds(pr, bi, bj);
        if (rc0) return fail(rc0);
    }
    double xmin[3], xmax[3], hq, hlat[3];
    const bool lattice = detect_lattice(pr, hlat, xmin, xmax, hq);
    s->lattice = lattice;

    // ---- layout.  Both forms take the stencil layout where it applies: it moves a fraction of the
    // record layout's bytes per entry and loads each neighbour record once per stencil row.
    // BBPD_LAYOUT=records asks for the record layout.
    StencilPlan sp;
    {
        const char *lay = std::getenv("BBPD_LAYOUT");
        const bool want_records = lay && std::strcmp(lay, "records") == 0;
        if (!lattice) sp.why = "not a lattice";
        else if (want_records) sp.why = "record layout requested";
        else plan_stencil(pr, X, bi, bj, hlat, sp);
    }
    s->stencil = sp.ok;
    s->layout = sp.ok ? "stencil" : "records (" + sp.why + ")";
    if (s->stencil && env_launch_variant() >= 0) s->occ = env_launch_variant();

    // ---- slots.  Stencil layout: the x-fastest lattice order.  Records: tiles of 64 x 2 x 2 nodes (64 along the longest axis) visited in Morton order,
    // lines of 64 inside a tile, so a wavefront is 64 consecutive nodes of a lattice line and the
    // k-th neighbours of its lanes are again 64 consecutive slots.
    int ax[3] = {0, 1, 2}; // ax[0] = longest axis
    std::sort(ax, ax + 3, [&](int a, int b) { return (xmax[a] - xmin[a]) > (xmax[b] - xmin[b]); });
    std::vector<uint64_t> code(sp.ok ? 0 : nN);
    for (int n = 0; n < (sp.ok ? 0 : nN); n++) {
        uint32_t q[3];
        for (int d = 0; d < 3; d++) {
            double t = std::floor((X[n + (size_t)nN * ax[d]] - xmin[ax[d]]) / hq + 0.5);
            q[d] = (uint32_t)std::min(std::max(t, 0.0), 2097151.0);
        }
        uint64_t tile = morton3(q[0] >> 6, q[1] >> 1, q[2] >> 1) & ((1ull << 55) - 1);
        uint64_t in = ((uint64_t)((q[2] & 1u) * 2u + (q[1] & 1u)) << 6) | (q[0] & 63u);
        code[n] = (tile << 8) | in;
    }
    const int nSlots = (int)(((long long)nN + SLOT_ALIGN - 1) / SLOT_ALIGN * SLOT_ALIGN);
    s->nSlots = nSlots;
    s->nSlices = nSlots / SLICE;
    if (sp.ok) {
        s->slot_of = sp.slot;
    } else {
        std::vector<int> order(nN);
        std::iota(order.begin(), order.end(), 0);
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return code[a] < code[b]; });
        s->slot_of.assign(nN, 0);
        for (int q = 0; q < nN; q++) s->slot_of[order[q]] = q;
    }

    // ---- rows in BONDLIST order
    std::vector<int> rowlen(nSlots, 0);
    for (int k = 0; k < nB; k++) {
        rowlen[s->slot_of[bi[k]]]++;
        rowlen[s->slot_of[bj[k]]]++;
    }
    // record layout: SELL-64 slices of the rows; stencil layout: one entry per (slot, offset)
    std::vector<long long> slice_off(s->nSlices + 1, 0);
    long long nE = 0;
    if (!sp.ok) {
        for (int sl = 0; sl < s->nSlices; sl++) {
            int w = 0;
            for (int t = 0; t < SLICE; t++) w = std::max(w, rowlen[sl * SLICE + t]);
            slice_off[sl + 1] = slice_off[sl] + (long long)w * SLICE;
        }
        nE = slice_off[s->nSlices];
    } else {
        nE = (long long)s->nSlices * sp.noff * SLICE;
    }
    s->nEntries = 2LL * nB;
    s->nPadded = nE;
    std::vector<int> fillpos(nSlots, 0);
    s->entry_i.assign(nB, 0);
    s->entry_j.assign(nB, 0);
    auto entry_at = [&](int slot, int kk) { return slice_off[slot / SLICE] + (long long)kk * SLICE + (slot % SLICE); };
    auto dense_at = [&](int slot, int o) {
        return ((long long)(slot / SLICE) * sp.noff + o) * SLICE + (slot % SLICE);
    };

    // ---- per-bond quantities: tables (FP64 form), inline (mixed form) or streams
    int cmode = PM_STREAM, s0mode = pr->s0 ? PM_STREAM : PM_SCALAR, scmode = pr->sc ? PM_STREAM : PM_SCALAR;
    std::vector<double> tabc, tabp;
    std::vector<uint32_t> cidx, pidx;
    if (sp.ok) {
        // key tables: c, then s0 and / or sc where they are functions of the key
        cmode = PM_TABLE;
        s0mode = sp.s0mode;
        scmode = sp.scmode;
        tabc = sp.tc;
        tabp = sp.ts0;
        tabp.insert(tabp.end(), sp.tsc.begin(), sp.tsc.end());
    } else if (!mixed) {
        if (make_table(pr->BONDSTIFFNESS, nB, TAB_MAX, tabc, cidx)) cmode = PM_TABLE;
        // one per-bond parameter may use the second table
        if (cmode == PM_TABLE && (pr->s0 != nullptr) != (pr->sc != nullptr)) {
            const double *pa = pr->s0 ? pr->s0 : pr->sc;
            if (make_table(pa, nB, TAB_MAX, tabp, pidx)) {
                if (pr->s0) s0mode = PM_TABLE;
                else scmode = PM_TABLE;
            }
        }
        if (cmode != PM_TABLE) tabc.clear();
        if (s0mode != PM_TABLE && scmode != PM_TABLE) tabp.clear();
    } else {
        // the first per-bond parameter travels inside the record; a second one is streamed
        if (pr->s0) s0mode = PM_INLINE;
        else if (pr->sc) scmode = PM_INLINE;
    }
    std::vector<uint4> h_rec(sp.ok ? 0 : nE);
    if (!sp.ok) std::memset(h_rec.data(), 0, sizeof(uint4) * nE);
    std::vector<double> h_c64, h_s064, h_sc64;
    std::vector<float> h_s032, h_sc032;
    if (!mixed && cmode == PM_STREAM) h_c64.assign(nE, 0.0);
    if (s0mode == PM_STREAM) {
        if (mixed) h_s032.assign(nE, 0.f);
        else h_s064.assign(nE, 0.0);
    }
    if (scmode == PM_STREAM) {
        if (mixed) h_sc032.assign(nE, 0.f);
        else h_sc64.assign(nE, 0.0);
    }
    const double V = pr->cellVolume;
    for (int k = 0; k < nB; k++) {
        const int si = s->slot_of[bi[k]], sj = s->slot_of[bj[k]];
        if (sp.ok) {
            const long long ei = dense_at(si, sp.oi[k]), ej = dense_at(sj, sp.oj[k]);
            s->entry_i[k] = ei;
            s->entry_j[k] = ej;
            for (long long e : {ei, ej}) {
                if (s0mode == PM_STREAM) {
                    if (mixed) h_s032[e] = (float)pr->s0[k];
                    else h_s064[e] = pr->s0[k];
                }
                if (scmode == PM_STREAM) {
                    if (mixed) h_sc032[e] = (float)pr->sc[k];
                    else h_sc64[e] = pr->sc[k];
                }
            }
            continue;
        }
        const uint32_t type = pr->BONDTYPE ? (uint32_t)pr->BONDTYPE[k] : 0u;
        const uint32_t flags = (type << TYPE_SHIFT) | ((pr->fail && pr->fail[k] == 0) ? FAILED_BIT : 0u);
        const long long ei = entry_at(si, fillpos[si]++);
        const long long ej = entry_at(sj, fillpos[sj]++);
        s->entry_i[k] = ei;
        s->entry_j[k] = ej;
        const long long ends[2] = {ei, ej};
        const int other[2] = {sj, si};
        for (int t = 0; t < 2; t++) {
            const long long e = ends[t];
            uint4 r;
            r.x = (uint32_t)other[t] | flags;
            if (!mixed) {
                uint32_t ti = 0;
                if (cmode == PM_TABLE) ti |= cidx[k];
                if (s0mode == PM_TABLE || scmode == PM_TABLE) ti |= pidx[k] << 16;
                r.y = ti;
                r.z = 0; // history starts at 0
                r.w = 0;
                if (cmode == PM_STREAM) h_c64[e] = pr->BONDSTIFFNESS[k];
                if (s0mode == PM_STREAM) h_s064[e] = pr->s0[k];
                if (scmode == PM_STREAM) h_sc64[e] = pr->sc[k];
            } else {
                float pin = 0.f;
                if (s0mode == PM_INLINE) pin = (float)pr->s0[k];
                else if (scmode == PM_INLINE) pin = (float)pr->sc[k];
                uint32_t pb;
                std::memcpy(&pb, &pin, 4);
                r.y = pb;
                r.z = 0;
                const float kap = (float)(pr->BONDSTIFFNESS[k] * V * pr->VOLUMECORRECTIONFACTORS[k]);
                std::memcpy(&r.w, &kap, 4);
                if (s0mode == PM_STREAM) h_s032[e] = (float)pr->s0[k];
                if (scmode == PM_STREAM) h_sc032[e] = (float)pr->sc[k];
            }
            h_rec[e] = r;
        }
    }
    // padding entries point at their own slot (their arithmetic is masked off)
    for (int sl = 0; sl < (sp.ok ? 0 : s->nSlices); sl++)
        for (int t = 0; t < SLICE; t++) {
            int slot = sl * SLICE + t;
            long long w = (slice_off[sl + 1] - slice_off[sl]) / SLICE;
            for (long long kk = rowlen[slot]; kk < w; kk++) h_rec[entry_at(slot, (int)kk)].x = (uint32_t)slot;
        }

    // ---- device allocations
    BBPD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    Params &P = s->P;
    P.nSlots = nSlots;
    int rc = 0;
    double *Xd, *ud = nullptr, *vd, *Fd, *Fpd = nullptr, *rhod, *massd = nullptr;
    NodeRec *posd[2], *Xrecd = nullptr;
    unsigned char *consd;
    if ((rc = dalloc(&Xd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    for (int q = 0; q < 2; q++)
        if ((rc = dalloc(&posd[q], nSlots, s->owned))) return fail(rc);
    if (!lattice && (rc = dalloc(&Xrecd, nSlots, s->owned))) return fail(rc);
    if (!mixed && (rc = dalloc(&ud, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&vd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&Fd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if (pr->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&Fpd, 3 * (size_t)nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&rhod, nSlots, s->owned))) return fail(rc);
    if (pr->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&massd, nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&consd, nSlots, s->owned))) return fail(rc);
    int *rowlend;
    long long *soffd;
    uint4 *recd;
    if ((rc = dalloc(&rowlend, nSlots, s->owned))) return fail(rc);
    if ((rc = dalloc(&soffd, s->nSlices + 1, s->owned))) return fail(rc);
    if (!sp.ok && (rc = dalloc(&recd, nE, s->owned))) return fail(rc);
    // stencil layout
    void *hsd = nullptr;
    uint4 *maskd = nullptr;
    unsigned long long *maskTd = nullptr, *snzd = nullptr;
    int4 *stoffd = nullptr;
    int2 *srowd = nullptr;
    StOff *sstd = nullptr;
    float *tabfd = nullptr;
    if (sp.ok) {
        if ((rc = dalloc((char **)&hsd, (size_t)(nE + 2 * SLICE) * (mixed ? 4 : 8), s->owned))) return fail(rc);
        if ((rc = dalloc(&maskd, nSlots, s->owned))) return fail(rc);
        if ((rc = dalloc(&maskTd, (size_t)s->nSlices * sp.noff + ST_PAD, s->owned))) return fail(rc);
        if ((rc = dalloc(&snzd, (size_t)s->nSlices * 2, s->owned))) return fail(rc);
        if ((rc = dalloc(&stoffd, sp.noff, s->owned))) return fail(rc);
        if ((rc = dalloc(&srowd, sp.rows.size(), s->owned))) return fail(rc);
        if ((rc = dalloc(&sstd, sp.noff + ST_PAD, s->owned))) return fail(rc);
    }
    double *cstrd = nullptr, *tabd = nullptr;
    void *s0d = nullptr, *scd = nullptr;
    const size_t ew = mixed ? 4 : 8;
    if (!h_c64.empty() && (rc = dalloc(&cstrd, nE, s->owned))) return fail(rc);
    if (s0mode == PM_STREAM && (rc = dalloc((char **)&s0d, (nE + 2 * SLICE) * ew, s->owned))) return fail(rc);
    if (scmode == PM_STREAM && (rc = dalloc((char **)&scd, (nE + 2 * SLICE) * ew, s->owned))) return fail(rc);
    const size_t ntab = tabc.size() + tabp.size();
    if ((rc = dalloc(&tabd, ntab > 0 ? ntab : 1, s->owned))) return fail(rc);
    if (sp.ok && mixed && (rc = dalloc(&tabfd, ntab > 0 ? ntab : 1, s->owned))) return fail(rc);
    double *partd;
    if ((rc = dalloc(&partd, 4 * (size_t)(nSlots / BLOCK) + 4, s->owned))) return fail(rc);

    if ((rc = dalloc(&s->counts, 4, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_hist, (size_t)HIST_CAP * N_HIST, s->owned))) return fail(rc);

    // ---- host staging in slot order
    {
        std::vector<double> hX(3 * (size_t)nSlots, 0.0), zero(3 * (size_t)nSlots, 0.0);
        std::vector<NodeRec> hpos(nSlots), hXr(lattice ? 0 : nSlots);
        std::memset(hpos.data(), 0, sizeof(NodeRec) * nSlots);
        std::vector<double> hrho(nSlots, 1.0), hm(nSlots, 1.0);
        std::vector<unsigned char> hc(nSlots, 0);
        for (int n = 0; n < nN; n++) {
            int q = s->slot_of[n];
            double xv[3];
            for (int d = 0; d < 3; d++) {
                xv[d] = X[n + (size_t)nN * d];
                hX[q + (size_t)nSlots * d] = xv[d];
                if (pr->CONSTRAINTFLAG && pr->CONSTRAINTFLAG[n + (size_t)nN * d] == 1) hc[q] |= (unsigned char)(1u << d);
            }
            // FP64 form: positions start at deformedCoordinates = undeformedCoordinates
            // (dynamicsolverdisplacementcontrolled.m:33); mixed form: displacements start at 0
            hpos[q].x = mixed ? 0.0 : xv[0];
            hpos[q].y = mixed ? 0.0 : xv[1];
            hpos[q].z = mixed ? 0.0 : xv[2];
            if (lattice) {
                uint64_t k0 = (uint64_t)std::floor(xv[0] / hlat[0] + 0.5);
                uint64_t k1 = (uint64_t)std::floor(xv[1] / hlat[1] + 0.5);
                uint64_t k2 = (uint64_t)std::floor(xv[2] / hlat[2] + 0.5);
                hpos[q].q = k0 | (k1 << LAT_BITS) | (k2 << (2 * LAT_BITS));
                if (sp.ok) hpos[q].q |= (uint64_t)sp.gidx[n] << G_SHIFT;
            } else {
                hXr[q].x = xv[0];
                hXr[q].y = xv[1];
                hXr[q].z = xv[2];
                hXr[q].q = 0;
            }
            hrho[q] = pr->DENSITY[n];
            if (pr->massVector) hm[q] = pr->massVector[n];
        }
        BBPD_HIP(hipMemcpy(Xd, hX.data(), sizeof(double) * hX.size(), hipMemcpyHostToDevice));
        for (int q = 0; q < 2; q++)
            BBPD_HIP(hipMemcpy(posd[q], hpos.data(), sizeof(NodeRec) * nSlots, hipMemcpyHostToDevice));
        if (Xrecd) BBPD_HIP(hipMemcpy(Xrecd, hXr.data(), sizeof(NodeRec) * nSlots, hipMemcpyHostToDevice));
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
    if (!sp.ok) BBPD_HIP(hipMemcpy(recd, h_rec.data(), sizeof(uint4) * nE, hipMemcpyHostToDevice));
    int fast64 = 0; // FP64 stencil kernel: per-offset nominal lengths apply (offset_records)
    if (sp.ok) {
        // per-slot masks; history 0, its sign bit set for bonds that start failed (brittle law)
        std::vector<uint4> hm(nSlots, make_uint4(0, 0, 0, 0));
        auto setbit = [&](int slot, int o) {
            uint32_t *w = reinterpret_cast<uint32_t *>(&hm[slot]);
            w[o >> 5] |= 1u << (o & 31);
        };
        for (int k = 0; k < nB; k++) {
            setbit(s->slot_of[bi[k]], sp.oi[k]);
            setbit(s->slot_of[bj[k]], sp.oj[k]);
        }
        BBPD_HIP(hipMemcpy(maskd, hm.data(), sizeof(uint4) * nSlots, hipMemcpyHostToDevice));
        const std::vector<unsigned long long> hmt = lane_masks(hm, nSlots, sp.noff);
        BBPD_HIP(hipMemcpy(maskTd, hmt.data(), sizeof(unsigned long long) * hmt.size(), hipMemcpyHostToDevice));
        if (mixed) {
            std::vector<float> h0(nE, 0.0f);
            for (int k = 0; k < nB; k++)
                if (pr->fail && pr->fail[k] == 0) h0[s->entry_i[k]] = h0[s->entry_j[k]] = -0.0f;
            BBPD_HIP(hipMemcpy(hsd, h0.data(), sizeof(float) * nE, hipMemcpyHostToDevice));
            const std::vector<unsigned long long> hw = history_words(h0.data(), s->nSlices, sp.noff);
            BBPD_HIP(hipMemcpy(snzd, hw.data(), sizeof(unsigned long long) * hw.size(), hipMemcpyHostToDevice));
        } else {
            std::vector<double> h0(nE, 0.0);
            for (int k = 0; k < nB; k++)
                if (pr->fail && pr->fail[k] == 0) h0[s->entry_i[k]] = h0[s->entry_j[k]] = -0.0;
            BBPD_HIP(hipMemcpy(hsd, h0.data(), sizeof(double) * nE, hipMemcpyHostToDevice));
            const std::vector<unsigned long long> hw = history_words(h0.data(), s->nSlices, sp.noff);
            BBPD_HIP(hipMemcpy(snzd, hw.data(), sizeof(unsigned long long) * hw.size(), hipMemcpyHostToDevice));
        }
        BBPD_HIP(hipMemcpy(stoffd, sp.off.data(), sizeof(int4) * sp.noff, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(srowd, sp.rows.data(), sizeof(int2) * sp.rows.size(), hipMemcpyHostToDevice));
        // mixed form per offset: xi = h * d rounded to FP32 (so the offset and its negation give
        // exactly negated vectors), |xi|, and the volume correction factor of |h * d|
        const std::vector<StOff> hst =
            offset_records(sp.off, hlat, hmr, hpr, horizon, two_r, mixed, sp.kmin, sp.kmax, &fast64);
        // BBPD_FAST64=0 asks for the lengths computed per entry
        const char *f64 = std::getenv("BBPD_FAST64");
        if (f64 && std::string(f64) == "0") fast64 = 0;
        if (!mixed && !fast64) s->layout = "stencil (FP64 lengths computed per entry)";
        BBPD_HIP(hipMemcpy(sstd, hst.data(), sizeof(StOff) * hst.size(), hipMemcpyHostToDevice));
        if (mixed) {
            // c * V, then the parameter tables, in FP32
            std::vector<float> tf(ntab, 0.0f);
            for (size_t i = 0; i < tabc.size(); i++) tf[i] = (float)(tabc[i] * V);
            for (size_t i = 0; i < tabp.size(); i++) tf[tabc.size() + i] = (float)tabp[i];
            if (ntab) BBPD_HIP(hipMemcpy(tabfd, tf.data(), sizeof(float) * ntab, hipMemcpyHostToDevice));
        }
    }
    if (cstrd) BBPD_HIP(hipMemcpy(cstrd, h_c64.data(), 8 * nE, hipMemcpyHostToDevice));
    if (s0d) BBPD_HIP(hipMemcpy(s0d, mixed ? (void *)h_s032.data() : (void *)h_s064.data(), ew * nE, hipMemcpyHostToDevice));
    if (scd) BBPD_HIP(hipMemcpy(scd, mixed ? (void *)h_sc032.data() : (void *)h_sc64.data(), ew * nE, hipMemcpyHostToDevice));
    {
        std::vector<double> t(tabc);
        t.insert(t.end(), tabp.begin(), tabp.end());
        if (!t.empty()) BBPD_HIP(hipMemcpy(tabd, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice));
    }
    BBPD_HIP(hipMemset(s->d_hist, 0, sizeof(double) * (size_t)HIST_CAP * N_HIST));

    auto comp = [&](double *base, int d) { return base ? base + (size_t)nSlots * d : nullptr; };
    P.X0 = comp(Xd, 0);
    P.X1 = comp(Xd, 1);
    P.X2 = comp(Xd, 2);
    P.pos[0] = posd[0];
    P.pos[1] = posd[1];
    P.Xrec = Xrecd;
    P.hx = hlat[0];
    P.hy = hlat[1];
    P.hz = hlat[2];
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
    P.rec = recd;
    P.cstr = cstrd;
    P.s0str = s0d;
    P.scstr = scd;
    P.tab = tabd;
    P.ntabc = (int)tabc.size();
    P.ntabp = (int)tabp.size();
    P.cmode = cmode;
    P.s0mode = s0mode;
    P.scmode = scmode;
    P.adr = pr->integrator == BBPD_INTEGRATOR_ADR ? 1 : 0;
    P.law_linear = pr->law == BBPD_LAW_LINEAR ? 1 : 0;
    P.V = V;
    P.horizon = horizon;
    P.hmr = hmr;
    P.hpr = hpr;
    P.two_r = two_r;
    P.rtwo_r = 1.0 / two_r;
    P.k = pr->k;
    P.alpha = pr->alpha;
    P.denom_exp = 1 - std::exp(-pr->k);
    P.sc_steel = pr->sc_steel;
    P.ff = pr->failureFunctionality;
    P.DT = pr->DT;
    P.D = pr->DAMPING;
    P.partials = partd;
    P.stencil = sp.ok ? 1 : 0;
    P.soff = stoffd;
    P.srow = srowd;
    P.sst = sstd;
    P.smask = maskd;
    P.smaskT = maskTd;
    P.hs = hsd;
    P.snz = snzd;
    P.tabf = tabfd;
    // key tables in LDS; with s0 and sc both tabled the stencil kernel keeps (c, law threshold) per key
    {
        const size_t w = mixed ? sizeof(float) : sizeof(double);
        const bool both = s0mode == PM_TABLE && scmode == PM_TABLE;
        s->lds_bytes = sp.ok ? w * (both ? (size_t)(2 * sp.G * sp.G) : ntab) : (mixed ? 0 : w * ntab);
    }
    if ((rc = allow_lds(s))) return fail(rc);

    // ---- region 0: contact bodies, load schedule, observables
    Region R0{};
    R0.base = 0;
    R0.nN = nN;
    R0.nSlots = nSlots;
    R0.blk0 = 0;
    R0.nblk = nSlots / BLOCK;
    R0.e0 = 0;
    R0.noff = sp.ok ? sp.noff : 0;
    R0.off0 = 0;
    R0.nrow = sp.ok ? sp.nrow : 0;
    R0.row0 = 0;
    R0.rx = sp.rx;
    R0.sst0 = 0;
    R0.fast64 = fast64;
    R0.kg = sp.G;
    R0.tab0 = 0;
    R0.ntab = (int)(tabc.size() + tabp.size());
    R0.ts0 = (int)tabc.size();
    R0.tsc = (int)(tabc.size() + (sp.ok ? sp.ts0.size() : 0));
    R0.s0 = pr->s0 ? 0.0 : pr->s0_scalar;
    R0.sc = pr->sc ? 0.0 : pr->sc_scalar;
    R0.s1 = pr->s1;
    R0.s0lo = pr->s0 && nB > 0 ? *std::min_element(pr->s0, pr->s0 + nB) : 0.0;
    R0.sclo = pr->sc && nB > 0 ? *std::min_element(pr->sc, pr->sc + nB) : 0.0;
    R0.body0 = 0;
    R0.nBodies = pr->nBodies;
    R0.schedule = pr->schedule;
    R0.nTimeSteps = pr->nTimeSteps;
    R0.ipi = pr->itersPerIncrement > 0 ? pr->itersPerIncrement : 1;
    R0.applied = pr->appliedDisplacement;
    R0.increment = pr->increment;
    R0.eq_tol = pr->eqTol;
    R0.ipi_min = pr->itersPerIncrementMin;
    auto slot_or_none = [&](int id1) { return (id1 >= 1 && id1 <= nN) ? s->slot_of[id1 - 1] : -1; };
    R0.ref_slot = slot_or_none(pr->referenceNode);
    R0.cmod1 = slot_or_none(pr->cmodNode1);
    R0.cmod2 = slot_or_none(pr->cmodNode2);
    R0.hist0 = 0;
    R0.hist_cap = HIST_CAP;
    R0.stop_frac = 0;
    R0.stop_disp = 0;
    // disjoint families are projected concurrently (one step-end block each); when families share
    // members, one block applies the bodies in order, as the reference's calls follow one another.
    // A family lists each member once (its members are projected in parallel).
    std::vector<unsigned char> in_family(nN, 0);
    std::vector<int> owner(nN, -1);
    R0.seq_contact = 0;
    std::vector<Body> hb(std::max(pr->nBodies, 1));
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
            if (owner[id - 1] == bdy) {
                bbpd::set_error("contact family " + std::to_string(bdy) + " lists node " + std::to_string(id) +
                                " twice");
                return fail(BBPD_ERR_ARG);
            }
            if (owner[id - 1] >= 0) R0.seq_contact = 1;
            owner[id - 1] = bdy;
            in_family[id - 1] = 1;
            fs[t] = s->slot_of[id - 1];
        }
        int *fd;
        if ((rc = dalloc(&fd, fs.size(), s->owned))) return fail(rc);
        BBPD_HIP(hipMemcpy(fd, fs.data(), sizeof(int) * fs.size(), hipMemcpyHostToDevice));
        Body &B = hb[bdy];
        B.fam = fd;
        B.nFam = nf;
        B.cx = pr->centreX[bdy];
        B.cz = pr->centreZ[bdy];
        B.R = pr->radius[bdy];
        B.moves = pr->moves[bdy];
        B.region = 0;
        B.index = bdy;
    }
    if (pr->integrator == BBPD_INTEGRATOR_ADR && pr->nBodies > 0) {
        std::vector<unsigned char> hc(nSlots);
        BBPD_HIP(hipMemcpy(hc.data(), consd, nSlots, hipMemcpyDeviceToHost));
        for (int n = 0; n < nN; n++)
            if (in_family[n]) hc[s->slot_of[n]] |= CONTACT_BIT;
        BBPD_HIP(hipMemcpy(consd, hc.data(), nSlots, hipMemcpyHostToDevice));
    }
    {
        // descriptors, state, block tables, step-end jobs, parity tick
        const int nblk = nSlots / BLOCK;
        std::vector<int> blkreg(nblk, 0), grid(nblk);
        std::iota(grid.begin(), grid.end(), 0);
        std::vector<int2> se;
        if (R0.seq_contact) se.push_back(make_int2(0, -2));
        else
            for (int bdy = 0; bdy < std::max(pr->nBodies, 1); bdy++)
                se.push_back(make_int2(0, pr->nBodies > 0 ? bdy : -1));
        int act = 0;
        if ((rc = dalloc(&s->d_reg, 1, s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_rst, 1, s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_blkreg, nblk, s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_grid, nblk, s->owned))) return fail(rc);
        const std::vector<int> gst = stripe_list(grid);
        if ((rc = dalloc(&s->d_grid_st, gst.size(), s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_active, 1, s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_se, se.size(), s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_bodies, hb.size(), s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_tick, 1, s->owned))) return fail(rc);
        if ((rc = dalloc(&s->d_tick_done, 1, s->owned))) return fail(rc);
        RegionState z{};
        z.inc = 1;
        long long t0 = 0;
        unsigned u0 = 0;
        BBPD_HIP(hipMemcpy(s->d_reg, &R0, sizeof(Region), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_rst, &z, sizeof(RegionState), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_blkreg, blkreg.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_grid, grid.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_grid_st, gst.data(), sizeof(int) * gst.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_active, &act, sizeof(int), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_se, se.data(), sizeof(int2) * se.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_bodies, hb.data(), sizeof(Body) * hb.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_tick, &t0, sizeof(t0), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_tick_done, &u0, sizeof(u0), hipMemcpyHostToDevice));
        s->reg.assign(1, R0);
        s->active.assign(1, 1);
        s->hist_read.assign(1, 0);
        s->grid_n = nblk;
        s->grid_n_st = (int)gst.size();
        s->nse = (int)se.size();
        s->nactive = 1;
        P.reg = s->d_reg;
        P.rst = s->d_rst;
        P.blkreg = s->d_blkreg;
        P.grid = s->d_grid;
        P.active = s->d_active;
        P.se = s->d_se;
        P.nse = s->nse;
        P.bodies = s->d_bodies;
        P.tick = s->d_tick;
        P.tick_done = s->d_tick_done;
        P.hist = s->d_hist;
    }

    s->device_bytes = sizeof(NodeRec) * (size_t)nSlots * (lattice ? 2 : 3) +
                      sizeof(double) * (size_t)nSlots * (3 * (mixed ? 3 : 4) + 2 + (Fpd ? 3 : 0)) + nSlots +
                      (sp.ok ? (mixed ? 4 : 8) * (size_t)nE + sizeof(uint4) * (size_t)nSlots : sizeof(uint4) * (size_t)nE) +
                      (cstrd ? 8 * nE : 0) + (s0d ? ew * nE : 0) + (scd ? ew * nE : 0);

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
    if (!s->gexec || s->graph_steps != n || s->graph_version != s->version) {
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
        s->graph_version = s->version;
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

namespace {

int region_observables(bbpd_solver *s, int ri, bbpd_observables *o)
{
    BBPD_HIP(hipSetDevice(s->device));
    const Region &RG = s->reg[ri];
    BBPD_HIP(hipMemsetAsync(s->counts, 0, sizeof(unsigned long long) * 4, s->stream));
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const unsigned g = (unsigned)RG.nblk;
    if (mixed) hipLaunchKernelGGL(k_count<true>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, ri, s->counts);
    else hipLaunchKernelGGL(k_count<false>, dim3(g), dim3(BLOCK), 0, s->stream, s->P, ri, s->counts);
    BBPD_HIP(hipGetLastError());
    unsigned long long c[4];
    RegionState d;
    BBPD_HIP(hipMemcpyAsync(c, s->counts, sizeof(c), hipMemcpyDeviceToHost, s->stream));
    BBPD_HIP(hipMemcpyAsync(&d, s->d_rst + ri, sizeof(d), hipMemcpyDeviceToHost, s->stream));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    std::memset(o, 0, sizeof(*o));
    o->step = d.step;
    o->reaction = RG.nBodies > 0 ? d.reaction[0] : 0.0;
    o->reaction_static = d.reaction_static;
    o->cn = d.cn;
    o->n_softening = (int64_t)(c[0] / 2);
    o->n_broken = (int64_t)(c[1] / 2);
    o->n_failed = (int64_t)(c[2] / 2);
    if (d.step > 0) {
        double h[N_HIST];
        BBPD_HIP(hipMemcpy(h, s->d_hist + (size_t)(RG.hist0 + (d.step - 1) % RG.hist_cap) * N_HIST, sizeof(h),
                           hipMemcpyDeviceToHost));
        o->deflection = h[2];
        o->cmod = h[3];
    }
    return BBPD_OK;
}

} // namespace

namespace {

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

// damage field (reference node order) and total damage energy of one region
// damage field, damage energy and (when asked) the strain energy of one region, in reference order
int region_damage(bbpd_solver *s, int ri, const std::vector<int> &slot_of, double *damage, double *energy,
                  double *density = nullptr, double *strain = nullptr)
{
    BBPD_HIP(hipSetDevice(s->device));
    const Region &RG = s->reg[ri];
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const bool want_sed = density || strain;
    double *dd = nullptr, *ep = nullptr, *sd = nullptr;
    BBPD_HIP(hipMalloc((void **)&dd, sizeof(double) * (size_t)RG.nSlots));
    hipError_t e = hipMalloc((void **)&ep, sizeof(double) * (size_t)RG.nblk);
    if (e == hipSuccess && want_sed) e = hipMalloc((void **)&sd, sizeof(double) * (size_t)RG.nSlots);
    if (e != hipSuccess) {
        (void)hipFree(dd);
        if (ep) (void)hipFree(ep);
        return bbpd::hip_fail(e, "hipMalloc (damage)");
    }
    DamageFn f = mixed ? pick_damage<true>(s->law) : pick_damage<false>(s->law);
    hipLaunchKernelGGL(f, dim3(RG.nblk), dim3(BLOCK), 0, s->stream, s->P, ri, dd, ep, sd);
    std::vector<double> hd(RG.nSlots), he(RG.nblk), hs(want_sed ? RG.nSlots : 0);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(hd.data(), dd, sizeof(double) * hd.size(), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(he.data(), ep, sizeof(double) * he.size(), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess && want_sed)
        e = hipMemcpyAsync(hs.data(), sd, sizeof(double) * hs.size(), hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(dd);
    (void)hipFree(ep);
    if (sd) (void)hipFree(sd);
    if (e != hipSuccess) return bbpd::hip_fail(e, "damage kernel");
    if (damage)
        for (size_t n = 0; n < slot_of.size(); n++) damage[n] = hd[slot_of[n] - RG.base];
    if (energy) {
        double t = 0;
        for (double x : he) t = t + x;
        *energy = t;
    }
    if (want_sed) {
        // calculatestrainenergy.m:26: sum(nodalStrainEnergyDensity) * nNodes * cellVolume, the sum in
        // node order
        double t = 0;
        for (size_t n = 0; n < slot_of.size(); n++) {
            const double v = hs[slot_of[n] - RG.base];
            if (density) density[n] = v;
            t = t + v;
        }
        if (strain) *strain = t * (double)slot_of.size() * s->P.V;
    }
    return BBPD_OK;
}

} // namespace

extern "C" int bbpd_solver_damage(bbpd_solver *s, double *damage, double *energy)
{
    BBPD_REQUIRE(s && !s->batch, "bad arguments (a batch takes bbpd_batch_damage)");
    BBPD_REQUIRE(!energy || s->law != BBPD_LAW_LINEAR,
                 "damage energy needs a softening law (the brittle latch keeps no stretch history)");
    return region_damage(s, 0, s->slot_of, damage, energy);
}

extern "C" int bbpd_solver_strain_energy(bbpd_solver *s, double *density, double *energy)
{
    BBPD_REQUIRE(s && !s->batch, "bbpd_solver_strain_energy takes a single-problem solver");
    return region_damage(s, 0, s->slot_of, nullptr, nullptr, density, energy);
}

extern "C" int bbpd_solver_observables(bbpd_solver *s, bbpd_observables *o)
{
    BBPD_REQUIRE(s && o, "null argument");
    return region_observables(s, 0, o);
}

namespace {

// stencil-layout history <-> host doubles (one per entry)
int fetch_hs(bbpd_solver *s, std::vector<double> &hv)
{
    const long long nE = s->nPadded;
    hv.assign(nE, 0.0);
    if (s->precision == BBPD_PRECISION_MIXED) {
        std::vector<float> f(nE);
        BBPD_HIP(hipMemcpy(f.data(), s->P.hs, sizeof(float) * nE, hipMemcpyDeviceToHost));
        for (long long e = 0; e < nE; e++) hv[e] = (double)f[e];
    } else {
        BBPD_HIP(hipMemcpy(hv.data(), s->P.hs, sizeof(double) * nE, hipMemcpyDeviceToHost));
    }
    return BBPD_OK;
}

int put_hs(bbpd_solver *s, const std::vector<double> &hv)
{
    const long long nE = s->nPadded;
    const std::vector<unsigned long long> hw = history_words(hv.data(), s->nSlices, (int)(nE / s->nSlices / SLICE));
    BBPD_HIP(hipMemcpy(s->P.snz, hw.data(), sizeof(unsigned long long) * hw.size(), hipMemcpyHostToDevice));
    if (s->precision == BBPD_PRECISION_MIXED) {
        std::vector<float> f(nE);
        for (long long e = 0; e < nE; e++) f[e] = (float)hv[e];
        BBPD_HIP(hipMemcpy(s->P.hs, f.data(), sizeof(float) * nE, hipMemcpyHostToDevice));
    } else {
        BBPD_HIP(hipMemcpy(s->P.hs, hv.data(), sizeof(double) * nE, hipMemcpyHostToDevice));
    }
    return BBPD_OK;
}

// device nodal state <-> reference layout (nNodes x 3, column-major) on the host
int fetch_soa(bbpd_solver *s, double *const *comps, double *dst)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<double> buf((size_t)nS * 3);
    for (int dd = 0; dd < 3; dd++)
        BBPD_HIP(hipMemcpy(buf.data() + (size_t)nS * dd, comps[dd], sizeof(double) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++)
        for (int dd = 0; dd < 3; dd++) dst[n + (size_t)nN * dd] = buf[s->slot_of[n] + (size_t)nS * dd];
    return BBPD_OK;
}

int put_soa(bbpd_solver *s, double *const *comps, const double *src)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<double> buf((size_t)nS * 3);
    for (int dd = 0; dd < 3; dd++)
        BBPD_HIP(hipMemcpy(buf.data() + (size_t)nS * dd, comps[dd], sizeof(double) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++)
        for (int dd = 0; dd < 3; dd++) buf[s->slot_of[n] + (size_t)nS * dd] = src[n + (size_t)nN * dd];
    for (int dd = 0; dd < 3; dd++)
        BBPD_HIP(hipMemcpy(comps[dd], buf.data() + (size_t)nS * dd, sizeof(double) * nS, hipMemcpyHostToDevice));
    return BBPD_OK;
}

int fetch_pos(bbpd_solver *s, int buf_id, double *dst)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<NodeRec> r(nS);
    BBPD_HIP(hipMemcpy(r.data(), s->P.pos[buf_id], sizeof(NodeRec) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++) {
        const NodeRec &q = r[s->slot_of[n]];
        dst[n] = q.x;
        dst[n + (size_t)nN] = q.y;
        dst[n + 2 * (size_t)nN] = q.z;
    }
    return BBPD_OK;
}

int put_pos(bbpd_solver *s, int buf_id, const double *src)
{
    const int nN = s->nN, nS = s->nSlots;
    std::vector<NodeRec> r(nS);
    BBPD_HIP(hipMemcpy(r.data(), s->P.pos[buf_id], sizeof(NodeRec) * nS, hipMemcpyDeviceToHost));
    for (int n = 0; n < nN; n++) {
        NodeRec &q = r[s->slot_of[n]];
        q.x = src[n];
        q.y = src[n + (size_t)nN];
        q.z = src[n + 2 * (size_t)nN];
    }
    BBPD_HIP(hipMemcpy(s->P.pos[buf_id], r.data(), sizeof(NodeRec) * nS, hipMemcpyHostToDevice));
    return BBPD_OK;
}

// steps completed by a region
long long current_step(bbpd_solver *s, int *rc, int ri = 0)
{
    RegionState d;
    hipError_t e = hipMemcpy(&d, s->d_rst + ri, sizeof(d), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        *rc = bbpd::hip_fail(e, "hipMemcpy (step counter)");
        return 0;
    }
    *rc = BBPD_OK;
    return d.step;
}

// the position buffer the next step reads (launch parity)
int current_buffer(bbpd_solver *s, int *rc)
{
    long long t = 0;
    hipError_t e = hipMemcpy(&t, s->d_tick, sizeof(t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) {
        *rc = bbpd::hip_fail(e, "hipMemcpy (launch counter)");
        return 0;
    }
    *rc = BBPD_OK;
    return (int)(t & 1);
}

// The force arrays as seen between launches, `cur` = current_buffer: F = the last step's force and
// Fp = the one before.  Under ADR the two sets alternate by launch parity (the launch with parity c
// writes set c), so F is set cur ^ 1 and Fp set cur; Euler-Cromer keeps F in set 0.
void force_sets(const bbpd_solver *s, int cur, double **F, double **Fp)
{
    double *const set[2][3] = {{s->P.F0, s->P.F1, s->P.F2}, {s->P.Fp0, s->P.Fp1, s->P.Fp2}};
    const int f = s->P.adr ? (cur ^ 1) : 0;
    for (int d = 0; d < 3; d++) {
        F[d] = set[f][d];
        Fp[d] = set[f ^ 1][d];
    }
}

} // namespace

extern "C" int bbpd_solver_get_nodal(bbpd_solver *s, double *u, double *v, double *F, double *x)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    int rc = 0;
    const int cur = current_buffer(s, &rc);
    if (rc) return rc;
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    double *uc[3] = {s->P.u0, s->P.u1, s->P.u2};
    double *vc[3] = {s->P.v0, s->P.v1, s->P.v2};
    double *fc[3], *fpc[3];
    force_sets(s, cur, fc, fpc);
    if (u && (rc = mixed ? fetch_pos(s, cur, u) : fetch_soa(s, uc, u))) return rc;
    if (v && (rc = fetch_soa(s, vc, v))) return rc;
    if (F && (rc = fetch_soa(s, fc, F))) return rc;
    if (x) {
        if (!mixed) {
            if ((rc = fetch_pos(s, cur, x))) return rc;
        } else {
            std::vector<double> uu((size_t)s->nN * 3), XX((size_t)s->nN * 3);
            double *Xc[3] = {(double *)s->P.X0, (double *)s->P.X1, (double *)s->P.X2};
            if ((rc = fetch_pos(s, cur, uu.data()))) return rc;
            if ((rc = fetch_soa(s, Xc, XX.data()))) return rc;
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
    if (s->stencil) {
        std::vector<double> hv;
        int rc = fetch_hs(s, hv);
        if (rc) return rc;
        const bool linear = s->law == BBPD_LAW_LINEAR;
        for (int k = 0; k < s->nB; k++) {
            const double a = hv[s->entry_i[k]], b = hv[s->entry_j[k]];
            if (!bits_equal(a, b)) {
                bbpd::set_error("internal: the two ends of bond " + std::to_string(k + 1) + " disagree");
                return BBPD_ERR_UNSUPPORTED;
            }
            const bool sgn = std::signbit(a);
            if (H) H[k] = std::fabs(a);
            if (flag) flag[k] = (!linear && sgn) ? 1.0 : 0.0;
            if (failout) failout[k] = (linear && sgn) ? 0.0 : 1.0;
        }
        return BBPD_OK;
    }
    std::vector<uint4> r(nE);
    BBPD_HIP(hipMemcpy(r.data(), s->P.rec, sizeof(uint4) * nE, hipMemcpyDeviceToHost));
    auto hist_of = [&](long long e) {
        if (mixed) {
            float f;
            std::memcpy(&f, &r[e].z, 4);
            return (double)f;
        }
        return bits_to_double(r[e].z, r[e].w);
    };
    for (int k = 0; k < s->nB; k++) {
        const long long e = s->entry_i[k], e2 = s->entry_j[k];
        if (H) {
            const double a = hist_of(e), b = hist_of(e2);
            if (!bits_equal(a, b)) {
                bbpd::set_error("internal: the two ends of bond " + std::to_string(k + 1) + " disagree");
                return BBPD_ERR_UNSUPPORTED;
            }
            H[k] = a;
        }
        if (flag) flag[k] = (r[e].x & FLAG_BIT) ? 1.0 : 0.0;
        if (failout) failout[k] = (r[e].x & FAILED_BIT) ? 0.0 : 1.0;
    }
    return BBPD_OK;
}

extern "C" int bbpd_solver_checkpoint(bbpd_solver *s, int64_t *step, double *u, double *v, double *x, double *F,
                                      double *Fp, double *H, double *flag, double *fail)
{
    BBPD_REQUIRE(s, "null solver");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    int rc = 0;
    const long long st = current_step(s, &rc);
    if (rc) return rc;
    if (step) *step = st;
    if ((rc = bbpd_solver_get_nodal(s, u, v, F, x))) return rc;
    if (Fp) {
        // at a step boundary the reference's ADR hand-over has made nodalForcePrevious equal to
        // nodalForce (dynamicsolverdisplacementcontrolledADR.m:153-155): the last step's force, which
        // is the next step's previous force; Euler-Cromer keeps none
        if (s->P.Fp0) {
            if (F) std::memcpy(Fp, F, sizeof(double) * 3 * (size_t)s->nN);
            else if ((rc = bbpd_solver_get_nodal(s, nullptr, nullptr, Fp, nullptr))) return rc;
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
    int rc = 0;
    const int cur = current_buffer(s, &rc);
    if (rc) return rc;
    double *uc[3] = {s->P.u0, s->P.u1, s->P.u2};
    double *vc[3] = {s->P.v0, s->P.v1, s->P.v2};
    double *fc[3], *fpc[3];
    force_sets(s, cur, fc, fpc);
    if (u && (rc = mixed ? put_pos(s, cur, u) : put_soa(s, uc, u))) return rc;
    if (!mixed && x && (rc = put_pos(s, cur, x))) return rc;
    if (v && (rc = put_soa(s, vc, v))) return rc;
    if (F && (rc = put_soa(s, fc, F))) return rc;
    if (Fp && s->P.Fp0 && (rc = put_soa(s, fpc, Fp))) return rc;
    const long long nE = s->nPadded;
    if (H)
        for (int k = 0; k < s->nB; k++)
            if (!(H[k] >= 0 && H[k] <= 1)) {
                bbpd::set_error("bondSofteningFactor must lie in [0, 1]");
                return BBPD_ERR_ARG;
            }
    if (s->stencil && (H || flag || fail)) {
        const bool linear = s->law == BBPD_LAW_LINEAR;
        if (fail && !linear)
            for (int k = 0; k < s->nB; k++)
                if (fail[k] != 1) {
                    bbpd::set_error("fail other than 1 with a softening law needs the record layout "
                                    "(BBPD_LAYOUT=records)");
                    return BBPD_ERR_UNSUPPORTED;
                }
        std::vector<double> hv;
        if ((rc = fetch_hs(s, hv))) return rc;
        for (int k = 0; k < s->nB; k++) {
            const double was = hv[s->entry_i[k]];
            const double mag = H ? H[k] : std::fabs(was);
            bool sgn = std::signbit(was);
            if (!linear && flag) sgn = flag[k] == 1;
            if (linear && fail) sgn = fail[k] == 0;
            const double hn = sgn ? -mag : mag;
            hv[s->entry_i[k]] = hv[s->entry_j[k]] = hn;
        }
        if ((rc = put_hs(s, hv))) return rc;
    } else if (H || flag || fail) {
        std::vector<uint4> r(nE);
        BBPD_HIP(hipMemcpy(r.data(), s->P.rec, sizeof(uint4) * nE, hipMemcpyDeviceToHost));
        for (int k = 0; k < s->nB; k++) {
            for (long long e : {s->entry_i[k], s->entry_j[k]}) {
                if (H) {
                    if (mixed) {
                        const float f = (float)H[k];
                        std::memcpy(&r[e].z, &f, 4);
                    } else {
                        double_to_bits(H[k], r[e].z, r[e].w);
                    }
                }
                if (flag) r[e].x = (r[e].x & ~FLAG_BIT) | (flag[k] == 1 ? FLAG_BIT : 0u);
                if (fail) r[e].x = (r[e].x & ~FAILED_BIT) | (fail[k] == 0 ? FAILED_BIT : 0u);
            }
        }
        BBPD_HIP(hipMemcpy(s->P.rec, r.data(), sizeof(uint4) * nE, hipMemcpyHostToDevice));
    }
    RegionState d;
    BBPD_HIP(hipMemcpy(&d, s->d_rst, sizeof(d), hipMemcpyDeviceToHost));
    d.step = step;
    d.done = 0u;
    BBPD_HIP(hipMemcpy(s->d_rst, &d, sizeof(d), hipMemcpyHostToDevice));
    s->hist_read[0] = step;
    return BBPD_OK;
}

namespace {

int region_history(bbpd_solver *s, int ri, double *out_rows, int cap, int *n_out)
{
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    int rc = 0;
    const long long step = current_step(s, &rc, ri);
    if (rc) return rc;
    const Region &RG = s->reg[ri];
    const long long hc = RG.hist_cap;
    long long from = std::max(s->hist_read[ri], step - hc);
    long long cnt = std::min((long long)cap, step - from);
    // copy only the ring rows [from, from + cnt), in at most two pieces
    std::vector<double> rows((size_t)std::max(cnt, 0LL) * N_HIST);
    long long done = 0;
    while (done < cnt) {
        const long long r0 = (from + done) % hc;
        const long long piece = std::min(cnt - done, hc - r0);
        BBPD_HIP(hipMemcpy(rows.data() + (size_t)done * N_HIST, s->d_hist + (size_t)(RG.hist0 + r0) * N_HIST,
                           sizeof(double) * N_HIST * (size_t)piece, hipMemcpyDeviceToHost));
        done += piece;
    }
    if (out_rows && cnt > 0) std::memcpy(out_rows, rows.data(), sizeof(double) * rows.size());
    s->hist_read[ri] = from + cnt;
    *n_out = (int)cnt;
    return BBPD_OK;
}

} // namespace

extern "C" int bbpd_solver_history(bbpd_solver *s, double *out_rows, int cap, int *n_out)
{
    BBPD_REQUIRE(s && n_out && cap >= 0, "bad arguments");
    return region_history(s, 0, out_rows, cap, n_out);
}

extern "C" int bbpd_solver_get_load_state(bbpd_solver *s, int64_t *inc, int64_t *iters)
{
    BBPD_REQUIRE(s && inc && iters, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    RegionState d;
    BBPD_HIP(hipMemcpyAsync(&d, s->d_rst, sizeof(d), hipMemcpyDeviceToHost, s->stream));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const Region &RG = s->reg[0];
    if (RG.schedule == 2) {
        *inc = d.inc;
        *iters = d.inc_iters;
    } else if (RG.schedule == 1) {
        *inc = d.step / RG.ipi + 1;
        *iters = d.step % RG.ipi;
    } else {
        *inc = 1;
        *iters = d.step;
    }
    return BBPD_OK;
}

extern "C" int bbpd_solver_set_load_state(bbpd_solver *s, int64_t inc, int64_t iters)
{
    BBPD_REQUIRE(s && inc >= 1 && iters >= 0, "bad arguments");
    BBPD_REQUIRE(s->reg[0].schedule == 2, "the load state is set for the equilibrium staircase only");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    RegionState d;
    BBPD_HIP(hipMemcpy(&d, s->d_rst, sizeof(d), hipMemcpyDeviceToHost));
    d.inc = inc;
    d.inc_iters = iters;
    BBPD_HIP(hipMemcpy(s->d_rst, &d, sizeof(d), hipMemcpyHostToDevice));
    return BBPD_OK;
}

extern "C" int bbpd_solver_info(bbpd_solver *s, int64_t *device_bytes, int64_t *bond_bytes, int64_t *n_entries,
                                int64_t *n_padded)
{
    BBPD_REQUIRE(s, "null solver");
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int64_t w = mixed ? 4 : 8;
    // algorithmic bytes of one bond-kernel launch as laid out here: per entry the 16-byte record
    // plus any streamed per-bond quantity; per node the state the kernel reads and writes
    int64_t per_entry = (s->stencil ? w : 16) + (s->P.cmode == PM_STREAM && !mixed ? 8 : 0) +
                        (s->P.s0mode == PM_STREAM ? w : 0) + (s->P.scmode == PM_STREAM ? w : 0);
    int64_t per_node;
    if (s->integrator == BBPD_INTEGRATOR_ADR)
        per_node = 32 /*pos r*/ + 24 /*X*/ + 24 /*F w*/ + 24 /*Fp*/ + 24 /*v*/ + (mixed ? 0 : 24) /*u*/ + 8 /*m*/ +
                   4 /*rowlen*/;
    else
        per_node = 32 /*pos r*/ + 24 /*X*/ + 24 /*pos w*/ + 48 /*v rw*/ + (mixed ? 0 : 48) /*u rw*/ + 24 /*F w*/ +
                   8 /*rho*/ + 1 /*cons*/ + 4 /*rowlen*/;
    if (!s->lattice) per_node += 32;
    // stencil layout: a 16-byte bond mask instead of X and the row length
    if (s->stencil) per_node += 16 - 24 - 4;
    if (device_bytes) *device_bytes = (int64_t)s->device_bytes;
    if (bond_bytes) *bond_bytes = per_entry * s->nEntries + per_node * s->nN;
    if (n_entries) *n_entries = s->nEntries;
    if (n_padded) *n_padded = s->nPadded;
    return BBPD_OK;
}

extern "C" int bbpd_solver_bond_bytes(bbpd_solver *s, int64_t *bytes, int64_t *offsets_set, int64_t *offsets_total)
{
    BBPD_REQUIRE(s && bytes, "bad arguments");
    int64_t all = 0;
    int rc = bbpd_solver_info(s, nullptr, &all, nullptr, nullptr);
    if (rc) return rc;
    if (!s->stencil) {
        *bytes = all;
        if (offsets_set) *offsets_set = 0;
        if (offsets_total) *offsets_total = 0;
        return BBPD_OK;
    }
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int64_t w = mixed ? 4 : 8;
    const int64_t streams = (s->P.s0mode == PM_STREAM ? 1 : 0) + (s->P.scmode == PM_STREAM ? 1 : 0);
    // the model of bbpd_solver_info counts a history word (and each streamed parameter) for every
    // entry; the kernel reads a history word only where its slice's history word for the offset is
    // set, 64 lanes at a time, and a streamed parameter only for an entry whose law runs, which such
    // an offset holds once it has run
    const int64_t nsl = (int64_t)s->nSlots / SLICE;
    std::vector<unsigned long long> hw((size_t)nsl * 2);
    BBPD_HIP(hipMemcpy(hw.data(), s->P.snz, sizeof(unsigned long long) * hw.size(), hipMemcpyDeviceToHost));
    int64_t set = 0, total = 0;
    for (size_t r = 0; r < s->reg.size(); r++) {
        if (r >= s->active.size() || !s->active[r]) continue;
        const Region &RG = s->reg[r];
        const int64_t sl0 = RG.base / SLICE, sl1 = (RG.base + ((int64_t)RG.nN + SLICE - 1) / SLICE * SLICE) / SLICE;
        for (int64_t q = sl0; q < sl1; q++) set += __builtin_popcountll(hw[2 * q]) + __builtin_popcountll(hw[2 * q + 1]);
        total += (sl1 - sl0) * RG.noff;
    }
    *bytes = all - (w + streams * w) * s->nEntries + set * SLICE * (w + streams * w);
    if (offsets_set) *offsets_set = set;
    if (offsets_total) *offsets_total = total;
    return BBPD_OK;
}

extern "C" const char *bbpd_solver_layout(bbpd_solver *s)
{
    return s ? s->layout.c_str() : "";
}

extern "C" int bbpd_solver_autotune(bbpd_solver *s, int reps, double *ms, int *choice)
{
    BBPD_REQUIRE(s && reps > 0, "bad arguments");
    if (choice) *choice = s->occ;
    if (!s->stencil || s->nactive == 0) return BBPD_OK; // one variant only
    BBPD_REQUIRE(s->integrator == BBPD_INTEGRATOR_ADR,
                 "autotuning reruns the bond kernel on the current state, which only the ADR step leaves unchanged");
    BBPD_HIP(hipSetDevice(s->device));
    // under ADR the bond kernel writes F, the Rayleigh partials and max(H, bsf(s)) from positions it
    // does not change: rerunning it on the same state writes the same values
    hipEvent_t e0, e1;
    BBPD_HIP(hipEventCreate(&e0));
    BBPD_HIP(hipEventCreate(&e1));
    double best = 1e300;
    int bestv = s->occ;
    for (int v = 0; v < N_OCC; v++) {
        KernelFn kb = pick_bond_kernel(s, v);
        if (!s->tuned) launch_bond(s, kb, v); // first run of the variant's code
        BBPD_HIP(hipGetLastError());
        BBPD_HIP(hipEventRecord(e0, s->stream));
        for (int r = 0; r < reps; r++) launch_bond(s, kb, v);
        BBPD_HIP(hipEventRecord(e1, s->stream));
        BBPD_HIP(hipEventSynchronize(e1));
        float t = 0;
        BBPD_HIP(hipEventElapsedTime(&t, e0, e1));
        const double avg = (double)t / reps;
        if (ms) ms[v] = avg;
        if (avg < best) {
            best = avg;
            bestv = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (bestv != s->occ) s->version++; // captured graphs hold the old kernel
    s->occ = bestv;
    s->tuned = true;
    if (choice) *choice = bestv;
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

// ================================================================================================
// Batches: a pool of regions in one set of device arrays, stepped by one launch of each kernel.
// Every region has fixed ranges in the shared pools (offsets, key tables, bodies, families, history
// rows) by its index, a block-aligned slot range from a first-fit allocator, and entries at
// ST_MAX per slot from its first slot.
// ================================================================================================
namespace {

constexpr int RB_TAB = 3 * ST_GMAX * ST_GMAX; // key-table entries per region: c, s0, sc
constexpr int RB_FAM = 1 << 16;               // contact-family members per region

// block -> region, launch list (active blocks, padded to a multiple of 8), step-end jobs, active list
int batch_rebuild(bbpd_solver *s)
{
    const int nblk = s->nSlots / BLOCK;
    std::vector<int> blkreg(nblk, 0), grid, act;
    std::vector<int2> se;
    size_t maxtab = 0;
    for (int r = 0; r < s->maxRegions; r++) {
        if (!s->active[r]) continue;
        const Region &RG = s->reg[r];
        for (int b = 0; b < RG.nblk; b++) {
            blkreg[RG.blk0 + b] = r;
            grid.push_back(RG.blk0 + b);
        }
        if (RG.nBodies == 0) se.push_back(make_int2(r, -1));
        else if (RG.seq_contact) se.push_back(make_int2(r, -2));
        else
            for (int b = 0; b < RG.nBodies; b++) se.push_back(make_int2(r, b));
        act.push_back(r);
        maxtab = std::max(maxtab, (size_t)RG.ntab);
    }
    while (grid.size() % 8) grid.push_back(-1);
    long long nn = 0, ne = 0;
    for (int r = 0; r < s->maxRegions; r++)
        if (s->active[r]) {
            nn += s->rh[r].nN;
            ne += 2LL * s->rh[r].nB;
        }
    s->nN = (int)nn;
    s->nEntries = ne;
    BBPD_HIP(hipStreamSynchronize(s->stream));
    BBPD_HIP(hipMemcpy(s->d_blkreg, blkreg.data(), sizeof(int) * nblk, hipMemcpyHostToDevice));
    if (!grid.empty()) BBPD_HIP(hipMemcpy(s->d_grid, grid.data(), sizeof(int) * grid.size(), hipMemcpyHostToDevice));
    const std::vector<int> gst = stripe_list(grid);
    if (!gst.empty()) BBPD_HIP(hipMemcpy(s->d_grid_st, gst.data(), sizeof(int) * gst.size(), hipMemcpyHostToDevice));
    if (!se.empty()) BBPD_HIP(hipMemcpy(s->d_se, se.data(), sizeof(int2) * se.size(), hipMemcpyHostToDevice));
    if (!act.empty()) BBPD_HIP(hipMemcpy(s->d_active, act.data(), sizeof(int) * act.size(), hipMemcpyHostToDevice));
    s->grid_n = (int)grid.size();
    s->grid_n_st = (int)gst.size();
    s->nse = (int)se.size();
    s->nactive = (int)act.size();
    s->P.nse = s->nse;
    s->lds_bytes = maxtab * (s->precision == BBPD_PRECISION_MIXED ? sizeof(float) : sizeof(double));
    s->version++;
    return allow_lds(s);
}

bool same_bits(double a, double b) { return bits_equal(a, b); }

} // namespace

extern "C" int bbpd_batch_create(const bbpd_batch_spec *spec, int precision, int device, bbpd_solver **out)
{
    BBPD_REQUIRE(spec && out && spec->shape, "null argument");
    *out = nullptr;
    const bbpd_problem *sh = spec->shape;
    BBPD_REQUIRE(spec->maxRegions > 0 && spec->maxSlots > 0 && spec->DX > 0, "bad batch capacity");
    BBPD_REQUIRE(precision == BBPD_PRECISION_FP64 || precision == BBPD_PRECISION_MIXED, "unknown precision");
    BBPD_REQUIRE(sh->law >= BBPD_LAW_LINEAR && sh->law <= BBPD_LAW_EXPONENTIAL, "unknown law");
    BBPD_REQUIRE(sh->integrator == BBPD_INTEGRATOR_EULER_CROMER || sh->integrator == BBPD_INTEGRATOR_ADR,
                 "unknown integrator");
    BBPD_REQUIRE(sh->DT > 0, "DT must be positive");
    int ndev = 0;
    BBPD_HIP(hipGetDeviceCount(&ndev));
    BBPD_REQUIRE(device >= 0 && device < ndev, "device ordinal out of range");
    BBPD_HIP(hipSetDevice(device));
    auto *s = new bbpd_solver();
    auto fail = [&](int rc) {
        bbpd_solver_destroy(s);
        return rc;
    };
    const bool mixed = precision == BBPD_PRECISION_MIXED;
    s->device = device;
    s->precision = precision;
    s->law = sh->law;
    s->integrator = sh->integrator;
    s->batch = true;
    s->stencil = true;
    if (env_launch_variant() >= 0) s->occ = env_launch_variant();
    s->lattice = true;
    s->layout = "stencil";
    s->maxRegions = spec->maxRegions;
    s->histRows = spec->historyRows > 0 ? spec->historyRows : 65536;
    s->shape = *sh;
    s->DX = spec->DX;
    const long long nSlots = (spec->maxSlots + SLOT_ALIGN - 1) / SLOT_ALIGN * SLOT_ALIGN;
    BBPD_REQUIRE(nSlots < (1LL << 31) / ST_MAX, "batch too large");
    s->nSlots = (int)nSlots;
    s->nSlices = s->nSlots / SLICE;
    const int nblk = s->nSlots / BLOCK;
    const long long nE = nSlots * ST_MAX;
    s->nPadded = nE;
    s->freeblk.assign(nblk, 1);
    s->reg.assign(s->maxRegions, Region{});
    s->active.assign(s->maxRegions, 0);
    s->hist_read.assign(s->maxRegions, 0);
    s->rh.assign(s->maxRegions, bbpd_solver::RegionHost{});

    BBPD_HIP(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    Params &P = s->P;
    int rc = 0;
    const size_t nS = (size_t)nSlots;
    const size_t ew = mixed ? 4 : 8;
    double *Xd, *ud = nullptr, *vd, *Fd, *Fpd = nullptr, *rhod, *massd = nullptr;
    NodeRec *posd[2];
    unsigned char *consd;
    uint4 *maskd;
    unsigned long long *maskTd, *snzd;
    void *hsd, *s0d = nullptr, *scd = nullptr;
    int4 *stoffd;
    int2 *srowd;
    StOff *sstd;
    float *tabfd = nullptr;
    double *tabd, *partd;
    if ((rc = dalloc(&Xd, 3 * nS, s->owned))) return fail(rc);
    for (int q = 0; q < 2; q++)
        if ((rc = dalloc(&posd[q], nS, s->owned))) return fail(rc);
    if (!mixed && (rc = dalloc(&ud, 3 * nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&vd, 3 * nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&Fd, 3 * nS, s->owned))) return fail(rc);
    if (sh->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&Fpd, 3 * nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&rhod, nS, s->owned))) return fail(rc);
    if (sh->integrator == BBPD_INTEGRATOR_ADR && (rc = dalloc(&massd, nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&consd, nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&maskd, nS, s->owned))) return fail(rc);
    if ((rc = dalloc(&maskTd, nS / SLICE * ST_MAX + ST_PAD, s->owned))) return fail(rc);
    if ((rc = dalloc(&snzd, nS / SLICE * 2, s->owned))) return fail(rc);
    if ((rc = dalloc((char **)&hsd, (size_t)(nE + 2 * SLICE) * ew, s->owned))) return fail(rc);
    if (sh->s0 && (rc = dalloc((char **)&s0d, (size_t)(nE + 2 * SLICE) * ew, s->owned))) return fail(rc);
    if (sh->sc && (rc = dalloc((char **)&scd, (size_t)(nE + 2 * SLICE) * ew, s->owned))) return fail(rc);
    if ((rc = dalloc(&stoffd, (size_t)s->maxRegions * ST_MAX, s->owned))) return fail(rc);
    if ((rc = dalloc(&srowd, (size_t)s->maxRegions * (ST_MAX + ST_PAD), s->owned))) return fail(rc);
    if ((rc = dalloc(&sstd, (size_t)s->maxRegions * (ST_MAX + ST_PAD), s->owned))) return fail(rc);
    if ((rc = dalloc(&tabd, (size_t)s->maxRegions * RB_TAB, s->owned))) return fail(rc);
    if (mixed && (rc = dalloc(&tabfd, (size_t)s->maxRegions * RB_TAB, s->owned))) return fail(rc);
    if ((rc = dalloc(&partd, 4 * (size_t)nblk + 4, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->counts, 4, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_hist, (size_t)s->maxRegions * s->histRows * N_HIST, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_reg, s->maxRegions, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_rst, s->maxRegions, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_blkreg, nblk, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_grid, nblk + 8, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_grid_st, nblk + 8 + 8 * ST_STRIPE, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_active, s->maxRegions, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_se, (size_t)s->maxRegions * BBPD_MAX_BODIES, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_bodies, (size_t)s->maxRegions * BBPD_MAX_BODIES, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_fampool, (size_t)s->maxRegions * RB_FAM, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_tick, 1, s->owned))) return fail(rc);
    if ((rc = dalloc(&s->d_tick_done, 1, s->owned))) return fail(rc);
    {
        std::vector<RegionState> z(s->maxRegions);
        std::memset(z.data(), 0, sizeof(RegionState) * z.size());
        BBPD_HIP(hipMemcpy(s->d_rst, z.data(), sizeof(RegionState) * z.size(), hipMemcpyHostToDevice));
        long long t0 = 0;
        unsigned u0 = 0;
        BBPD_HIP(hipMemcpy(s->d_tick, &t0, sizeof(t0), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(s->d_tick_done, &u0, sizeof(u0), hipMemcpyHostToDevice));
        std::vector<Region> zr(s->maxRegions);
        std::memset(zr.data(), 0, sizeof(Region) * zr.size());
        BBPD_HIP(hipMemcpy(s->d_reg, zr.data(), sizeof(Region) * zr.size(), hipMemcpyHostToDevice));
    }
    auto comp = [&](double *base, int d) { return base ? base + nS * d : nullptr; };
    P.nSlots = s->nSlots;
    P.X0 = comp(Xd, 0);
    P.X1 = comp(Xd, 1);
    P.X2 = comp(Xd, 2);
    P.pos[0] = posd[0];
    P.pos[1] = posd[1];
    P.Xrec = nullptr;
    P.hx = P.hy = P.hz = spec->DX;
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
    P.tab = tabd;
    P.tabf = tabfd;
    P.cmode = PM_TABLE;
    P.s0mode = sh->s0 ? PM_STREAM : PM_SCALAR;
    P.scmode = sh->sc ? PM_STREAM : PM_SCALAR;
    P.s0str = s0d;
    P.scstr = scd;
    P.adr = sh->integrator == BBPD_INTEGRATOR_ADR ? 1 : 0;
    P.law_linear = sh->law == BBPD_LAW_LINEAR ? 1 : 0;
    P.V = sh->cellVolume;
    P.horizon = sh->horizon;
    P.hmr = sh->horizon - sh->RADIJ;
    P.hpr = sh->horizon + sh->RADIJ;
    P.two_r = 2 * sh->RADIJ;
    P.rtwo_r = 1.0 / P.two_r;
    P.k = sh->k;
    P.alpha = sh->alpha;
    P.denom_exp = 1 - std::exp(-sh->k);
    P.sc_steel = sh->sc_steel;
    P.ff = sh->failureFunctionality;
    P.DT = sh->DT;
    P.D = sh->DAMPING;
    P.partials = partd;
    P.hist = s->d_hist;
    P.reg = s->d_reg;
    P.rst = s->d_rst;
    P.blkreg = s->d_blkreg;
    P.grid = s->d_grid;
    P.active = s->d_active;
    P.se = s->d_se;
    P.nse = 0;
    P.bodies = s->d_bodies;
    P.tick = s->d_tick;
    P.tick_done = s->d_tick_done;
    P.stencil = 1;
    P.soff = stoffd;
    P.srow = srowd;
    P.sst = sstd;
    P.smask = maskd;
    P.smaskT = maskTd;
    P.hs = hsd;
    P.snz = snzd;
    s->device_bytes = nS * (2 * sizeof(NodeRec) + 8 * (3 * (mixed ? 3 : 4) + 2 + (Fpd ? 3 : 0)) + 1 + 16) +
                      (size_t)nE * ew * (1 + (s0d ? 1 : 0) + (scd ? 1 : 0));
    s->ev.resize(2048);
    for (auto &e : s->ev) BBPD_HIP(hipEventCreate(&e));
    if ((rc = batch_rebuild(s))) return fail(rc);
    BBPD_HIP(hipDeviceSynchronize());
    *out = s;
    return BBPD_OK;
}

extern "C" int bbpd_batch_admit(bbpd_solver *s, const bbpd_problem *pr, double stop_frac, double stop_disp,
                                int *region)
{
    BBPD_REQUIRE(s && s->batch && pr && region, "bad arguments");
    *region = -1;
    const bbpd_problem &sh = s->shape;
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const int nN = pr->nNodes, nB = pr->nBonds;
    BBPD_REQUIRE(nN > 0 && nB >= 0, "need nNodes > 0 and nBonds >= 0");
    BBPD_REQUIRE(pr->undeformedCoordinates && pr->BONDLIST && pr->UNDEFORMEDLENGTH && pr->VOLUMECORRECTIONFACTORS &&
                     pr->BONDSTIFFNESS && pr->DENSITY,
                 "missing input array");
    BBPD_REQUIRE(pr->law == sh.law && pr->integrator == sh.integrator, "law or integrator differs from the batch");
    BBPD_REQUIRE(same_bits(pr->cellVolume, sh.cellVolume) && same_bits(pr->horizon, sh.horizon) &&
                     same_bits(pr->RADIJ, sh.RADIJ) && same_bits(pr->k, sh.k) && same_bits(pr->alpha, sh.alpha) &&
                     same_bits(pr->sc_steel, sh.sc_steel) &&
                     same_bits(pr->failureFunctionality, sh.failureFunctionality) && same_bits(pr->DT, sh.DT) &&
                     same_bits(pr->DAMPING, sh.DAMPING),
                 "a shared scalar (cellVolume, horizon, RADIJ, k, alpha, sc_steel, failureFunctionality, DT, "
                 "DAMPING) differs from the batch");
    BBPD_REQUIRE((pr->s0 != nullptr) == (sh.s0 != nullptr) && (pr->sc != nullptr) == (sh.sc != nullptr),
                 "s0 / sc must be per bond exactly when the batch's are");
    BBPD_REQUIRE(pr->integrator != BBPD_INTEGRATOR_ADR || pr->massVector, "ADR needs massVector");
    BBPD_REQUIRE(pr->nBodies >= 0 && pr->nBodies <= BBPD_MAX_BODIES, "too many contact bodies");
    BBPD_REQUIRE(pr->schedule >= 0 && pr->schedule <= 2, "unknown load schedule");
    BBPD_REQUIRE(pr->schedule == 0 ? pr->nTimeSteps > 1 : pr->itersPerIncrement > 0, "bad load schedule");
    BBPD_REQUIRE(pr->schedule != 2 || (pr->integrator == BBPD_INTEGRATOR_ADR && pr->eqTol >= 0 &&
                                       pr->itersPerIncrementMin >= 0),
                 "the equilibrium staircase needs ADR, eqTol >= 0 and itersPerIncrementMin >= 0");
    BBPD_HIP(hipSetDevice(s->device));
    int rc;
    std::vector<int> bi, bj;
    if ((rc = check_bonds(pr, bi, bj))) return rc;
    double xmin[3], xmax[3], hq, hlat[3];
    if (!detect_lattice(pr, hlat, xmin, xmax, hq) || !same_bits(hlat[0], s->DX) || !same_bits(hlat[1], s->DX) ||
        !same_bits(hlat[2], s->DX)) {
        bbpd::set_error("a batch problem must be a lattice of spacing DX (X == DX * k)");
        return BBPD_ERR_UNSUPPORTED;
    }
    StencilPlan sp;
    plan_stencil(pr, pr->undeformedCoordinates, bi, bj, hlat, sp);
    if (!sp.ok) {
        bbpd::set_error("a batch problem needs the stencil layout: " + sp.why);
        return BBPD_ERR_UNSUPPORTED;
    }
    int famtot = 0;
    for (int b = 0; b < pr->nBodies; b++) famtot += pr->nFamily[b];
    BBPD_REQUIRE(famtot <= RB_FAM, "contact families too large for a batch region");
    // a problem larger than the whole slot pool can never be admitted: a distinct error, so that a
    // caller does not wait for slots that never free up
    if ((long long)(nN + BLOCK - 1) / BLOCK > (long long)s->freeblk.size()) {
        bbpd::set_error("the problem needs " + std::to_string(((long long)nN + BLOCK - 1) / BLOCK * BLOCK) +
                        " node slots, more than the batch holds (" +
                        std::to_string((long long)s->freeblk.size() * BLOCK) + ")");
        return BBPD_ERR_NOMEM;
    }
    // region and slot range
    int rid = -1;
    for (int r = 0; r < s->maxRegions && rid < 0; r++)
        if (!s->active[r]) rid = r;
    const int need = (nN + BLOCK - 1) / BLOCK;
    int blk0 = -1;
    {
        const int nblk = (int)s->freeblk.size();
        int run = 0;
        for (int b = 0; b < nblk; b++) {
            run = s->freeblk[b] ? run + 1 : 0;
            if (run == need) {
                blk0 = b - need + 1;
                break;
            }
        }
    }
    if (rid < 0 || blk0 < 0) {
        bbpd::set_error(rid < 0 ? "batch full: no free region" : "batch full: no free slot range");
        return BBPD_ERR_FULL;
    }
    const int base = blk0 * BLOCK, nSl = need * BLOCK;
    const size_t nS = (size_t)s->nSlots;
    const long long e0 = (long long)base * ST_MAX;
    const int noff = sp.noff;
    BBPD_HIP(hipStreamSynchronize(s->stream));
    long long tick = 0;
    BBPD_HIP(hipMemcpy(&tick, s->d_tick, sizeof(tick), hipMemcpyDeviceToHost));

    // node data of the region's slots
    std::vector<NodeRec> hpos(nSl);
    std::memset(hpos.data(), 0, sizeof(NodeRec) * nSl);
    std::vector<double> hX(3 * (size_t)nSl, 0.0), zero(3 * (size_t)nSl, 0.0), hrho(nSl, 1.0), hm(nSl, 1.0);
    std::vector<unsigned char> hc(nSl, 0);
    std::vector<uint4> hmask(nSl, make_uint4(0, 0, 0, 0));
    auto &RH = s->rh[rid];
    RH.nN = nN;
    RH.nB = nB;
    RH.slot_of.assign(nN, 0);
    for (int n = 0; n < nN; n++) {
        const int q = sp.slot[n]; // local slot
        RH.slot_of[n] = base + q;
        double xv[3];
        for (int d = 0; d < 3; d++) {
            xv[d] = pr->undeformedCoordinates[n + (size_t)nN * d];
            hX[q + (size_t)nSl * d] = xv[d];
            if (pr->CONSTRAINTFLAG && pr->CONSTRAINTFLAG[n + (size_t)nN * d] == 1) hc[q] |= (unsigned char)(1u << d);
        }
        hpos[q].x = mixed ? 0.0 : xv[0];
        hpos[q].y = mixed ? 0.0 : xv[1];
        hpos[q].z = mixed ? 0.0 : xv[2];
        hpos[q].q = (uint64_t)sp.k[3 * (size_t)n] | ((uint64_t)sp.k[3 * (size_t)n + 1] << LAT_BITS) |
                    ((uint64_t)sp.k[3 * (size_t)n + 2] << (2 * LAT_BITS)) | ((uint64_t)sp.gidx[n] << G_SHIFT);
        hrho[q] = pr->DENSITY[n];
        if (pr->massVector) hm[q] = pr->massVector[n];
    }
    // entries: masks, history (sign bit set for bonds that start failed), streamed parameters
    const long long nEr = (long long)nSl * noff;
    std::vector<double> h0(nEr, 0.0), hs0(pr->s0 ? nEr : 0, 0.0), hsc(pr->sc ? nEr : 0, 0.0);
    RH.entry_i.assign(nB, 0);
    RH.entry_j.assign(nB, 0);
    auto dense_local = [&](int q, int o) { return ((long long)(q / SLICE) * noff + o) * SLICE + (q % SLICE); };
    for (int k = 0; k < nB; k++) {
        const int qi = sp.slot[bi[k]], qj = sp.slot[bj[k]];
        const long long ei = dense_local(qi, sp.oi[k]), ej = dense_local(qj, sp.oj[k]);
        RH.entry_i[k] = e0 + ei;
        RH.entry_j[k] = e0 + ej;
        reinterpret_cast<uint32_t *>(&hmask[qi])[sp.oi[k] >> 5] |= 1u << (sp.oi[k] & 31);
        reinterpret_cast<uint32_t *>(&hmask[qj])[sp.oj[k] >> 5] |= 1u << (sp.oj[k] & 31);
        if (pr->fail && pr->fail[k] == 0) h0[ei] = h0[ej] = -0.0;
        if (pr->s0) hs0[ei] = hs0[ej] = pr->s0[k];
        if (pr->sc) hsc[ei] = hsc[ej] = pr->sc[k];
    }
    // contact bodies (families as absolute slots), contact-member bits
    std::vector<Body> hb(BBPD_MAX_BODIES);
    std::vector<int> fam;
    std::vector<unsigned char> in_family(nN, 0);
    std::vector<int> owner(nN, -1);
    int seq_contact = 0; // families that share members are applied in order by one block
    int *fambase = s->d_fampool + (size_t)rid * RB_FAM;
    for (int bdy = 0; bdy < pr->nBodies; bdy++) {
        const int nf = pr->nFamily[bdy];
        BBPD_REQUIRE(nf >= 0 && (nf == 0 || pr->family[bdy]), "contact family missing");
        Body &B = hb[bdy];
        B.fam = fambase + fam.size();
        B.nFam = nf;
        for (int t = 0; t < nf; t++) {
            const double idd = pr->family[bdy][t];
            const int id = (int)idd;
            if (idd != (double)id || id < 1 || id > nN) {
                bbpd::set_error("contact family holds an id outside 1..nNodes");
                return BBPD_ERR_ARG;
            }
            if (owner[id - 1] == bdy) {
                bbpd::set_error("a contact family lists a node twice");
                return BBPD_ERR_ARG;
            }
            if (owner[id - 1] >= 0) seq_contact = 1;
            owner[id - 1] = bdy;
            in_family[id - 1] = 1;
            fam.push_back(base + sp.slot[id - 1]);
        }
        B.cx = pr->centreX[bdy];
        B.cz = pr->centreZ[bdy];
        B.R = pr->radius[bdy];
        B.moves = pr->moves[bdy];
        B.region = rid;
        B.index = bdy;
    }
    if (pr->integrator == BBPD_INTEGRATOR_ADR)
        for (int n = 0; n < nN; n++)
            if (in_family[n]) hc[sp.slot[n]] |= CONTACT_BIT;
    // per-offset data and key tables
    const double hmr = sh.horizon - sh.RADIJ, hpr = sh.horizon + sh.RADIJ, two_r = 2 * sh.RADIJ;
    const double hdx[3] = {s->DX, s->DX, s->DX};
    int fast64 = 0;
    const std::vector<StOff> hst =
        offset_records(sp.off, hdx, hmr, hpr, sh.horizon, two_r, mixed, sp.kmin, sp.kmax, &fast64);
    std::vector<double> tab(sp.tc);
    tab.insert(tab.end(), sp.ts0.begin(), sp.ts0.end());
    tab.insert(tab.end(), sp.tsc.begin(), sp.tsc.end());
    BBPD_REQUIRE((int)tab.size() <= RB_TAB, "key tables too large for a batch region");
    std::vector<float> tabf(tab.size());
    for (size_t i = 0; i < sp.tc.size(); i++) tabf[i] = (float)(sp.tc[i] * sh.cellVolume);
    for (size_t i = sp.tc.size(); i < tab.size(); i++) tabf[i] = (float)tab[i];

    // descriptor
    Region RG{};
    RG.base = base;
    RG.nN = nN;
    RG.nSlots = nSl;
    RG.blk0 = blk0;
    RG.nblk = need;
    RG.e0 = e0;
    RG.noff = noff;
    RG.off0 = rid * ST_MAX;
    RG.nrow = sp.nrow;
    RG.row0 = rid * (ST_MAX + ST_PAD);
    RG.rx = sp.rx;
    RG.sst0 = rid * (ST_MAX + ST_PAD);
    RG.fast64 = fast64;
    RG.kg = sp.G;
    RG.tab0 = rid * RB_TAB;
    RG.ntab = (int)tab.size();
    RG.ts0 = (int)sp.tc.size();
    RG.tsc = (int)(sp.tc.size() + sp.ts0.size());
    // key-tabled s0 / sc are not used in a batch: per-bond values are streamed
    RG.s0 = pr->s0 ? 0.0 : pr->s0_scalar;
    RG.sc = pr->sc ? 0.0 : pr->sc_scalar;
    RG.s1 = pr->s1;
    RG.s0lo = pr->s0 && nB > 0 ? *std::min_element(pr->s0, pr->s0 + nB) : 0.0;
    RG.sclo = pr->sc && nB > 0 ? *std::min_element(pr->sc, pr->sc + nB) : 0.0;
    RG.body0 = rid * BBPD_MAX_BODIES;
    RG.nBodies = pr->nBodies;
    RG.seq_contact = seq_contact;
    RG.schedule = pr->schedule;
    RG.nTimeSteps = pr->nTimeSteps;
    RG.ipi = pr->itersPerIncrement > 0 ? pr->itersPerIncrement : 1;
    RG.applied = pr->appliedDisplacement;
    RG.increment = pr->increment;
    RG.eq_tol = pr->eqTol;
    RG.ipi_min = pr->itersPerIncrementMin;
    auto slot_or_none = [&](int id1) { return (id1 >= 1 && id1 <= nN) ? base + sp.slot[id1 - 1] : -1; };
    RG.ref_slot = slot_or_none(pr->referenceNode);
    RG.cmod1 = slot_or_none(pr->cmodNode1);
    RG.cmod2 = slot_or_none(pr->cmodNode2);
    RG.hist0 = (long long)rid * s->histRows;
    RG.hist_cap = s->histRows;
    RG.stop_frac = stop_frac;
    RG.stop_disp = stop_disp;

    // upload
    Params &P = s->P;
    for (int q = 0; q < 2; q++) BBPD_HIP(hipMemcpy(P.pos[q] + base, hpos.data(), sizeof(NodeRec) * nSl, hipMemcpyHostToDevice));
    double *Xc[3] = {(double *)P.X0, (double *)P.X1, (double *)P.X2};
    double *uc[3] = {P.u0, P.u1, P.u2}, *vc[3] = {P.v0, P.v1, P.v2}, *fc[3] = {P.F0, P.F1, P.F2};
    double *fpc[3] = {P.Fp0, P.Fp1, P.Fp2};
    for (int d = 0; d < 3; d++) {
        BBPD_HIP(hipMemcpy(Xc[d] + base, hX.data() + (size_t)nSl * d, sizeof(double) * nSl, hipMemcpyHostToDevice));
        if (uc[d]) BBPD_HIP(hipMemcpy(uc[d] + base, zero.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(vc[d] + base, zero.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(fc[d] + base, zero.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
        if (fpc[d]) BBPD_HIP(hipMemcpy(fpc[d] + base, zero.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
    }
    BBPD_HIP(hipMemcpy((double *)P.rho + base, hrho.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
    if (P.mass) BBPD_HIP(hipMemcpy((double *)P.mass + base, hm.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((unsigned char *)P.cons + base, hc.data(), nSl, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((uint4 *)P.smask + base, hmask.data(), sizeof(uint4) * nSl, hipMemcpyHostToDevice));
    {
        const std::vector<unsigned long long> hmt = lane_masks(hmask, nSl, noff);
        BBPD_HIP(hipMemcpy((unsigned long long *)P.smaskT + RG.e0 / SLICE, hmt.data(),
                           sizeof(unsigned long long) * hmt.size(), hipMemcpyHostToDevice));
    }
    auto put_entries = [&](void *dst, const std::vector<double> &v) -> int {
        if (mixed) {
            std::vector<float> f(v.size());
            for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
            BBPD_HIP(hipMemcpy((float *)dst + e0, f.data(), sizeof(float) * f.size(), hipMemcpyHostToDevice));
        } else {
            BBPD_HIP(hipMemcpy((double *)dst + e0, v.data(), sizeof(double) * v.size(), hipMemcpyHostToDevice));
        }
        return BBPD_OK;
    };
    if ((rc = put_entries(P.hs, h0))) return rc;
    {
        const std::vector<unsigned long long> hw = history_words(h0.data(), nSl / SLICE, noff);
        BBPD_HIP(hipMemcpy(P.snz + 2 * (base / SLICE), hw.data(), sizeof(unsigned long long) * hw.size(),
                           hipMemcpyHostToDevice));
    }
    if (pr->s0 && (rc = put_entries((void *)P.s0str, hs0))) return rc;
    if (pr->sc && (rc = put_entries((void *)P.scstr, hsc))) return rc;
    BBPD_HIP(hipMemcpy((int4 *)P.soff + RG.off0, sp.off.data(), sizeof(int4) * noff, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((int2 *)P.srow + RG.row0, sp.rows.data(), sizeof(int2) * sp.rows.size(),
                       hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((StOff *)P.sst + RG.sst0, hst.data(), sizeof(StOff) * hst.size(), hipMemcpyHostToDevice));
    if (!tab.empty()) {
        BBPD_HIP(hipMemcpy((double *)P.tab + RG.tab0, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice));
        if (P.tabf)
            BBPD_HIP(hipMemcpy((float *)P.tabf + RG.tab0, tabf.data(), sizeof(float) * tabf.size(), hipMemcpyHostToDevice));
    }
    if (!fam.empty()) BBPD_HIP(hipMemcpy(fambase, fam.data(), sizeof(int) * fam.size(), hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy(s->d_bodies + RG.body0, hb.data(), sizeof(Body) * BBPD_MAX_BODIES, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy(s->d_reg + rid, &RG, sizeof(Region), hipMemcpyHostToDevice));
    RegionState z;
    std::memset(&z, 0, sizeof(z));
    z.inc = 1;
    BBPD_HIP(hipMemcpy(s->d_rst + rid, &z, sizeof(z), hipMemcpyHostToDevice));
    for (int b = 0; b < need; b++) s->freeblk[blk0 + b] = 0;
    s->reg[rid] = RG;
    s->active[rid] = 1;
    s->hist_read[rid] = 0;
    if ((rc = batch_rebuild(s))) return rc;
    *region = rid;
    return BBPD_OK;
}

extern "C" int bbpd_batch_retire(bbpd_solver *s, int region)
{
    BBPD_REQUIRE(s && s->batch && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const Region &RG = s->reg[region];
    for (int b = 0; b < RG.nblk; b++) s->freeblk[RG.blk0 + b] = 1;
    s->active[region] = 0;
    s->rh[region] = bbpd_solver::RegionHost{};
    return batch_rebuild(s);
}

extern "C" int bbpd_batch_status(bbpd_solver *s, bbpd_region_status *st, int n)
{
    BBPD_REQUIRE(s && s->batch && st && n >= 0 && n <= s->maxRegions, "bad arguments");
    BBPD_HIP(hipSetDevice(s->device));
    std::vector<RegionState> d(std::max(n, 1));
    if (n > 0) {
        BBPD_HIP(hipMemcpyAsync(d.data(), s->d_rst, sizeof(RegionState) * n, hipMemcpyDeviceToHost, s->stream));
        BBPD_HIP(hipStreamSynchronize(s->stream));
    }
    for (int r = 0; r < n; r++) {
        std::memset(&st[r], 0, sizeof(st[r]));
        st[r].active = s->active[r];
        if (!s->active[r]) continue;
        st[r].finished = d[r].finished;
        st[r].step = d[r].step;
        st[r].peak = d[r].peak;
        st[r].reaction_static = d[r].reaction_static;
        st[r].nNodes = s->rh[r].nN;
        st[r].nBonds = s->rh[r].nB;
    }
    return BBPD_OK;
}

extern "C" int bbpd_batch_observables(bbpd_solver *s, int region, bbpd_observables *obs)
{
    BBPD_REQUIRE(s && s->batch && obs && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    return region_observables(s, region, obs);
}

extern "C" int bbpd_batch_history(bbpd_solver *s, int region, double *rows, int cap, int *n_out)
{
    BBPD_REQUIRE(s && s->batch && n_out && cap >= 0 && region >= 0 && region < s->maxRegions && s->active[region],
                 "no such region");
    return region_history(s, region, rows, cap, n_out);
}

extern "C" int bbpd_batch_get_nodal(bbpd_solver *s, int region, double *u, double *v, double *F, double *x)
{
    BBPD_REQUIRE(s && s->batch && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    int rc = 0;
    const int cur = current_buffer(s, &rc);
    if (rc) return rc;
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const Region &RG = s->reg[region];
    const auto &RH = s->rh[region];
    const int nN = RH.nN, nSl = RG.nSlots, base = RG.base;
    auto soa = [&](double *const *comps, double *dst) -> int {
        std::vector<double> buf(nSl);
        for (int d = 0; d < 3; d++) {
            BBPD_HIP(hipMemcpy(buf.data(), comps[d] + base, sizeof(double) * nSl, hipMemcpyDeviceToHost));
            for (int n = 0; n < nN; n++) dst[n + (size_t)nN * d] = buf[RH.slot_of[n] - base];
        }
        return BBPD_OK;
    };
    std::vector<NodeRec> rec(nSl);
    BBPD_HIP(hipMemcpy(rec.data(), s->P.pos[cur] + base, sizeof(NodeRec) * nSl, hipMemcpyDeviceToHost));
    auto posv = [&](int n, int d) {
        const NodeRec &q = rec[RH.slot_of[n] - base];
        return d == 0 ? q.x : (d == 1 ? q.y : q.z);
    };
    double *uc[3] = {s->P.u0, s->P.u1, s->P.u2}, *vc[3] = {s->P.v0, s->P.v1, s->P.v2};
    double *fc[3], *fpc[3];
    force_sets(s, cur, fc, fpc);
    double *Xc[3] = {(double *)s->P.X0, (double *)s->P.X1, (double *)s->P.X2};
    if (u) {
        if (mixed) {
            for (int n = 0; n < nN; n++)
                for (int d = 0; d < 3; d++) u[n + (size_t)nN * d] = posv(n, d);
        } else if ((rc = soa(uc, u))) {
            return rc;
        }
    }
    if (v && (rc = soa(vc, v))) return rc;
    if (F && (rc = soa(fc, F))) return rc;
    if (x) {
        if (!mixed) {
            for (int n = 0; n < nN; n++)
                for (int d = 0; d < 3; d++) x[n + (size_t)nN * d] = posv(n, d);
        } else {
            if ((rc = soa(Xc, x))) return rc;
            for (int n = 0; n < nN; n++)
                for (int d = 0; d < 3; d++) x[n + (size_t)nN * d] += posv(n, d);
        }
    }
    return BBPD_OK;
}

extern "C" int bbpd_batch_get_bonds(bbpd_solver *s, int region, double *H, double *flag, double *failout)
{
    BBPD_REQUIRE(s && s->batch && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    BBPD_HIP(hipSetDevice(s->device));
    BBPD_HIP(hipStreamSynchronize(s->stream));
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    const Region &RG = s->reg[region];
    const auto &RH = s->rh[region];
    const long long e0 = RG.e0, nEr = (long long)RG.nSlots * RG.noff;
    std::vector<double> hv(nEr);
    if (mixed) {
        std::vector<float> f(nEr);
        BBPD_HIP(hipMemcpy(f.data(), (float *)s->P.hs + e0, sizeof(float) * nEr, hipMemcpyDeviceToHost));
        for (long long e = 0; e < nEr; e++) hv[e] = (double)f[e];
    } else {
        BBPD_HIP(hipMemcpy(hv.data(), (double *)s->P.hs + e0, sizeof(double) * nEr, hipMemcpyDeviceToHost));
    }
    const bool linear = s->law == BBPD_LAW_LINEAR;
    for (int k = 0; k < RH.nB; k++) {
        const double a = hv[RH.entry_i[k] - e0], b = hv[RH.entry_j[k] - e0];
        if (!bits_equal(a, b)) {
            bbpd::set_error("internal: the two ends of bond " + std::to_string(k + 1) + " disagree");
            return BBPD_ERR_UNSUPPORTED;
        }
        const bool sgn = std::signbit(a);
        if (H) H[k] = std::fabs(a);
        if (flag) flag[k] = (!linear && sgn) ? 1.0 : 0.0;
        if (failout) failout[k] = (linear && sgn) ? 0.0 : 1.0;
    }
    return BBPD_OK;
}

extern "C" int bbpd_batch_strain_energy(bbpd_solver *s, int region, double *density, double *energy)
{
    BBPD_REQUIRE(s && s->batch && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    return region_damage(s, region, s->rh[region].slot_of, nullptr, nullptr, density, energy);
}

extern "C" int bbpd_batch_damage(bbpd_solver *s, int region, double *damage, double *energy)
{
    BBPD_REQUIRE(s && s->batch && region >= 0 && region < s->maxRegions && s->active[region], "no such region");
    BBPD_REQUIRE(!energy || s->law != BBPD_LAW_LINEAR,
                 "damage energy needs a softening law (the brittle latch keeps no stretch history)");
    return region_damage(s, region, s->rh[region].slot_of, damage, energy);
}
