// bbpd_solver_impl.hpp -- private header of the fused solver (layer 2): device layout types,
// the launch parameters, the solver object, and the functions the solver's translation units
// share (kernels and launches in bbpd_kernels.hip, stencil planning in bbpd_plan.hip, the
// single-problem C ABI in bbpd_solver.hip, batches in bbpd_batch.hip).
#pragma once
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

namespace bbpd_sv {


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
// Stencil tiles (mixed form): a wavefront takes 16 consecutive slots (x) on each of 4 lattice lines
// (slot stride = the box's length along x), and stages each stencil row as 4 lines of TILE_W sites
// through a window that slides one line per step in dy
constexpr int TILE_X = 16;
constexpr int TILE_L = 4;
constexpr int TILE_RX = 3;     // largest |dx| the tiles take
constexpr int TILE_W = TILE_X + 2 * TILE_RX; // sites of a staged line
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
// stencil bond kernel variants by where s0 and sc come from (PV_GENTAB: PV_GENERAL with a key table
// for one of them)
enum ParamVariant { PV_SCALAR = 0, PV_TABLE = 1, PV_GENERAL = 2, PV_GENTAB = 3 };

// A region is one problem instance: a block-aligned range of slots with its own stencil, tables,
// law scalars, contact bodies, load schedule, history and step counter.  A single problem is one
// region; a batch packs many into one launch of each kernel.
struct Region {
    int base, nN, nSlots;          // slots [base, base + nSlots); the first nN hold nodes
    int blk0, nblk;                // global blocks [blk0, blk0 + nblk)
    long long e0;                  // stencil layout: first entry of the region
    int noff, off0;                // stencil layout: offsets off0 .. off0 + noff - 1
    int nrow, row0, rx;            // stencil layout: rows row0 .. row0 + nrow - 1, largest |dx|
    int tnx, ntile;                // stencil layout in tiles (0 = slices): the box's length along x
                                   // and the tile count (see TILE_X)
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
    const double *L0str, *vcfstr; // FP64 record layout: UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS per
                                  // entry where the given arrays differ from their formulas (else
                                  // nullptr: both recomputed from the coordinates)
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
    const int4 *se;       // step-end blocks: (region, body index in the region or -1 (no bodies) or
                          // -2 (overlapping families), the body's index in bodies, 0)
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

inline double bits_to_double(uint32_t lo, uint32_t hi)
{
    uint64_t v = ((uint64_t)hi << 32) | lo;
    double d;
    std::memcpy(&d, &v, 8);
    return d;
}

inline void double_to_bits(double d, uint32_t &lo, uint32_t &hi)
{
    uint64_t v;
    std::memcpy(&v, &d, 8);
    lo = (uint32_t)v;
    hi = (uint32_t)(v >> 32);
}

uint64_t morton3(uint32_t x, uint32_t y, uint32_t z);

} // namespace bbpd_sv

using namespace bbpd_sv;

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
    int4 *d_se = nullptr;
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
    bool tiles = false; // the stencil bond kernel runs in tiles of TILE_X x TILE_L nodes
};

namespace bbpd_sv {

using KernelFn = void (*)(Params);
// launch variants of the stencil bond kernel: (minimum waves per SIMD, block order); orders: XCD
// remap, launch order, and launch order over the stripe list (variants 4 and 5)
constexpr int N_OCC = BBPD_AUTOTUNE_VARIANTS;
constexpr int ST_STRIPE = 8; // consecutive global blocks per XCD in the stripe list

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

// Stencil tiles: the slot of lane l of tile t in a box of nx sites per line (TILE_X x TILE_L lanes;
// tiles run along x, then over groups of TILE_L lines)
__host__ __device__ inline int tile_slot(int t, int l, int nx)
{
    const int tx = t % (nx / TILE_X), g = t / (nx / TILE_X);
    return TILE_X * tx + (l % TILE_X) + nx * (TILE_L * g + l / TILE_X);
}

// The history words of `nt` tiles (as history_words, per tile): bit o set when an entry of offset o
// of a lane of the tile holds a history word other than +0; the history `h` is in the slice order
// (slice, offset, lane) of the slots.
template <typename V>
std::vector<unsigned long long> tile_history_words(const V *h, long long nt, int noff, int nx)
{
    std::vector<unsigned long long> w((size_t)nt * 2, 0ull);
    for (long long t = 0; t < nt; t++)
        for (int l = 0; l < SLICE; l++) {
            const long long q = tile_slot((int)t, l, nx);
            const V *e = h + (q / SLICE) * noff * SLICE + q % SLICE;
            for (int o = 0; o < noff; o++)
                if (e[(long long)o * SLICE] != 0 || std::signbit(e[(long long)o * SLICE]))
                    w[(size_t)t * 2 + (o >> 6)] |= 1ull << (o & 63);
        }
    return w;
}

// ---- bbpd_kernels.hip: kernel choice and launches
std::vector<int> stripe_list(const std::vector<int> &list);
int env_launch_variant();
void launch_bond(bbpd_solver *s, KernelFn kb, int v);
KernelFn pick_bond_kernel(const bbpd_solver *s, int variant = -1);
int allow_lds(bbpd_solver *s);
int launch_step(bbpd_solver *s, bool allow_timing);
// softening / broken / failed entry counts of region ri into s->counts (k_count)
void launch_count(bbpd_solver *s, int ri);
// damage field, per-block damage energy and (sed != nullptr) strain-energy density of region ri (k_damage)
void launch_damage(bbpd_solver *s, int ri, double *dmg, double *epart, double *sed);

// ---- bbpd_plan.hip: set-up checks and the stencil plan
bool bits_equal(double a, double b);
bool make_table(const double *a, int n, int cap, std::vector<double> &tab, std::vector<uint32_t> &idx);
int check_bonds(const bbpd_problem *pr, std::vector<int> &bi, std::vector<int> &bj, bool *geom_differs = nullptr);
bool detect_lattice(const bbpd_problem *pr, double hlat[3], double xmin[3], double xmax[3], double &hq);
std::vector<unsigned long long> lane_masks(const std::vector<uint4> &m, long long nslots, int noff);
std::vector<unsigned long long> tile_lane_masks(const std::vector<uint4> &m, long long ntiles, int noff, int nx);
std::vector<StOff> offset_records(const std::vector<int4> &off, const double *h, double hmr, double hpr,
                                  double horizon, double two_r, bool mixed, const int *kmin, const int *kmax,
                                  int *fast64);
void plan_stencil(const bbpd_problem *pr, const double *X, const std::vector<int> &bi, const std::vector<int> &bj,
                  const double *hlat, StencilPlan &sp);

// ---- bbpd_solver.hip: region state and observables shared with the batch ABI
int region_observables(bbpd_solver *s, int ri, bbpd_observables *o);
int region_damage(bbpd_solver *s, int ri, const std::vector<int> &slot_of, double *damage, double *energy,
                  double *density = nullptr, double *strain = nullptr);
int region_history(bbpd_solver *s, int ri, double *out_rows, int cap, int *n_out);
long long current_step(bbpd_solver *s, int *rc, int ri = 0);
int current_buffer(bbpd_solver *s, int *rc);
void force_sets(const bbpd_solver *s, int cur, double **F, double **Fp);


} // namespace bbpd_sv
