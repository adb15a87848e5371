/*
 * bbpd.h -- C ABI of the MI355X bond-based peridynamics step library (libbbpd_hip.so).
 *
 * Two layers:
 *
 *  1. Drop-in routines.  One entry point per computational routine of the reference's MEX
 *     files and per MATLAB-only hot-path function, with the reference's argument order, array
 *     layout (FP64, column-major, BONDLIST as nBonds x 2 doubles of 1-based ids, flags as doubles)
 *     and in-place semantics.  Arrays are DEVICE pointers (hipMalloc'd or torch CUDA tensors);
 *     `stream` is a hipStream_t (NULL = default stream).  Results are identical bit for bit to the
 *     reference C routines where those use only +,-,*,/,sqrt (everything except exp()).
 *     A MEX gateway binds each routine exactly as the reference's gateway does; the gateway code
 *     for every name is given in INTEGRATION.md.
 *
 *  2. Fused solver.  A device-resident state built once from the same reference arrays, that runs
 *     whole steps (stretch -> law -> force -> node-centric gather -> update -> contact) with no
 *     per-step host traffic, in plain FP64 (bit-identical to the reference order) or FP32-mixed.
 *
 * Every call returns BBPD_OK (0) or a negative code; bbpd_last_error() gives the message (the
 * reference reports only a wrong argument count, through mexErrMsgTxt).
 */
#ifndef BBPD_H
#define BBPD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BBPD_OK 0
#define BBPD_ERR_ARG (-1)
#define BBPD_ERR_HIP (-2)
#define BBPD_ERR_NOMEM (-3)
#define BBPD_ERR_UNSUPPORTED (-4)

/* constitutive laws of the fused solver */
#define BBPD_LAW_LINEAR 0      /* brittle latch, calculatebondfailureMex.c:26-44 */
#define BBPD_LAW_BILINEAR 1    /* calculatebondsofteningfactor.m:42-48 */
#define BBPD_LAW_TRILINEAR 2   /* calculatebsftrilinear.m:40-112 */
#define BBPD_LAW_EXPONENTIAL 3 /* calculatebsfexponentialMex.c:26-66 */

#define BBPD_INTEGRATOR_EULER_CROMER 0 /* timeintegrationeulercromerMex.c:26-53 */
#define BBPD_INTEGRATOR_ADR 1          /* calculatedampingcoefficient.m + timeintegrationADR.m */

#define BBPD_PRECISION_FP64 0
#define BBPD_PRECISION_MIXED 1

const char *bbpd_last_error(void);
int bbpd_version(void);
/* number of visible HIP devices (0 when none); never fails */
int bbpd_device_count(void);

/* ============================================================================================
 * Layer 1: drop-in routines (device pointers)
 * ========================================================================================== */

/* replaces calculatedeformedlength, solvers/functions/mex/calculatedeformedlengthMex.c:27 */
int bbpd_calculatedeformedlength(const double *deformedCoordinates, const double *BONDLIST,
                                 const double *undeformedLength, int nBonds, int nNodes, double *deformedLength,
                                 double *deformedX, double *deformedY, double *deformedZ, double *stretch,
                                 void *stream);

/* replaces calculatebsfexponential, solvers/functions/mex/calculatebsfexponentialMex.c:26 */
int bbpd_calculatebsfexponential(const double *stretch, const double *s0, const double *sc, double k, double alpha,
                                 double *bsf, double *bondSofteningFactor, const double *BONDTYPE,
                                 double *flagBondSoftening, int nBonds, void *stream);

/* replaces calculatebondforces, solvers/functions/mex/calculatebondforcesMex.c:26 */
int bbpd_calculatebondforces(double *bForceX, double *bForceY, double *bForceZ, const double *fail,
                             const double *deformedX, const double *deformedY, const double *deformedZ,
                             const double *deformedLength, const double *stretch, const double *stretchPlastic,
                             const double *BONDSTIFFNESS, double cellVolume, const double *VOLUMECORRECTIONFACTORS,
                             const double *bondSofteningFactor, int nBonds, void *stream);

/* replaces calculatenodalforces of solvers/functions/mex/calculatenodalforcesopenmpMex.c:29
 * (nThreads = 8 reproduces that file, which forces 8 threads at :39) and of the serial
 * solvers/functions/mex/calculatenodalforcesMex.c:19 (nThreads = 1, nfShared* may be NULL).
 * nodalForce and the nNodes x nThreads scratch arrays must be zeroed by the caller, as in the
 * reference; the scratch columns receive the per-thread partial sums the reference leaves there.
 * The summation order reproduces the reference's for that thread count, with no atomics. */
int bbpd_calculatenodalforces(int nBonds, int nNodes, const double *BONDLIST, double *nodalForce,
                              const double *bForceX, const double *bForceY, const double *bForceZ, double *nfSharedX,
                              double *nfSharedY, double *nfSharedZ, int nThreads, void *stream);

/* replaces timeintegrationeulercromer, solvers/functions/mex/timeintegrationeulercromerMex.c:26 */
int bbpd_timeintegrationeulercromer(double *nodalAcceleration, const double *nodalForce, double DAMPING,
                                    double *nodalVelocity, const double *DENSITY, const double *constraintFlag,
                                    double DT, double *nodalDisplacement, double *deformedCoordinates,
                                    const double *undeformedCoordinates, int nNodes, int NOD, void *stream);

/* replaces calculatebondfailure, solvers/functions/mex/calculatebondfailureMex.c:26 */
int bbpd_calculatebondfailure(double *fail, double failureFunctionality, const double *BONDTYPE,
                              const double *stretch, double criticalStretchConcrete, double criticalStretchSteel,
                              int nBonds, void *stream);

/* replaces solvers/functions/calculatebondsofteningfactor.m:1 (bilinear law).  s0/sc are arrays
 * with stride 1 or scalars (stride 0, pointer to one device double), as MATLAB broadcasts. */
int bbpd_calculatebondsofteningfactor(const double *stretch, const double *linearElasticLimit, int s0_stride,
                                      const double *criticalStretchConcrete, int sc_stride,
                                      double *flagBondSoftening, double *bondSofteningFactor,
                                      const double *BONDTYPE, int nBonds, void *stream);

/* replaces solvers/functions/calculatebsftrilinear.m:1 (scalar s0, s1, sc as the .m requires) */
int bbpd_calculatebsftrilinear(const double *stretch, double s0, double s1, double sc, double *bondSofteningFactor,
                               const double *BONDTYPE, double *flagBondSoftening, int nBonds, void *stream);

/* replaces solvers/functions/calculatedampingcoefficient.m:1 (NOD = 3).  Writes cn to *cn_host
 * (host memory) and synchronises the stream.  The two sums use a fixed-order device reduction. */
int bbpd_calculatedampingcoefficient(const double *nodalForce, const double *massVector,
                                     const double *nodalForcePrevious, double DT,
                                     const double *nodalVelocityPreviousHalf, const double *nodalDisplacement,
                                     int nNodes, double *cn_host, void *stream);

/* No counterpart in the reference.  The same routine as bbpd_calculatedampingcoefficient (same
 * arguments, kernels, summation order and result) for callers that cannot allocate device memory:
 * tight serving loops and harnesses that forbid hipMalloc.  It calls no hipMalloc, hipFree or other
 * allocating routine; its block partials live in a fixed device array of the library (about 1 MiB
 * per GPU, held until the process ends), which bounds nNodes by BBPD_NOALLOC_MAX_NODES (larger
 * inputs are rejected with BBPD_ERR_ARG).  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (other threads or streams) race. */
#define BBPD_NOALLOC_MAX_NODES 16777216
int bbpd_calculatedampingcoefficient_noalloc(const double *nodalForce, const double *massVector,
                                             const double *nodalForcePrevious, double DT,
                                             const double *nodalVelocityPreviousHalf,
                                             const double *nodalDisplacement, int nNodes, double *cn_host,
                                             void *stream);

/* replaces solvers/functions/timeintegrationADR.m:1 (NOD = 3) */
int bbpd_timeintegrationADR(int iTimeStep, double DT, double *nodalVelocityForwardHalf, const double *nodalForce,
                            const double *massVector, double cn, const double *nodalVelocityPreviousHalf,
                            double *nodalDisplacement, const double *CONSTRAINTFLAG, int nNodes, void *stream);

/* replaces solvers/functions/calculatecontactforce.m:1.  family holds nFamily 1-based node ids
 * (doubles, device); centre (cx, cz), radius as in the penetrator struct.  Writes penetratorfz to
 * *fz_host and synchronises the stream. */
int bbpd_calculatecontactforce(const double *family, int nFamily, double centreX, double centreZ, double radius,
                               double displacementIncrement, const double *undeformedCoordinates,
                               double *deformedCoordinates, double *nodalDisplacement, double *nodalVelocity,
                               double DT, double cellVolume, const double *DENSITY, int nNodes, double *fz_host,
                               void *stream);

/* ============================================================================================
 * Input generator (host memory; no GPU needed)
 * ========================================================================================== */
typedef struct bbpd_beam_spec {
    int nx, ny, nz;         /* nodes along x (length), y (width), z (depth) */
    double DX;              /* lattice spacing (m) */
    double horizon_factor;  /* delta = horizon_factor * DX */
    int notch;              /* 1: remove bonds crossing the notch rectangle */
    double notch_x;         /* plane x = notch_x * DX (half-integer) */
    double notch_depth;     /* notch top z = notch_depth * DX */
    double E;               /* Young's modulus (Pa) */
} bbpd_beam_spec;

int bbpd_beam_count(const bbpd_beam_spec *spec, int64_t *nNodes, int64_t *nBonds);
/* X: nNodes x 3; BONDLIST: nBonds x 2; others nBonds (nFAMILYMEMBERS: nNodes), all host */
int bbpd_beam_build(const bbpd_beam_spec *spec, int64_t nBondsCap, double *X, double *BONDLIST,
                    double *UNDEFORMEDLENGTH, double *VOLUMECORRECTIONFACTORS, double *BONDSTIFFNESS,
                    double *nFAMILYMEMBERS);

/* ============================================================================================
 * Layer 2: fused device-resident solver
 * ========================================================================================== */
#define BBPD_MAX_BODIES 8

typedef struct bbpd_problem {
    int nNodes, nBonds;
    /* reference input arrays (HOST memory, reference layout) */
    const double *undeformedCoordinates;   /* nNodes x 3 */
    const double *BONDLIST;                /* nBonds x 2, 1-based */
    const double *UNDEFORMEDLENGTH;        /* nBonds; must equal the length recomputed from X */
    const double *VOLUMECORRECTIONFACTORS; /* nBonds; must equal calculatevolumecorrectionfactors(L0) */
    const double *BONDSTIFFNESS;           /* nBonds */
    const double *BONDTYPE;                /* nBonds or NULL (all concrete) */
    const double *fail;                    /* nBonds or NULL (all 1) */
    const double *DENSITY;                 /* nNodes */
    const double *CONSTRAINTFLAG;          /* nNodes x 3 or NULL */
    const double *massVector;              /* nNodes (ADR) or NULL */
    double cellVolume, horizon, RADIJ;     /* RADIJ as in calculatevolumecorrectionfactors */
    /* law */
    int law;
    const double *s0;  /* nBonds or NULL -> s0_scalar */
    const double *sc;  /* nBonds or NULL -> sc_scalar */
    double s0_scalar, s1, sc_scalar, k, alpha;
    double sc_steel, failureFunctionality; /* linear law */
    /* integration */
    int integrator;
    double DT, DAMPING;
    /* contact bodies (calculatecontactforce.m), body 0 gives the reaction */
    int nBodies;
    const double *family[BBPD_MAX_BODIES]; /* 1-based ids, host */
    int nFamily[BBPD_MAX_BODIES];
    double centreX[BBPD_MAX_BODIES], centreZ[BBPD_MAX_BODIES], radius[BBPD_MAX_BODIES];
    int moves[BBPD_MAX_BODIES]; /* 1: centre follows the load schedule */
    /* load schedule, the load-point displacement of step i:
     * 0 = smoothstepdata(i, 1, nTimeSteps, 0, appliedDisplacement);
     * 1 = staircase: increment * ceil(i / itersPerIncrement);
     * 2 = equilibrium staircase (ADR): increment * inc, where inc (from 1) advances after the step at
     *     whose end the equilibrium ratio is below eqTol, once at least itersPerIncrementMin
     *     iterations ran at inc, or after itersPerIncrement iterations at inc.  The equilibrium ratio
     *     (after calculateequilibriumstate.m:6-8) is sqrt(sum of F_d^2 over the DOFs that neither
     *     CONSTRAINTFLAG nor a contact body holds) * cellVolume / |static reaction| of the step's
     *     forces, +infinity while the reaction is 0. */
    int schedule;
    int nTimeSteps;
    double appliedDisplacement;
    double increment;
    int itersPerIncrement;
    /* observables */
    int referenceNode; /* 1-based; 0 = none */
    int cmodNode1, cmodNode2;
    /* equilibrium staircase (schedule 2) */
    double eqTol;
    int itersPerIncrementMin;
} bbpd_problem;

typedef struct bbpd_solver bbpd_solver;

typedef struct bbpd_observables {
    int64_t step;          /* steps done */
    double reaction;       /* body-0 contact reaction of the last step (calculatecontactforce.m:121) */
    double reaction_static;/* sum of F_z * cellVolume over body-0 family (calculatereactionforceFast.m:3-5) */
    double deflection;     /* u_z at referenceNode */
    double cmod;           /* u_x(cmod2) - u_x(cmod1) */
    double cn;             /* last ADR damping coefficient */
    int64_t n_softening;   /* sum(flagBondSoftening) over bonds */
    int64_t n_broken;      /* bonds with bondSofteningFactor == 1 or fail == 0 */
    int64_t n_failed;      /* nBonds - sum(fail) (printoutput.m:40) */
} bbpd_observables;

int bbpd_solver_create(const bbpd_problem *problem, int precision, int device, bbpd_solver **out);
int bbpd_solver_destroy(bbpd_solver *s);
/* run n steps on the solver's stream (asynchronous) */
int bbpd_solver_step(bbpd_solver *s, int n);
/* capture `n` steps as one HIP graph and replay it `reps` times */
int bbpd_solver_run_graph(bbpd_solver *s, int n, int reps);
int bbpd_solver_sync(bbpd_solver *s);
void *bbpd_solver_stream(bbpd_solver *s);
int bbpd_solver_observables(bbpd_solver *s, bbpd_observables *obs);
/* state in reference layout, host memory (NULL pointers skipped):
 * nodal arrays nNodes x 3, bond arrays nBonds in BONDLIST order */
int bbpd_solver_get_nodal(bbpd_solver *s, double *nodalDisplacement, double *nodalVelocity, double *nodalForce,
                          double *deformedCoordinates);
int bbpd_solver_get_bonds(bbpd_solver *s, double *bondSofteningFactor, double *flagBondSoftening, double *fail);
/* Checkpoint / resume of the device state (replaces the workspace save of
 * dynamicsolverdisplacementcontrolled.m:133-141 and its reload at :46-50).  Host memory, reference
 * layout: nodal arrays nNodes x 3 (u, v, x = deformedCoordinates, F = nodalForce,
 * Fp = nodalForcePrevious for ADR), bond arrays nBonds in BONDLIST order (H =
 * bondSofteningFactor, flag = flagBondSoftening, fail), `step` = steps completed
 * (timeStepTracker - 1).  restore() takes the same arrays; NULL leaves that part unchanged.
 * Under ADR, F is the last step's force, which the next step takes as its previous force: at a
 * step boundary the reference's hand-over (dynamicsolverdisplacementcontrolledADR.m:153-155) has
 * made nodalForcePrevious equal nodalForce, so checkpoint() returns Fp equal to F, and restore()
 * takes the previous force from F (Fp is accepted for symmetry and has no effect on the steps).
 * The FP32-mixed form keeps displacements, not coordinates: its x is X + u on output and ignored
 * on input. */
int bbpd_solver_checkpoint(bbpd_solver *s, int64_t *step, double *u, double *v, double *x, double *F, double *Fp,
                           double *H, double *flag, double *fail);
int bbpd_solver_restore(bbpd_solver *s, int64_t step, const double *u, const double *v, const double *x,
                        const double *F, const double *Fp, const double *H, const double *flag, const double *fail);
/* Per-step record since the previous call, BBPD_HISTORY_COLUMNS doubles per step in this order:
 * body-0 contact reaction (calculatecontactforce.m:121), static reaction
 * (calculatereactionforceFast.m:3-5 over the body-0 family, from the step's forces), deflection
 * u_z(referenceNode), CMOD u_x(cmod2) - u_x(cmod1) (dynamicsolverdisplacementcontrolled.m:115-124),
 * the step's load-point displacement, its equilibrium ratio (see schedule 2), the ADR damping
 * coefficient (0 under Euler-Cromer), and 1 when the step ended a load increment (its static
 * reaction is then the increment's converged load), else 0.  At most `cap` steps are returned (the
 * device keeps the last 2^20). */
#define BBPD_HISTORY_COLUMNS 8
int bbpd_solver_history(bbpd_solver *s, double *rows, int cap, int *n_out);
/* Load state of the equilibrium staircase: the current increment (from 1) and the iterations done at
 * it.  get returns them (and, for the other schedules, the increment and iteration count the step
 * counter implies); set loads them, after bbpd_solver_restore, for schedule 2.  No counterpart in the
 * reference, whose load schedule is a function of the step (smoothstepdata.m:40-42). */
int bbpd_solver_get_load_state(bbpd_solver *s, int64_t *increment, int64_t *itersAtIncrement);
int bbpd_solver_set_load_state(bbpd_solver *s, int64_t increment, int64_t itersAtIncrement);
/* Damage field (calculatedamage.m:34-45: 1 - unbroken / nFAMILY per node, broken meaning H == 1 or
 * failed; damage is nNodes, host, reference order, may be NULL) and total dissipated damage energy
 * (calculatedamageenergy.m:35-62 summed over bonds, with the solver's k and alpha; may be NULL).
 * The reference keeps each bond's energy as a running max over steps; here it is evaluated from
 * the largest stretch the bond has reached, which its softening history fixes (the law inverted at
 * H).  That equals the running max for every loading path except the bilinear law's compression
 * branch, where the reference sets H = 1 without the stretch passing sc.  Energy needs a softening
 * law. */
int bbpd_solver_damage(bbpd_solver *s, double *damage, double *energy);
/* Strain energy (calculatestrainenergy.m:10-27) of the current state: per node (host, reference
 * order, may be NULL) the strain energy density 0.5 * sum over its bonds of the micropotential
 * 0.5 * fail * c * s^2 * L0 * cellVolume * vcf (left to right, BFMULTIPLIER = 1, s from the current
 * positions in FP64, bonds in BONDLIST order), and in *energy (may be NULL) the reference's total,
 * sum(density) * nNodes * cellVolume with the sum in node order. */
int bbpd_solver_strain_energy(bbpd_solver *s, double *nodalStrainEnergyDensity, double *strainEnergy);
/* device bytes held and the algorithmic HBM bytes one step of the bond kernel moves */
int bbpd_solver_info(bbpd_solver *s, int64_t *device_bytes, int64_t *bond_kernel_bytes, int64_t *n_entries,
                     int64_t *n_padded_entries);
/* Algorithmic HBM bytes of one bond-kernel launch at the current state.  bbpd_solver_info counts a
 * history word for every entry; the stencil layout's kernel reads the history of an offset only in
 * the slices (64 nodes) whose history word for it is set, so here the history (and any per-bond
 * parameter stream, read only by entries whose law runs) counts 64 entries per set (slice, offset)
 * pair: offsets_set of offsets_total such pairs.  The record layout returns the bbpd_solver_info
 * figure (and 0 pairs).  No counterpart in the reference. */
int bbpd_solver_bond_bytes(bbpd_solver *s, int64_t *bytes, int64_t *offsets_set, int64_t *offsets_total);
/* the bond-table layout chosen at set-up: "stencil" (full lattice box, rows in lattice-offset
 * order: neighbours implicit, per entry only the history) or "records (<why not stencil>)".  The
 * string lives as long as the solver.  Default: stencil where it applies, for both forms;
 * BBPD_LAYOUT=records in the environment asks for the record layout.  The FP64 stencil kernel takes
 * the undeformed lengths, their reciprocals and the volume correction factors from per-offset
 * nominal values where set-up finds that exact for every value they take on the lattice, else
 * computes them per entry and reports "stencil (FP64 lengths computed per entry)"; BBPD_FAST64=0
 * asks for the latter. */
const char *bbpd_solver_layout(bbpd_solver *s);
/* Autotuner: times each compiled launch variant of the stencil bond kernel over `reps` launches on
 * the current state and keeps the fastest for the solver or batch.  Variants: 0 = compiler's
 * register budget, XCD-contiguous block order; 1 = the same, blocks in launch order; 2 / 3 = at
 * least 5 waves per SIMD (fewer registers per wave), XCD order / launch order; 4 / 5 = the
 * register budgets of 0 / 2 with the blocks in XCD stripes (each XCD takes runs of 8 consecutive
 * blocks, the runs dealt out in launch order).  XCD order keeps a block's neighbour rows in its
 * XCD's L2; launch order spreads a localised process zone over all XCDs; stripes do some of both.
 * The first call runs each variant once untimed before timing it (loading its code); later calls
 * time at once.  ms (may be NULL) receives BBPD_AUTOTUNE_VARIANTS averages, choice the index kept.
 * Needs the ADR integrator (whose bond kernel leaves the state unchanged when rerun); a no-op on the
 * record layout.  BBPD_LAUNCH_VARIANT=v in the environment fixes the variant at set-up instead of 0.
 * No counterpart in the reference. */
#define BBPD_AUTOTUNE_VARIANTS 6
int bbpd_solver_autotune(bbpd_solver *s, int reps, double *ms, int *choice);
/* average duration (ms) of the bond kernel over the steps since the last call, from HIP events
 * recorded on the solver's stream around each launch (0 when timing is off) */
int bbpd_solver_set_timing(bbpd_solver *s, int on);
int bbpd_solver_kernel_time(bbpd_solver *s, double *avg_ms, int64_t *launches);

/* ============================================================================================
 * Batches: many problems stepped together in one launch of each kernel.  No counterpart in the
 * reference, whose only multi-specimen workflow runs specimens one after another
 * (solvers/scripts/randomfield/dynamicsolverdisplacementcontrolledRF.m:20-114, its CMOD stop being
 * the analogue of the stop criterion here).  A batch is a pool of regions: problems are admitted
 * into free node slots while others run, step together, and are retired to free their slots.
 * Batches use the stencil layout, so every problem must be a full lattice box with a BONDLIST sorted
 * by node id (what the input generator produces).  bbpd_solver_step / run_graph / sync / info /
 * set_timing / kernel_time / destroy work on a batch as on a single problem.
 * ========================================================================================== */
#define BBPD_ERR_FULL (-5) /* the batch has no free region or slots for the problem now */

typedef struct bbpd_batch_spec {
    int maxRegions;            /* problems held at once */
    int64_t maxSlots;          /* node slots over all problems (a problem takes nNodes rounded up to 256) */
    int historyRows;           /* per-problem history ring rows (0: 65536) */
    double DX;                 /* lattice spacing every problem shares */
    const bbpd_problem *shape; /* every admitted problem shares its cellVolume, horizon, RADIJ, law, k,
                                  alpha, sc_steel, failureFunctionality, integrator, DT, DAMPING, and
                                  whether s0 / sc are per bond (arrays) or scalars */
} bbpd_batch_spec;

typedef struct bbpd_region_status {
    int active;             /* the region holds a problem */
    int finished;           /* its stop criterion was met */
    int64_t step;           /* steps done */
    double peak;            /* largest static reaction at a load-increment end */
    double reaction_static; /* static reaction of the last step */
    int64_t nNodes, nBonds;
} bbpd_region_status;

int bbpd_batch_create(const bbpd_batch_spec *spec, int precision, int device, bbpd_solver **out);
/* Admit a problem: it starts at step 0 with the next launch.  stopFraction / stopDisplacement (0 =
 * none): the problem is marked finished at the first load-increment end where its static reaction
 * is below stopFraction of its peak, or the load point has travelled stopDisplacement.  Returns
 * BBPD_ERR_FULL when no region or slot range is free now, BBPD_ERR_NOMEM when the problem needs more
 * node slots than the whole batch holds (it can never be admitted). */
int bbpd_batch_admit(bbpd_solver *s, const bbpd_problem *problem, double stopFraction, double stopDisplacement,
                     int *region);
/* Admission in two parts, so that a service can check and lay out arriving problems on host threads
 * while the batch steps: bbpd_batch_prepare does everything bbpd_batch_admit does on the host (the
 * checks, the stencil plan, the node and entry arrays in the problem's own slot order) and touches
 * neither the device nor the batch's membership, so it may run on another host thread while the batch
 * steps; bbpd_batch_admit_prepared then finds a region and a slot range and uploads.  Errors are those
 * of bbpd_batch_admit; after BBPD_ERR_FULL the admission stays valid for a later call.  Free it with
 * bbpd_admission_free once admitted or abandoned (not while a call that uses it runs). */
typedef struct bbpd_admission bbpd_admission;
int bbpd_batch_prepare(bbpd_solver *s, const bbpd_problem *problem, bbpd_admission **out);
int bbpd_batch_admit_prepared(bbpd_solver *s, bbpd_admission *admission, double stopFraction,
                              double stopDisplacement, int *region);
void bbpd_admission_free(bbpd_admission *admission);
int bbpd_batch_retire(bbpd_solver *s, int region);
/* status of regions 0 .. n-1 (n <= maxRegions) in one device read */
int bbpd_batch_status(bbpd_solver *s, bbpd_region_status *status, int n);
int bbpd_batch_observables(bbpd_solver *s, int region, bbpd_observables *obs);
int bbpd_batch_history(bbpd_solver *s, int region, double *rows, int cap, int *n_out);
int bbpd_batch_get_nodal(bbpd_solver *s, int region, double *nodalDisplacement, double *nodalVelocity,
                         double *nodalForce, double *deformedCoordinates);
int bbpd_batch_get_bonds(bbpd_solver *s, int region, double *bondSofteningFactor, double *flagBondSoftening,
                         double *fail);
int bbpd_batch_damage(bbpd_solver *s, int region, double *damage, double *energy);
int bbpd_batch_strain_energy(bbpd_solver *s, int region, double *nodalStrainEnergyDensity, double *strainEnergy);

#ifdef __cplusplus
}
#endif
#endif /* BBPD_H */
