// bbpd_batch.hip -- batches (bbpd_batch_*): many problem instances as regions of one set of
// device arrays, stepped by one launch of each kernel.
#include <memory>

#include "bbpd_solver_impl.hpp"

// ================================================================================================
// Batches: a pool of regions in one set of device arrays, stepped by one launch of each kernel.
// Every region has fixed ranges in the shared pools (offsets, key tables, bodies, families, history
// rows) by its index, a block-aligned slot range from a first-fit allocator, and entries at
// ST_MAX per slot from its first slot.
// ================================================================================================
namespace bbpd_sv {

constexpr int RB_TAB = 3 * ST_GMAX * ST_GMAX; // key-table entries per region: c, s0, sc
constexpr int RB_FAM = 1 << 16;               // contact-family members per region

// block -> region, launch list (active blocks, padded to a multiple of 8), step-end jobs, active list
int batch_rebuild(bbpd_solver *s)
{
    const int nblk = s->nSlots / BLOCK;
    std::vector<int> blkreg(nblk, 0), grid, act;
    std::vector<int4> se;
    size_t maxtab = 0;
    for (int r = 0; r < s->maxRegions; r++) {
        if (!s->active[r]) continue;
        const Region &RG = s->reg[r];
        for (int b = 0; b < RG.nblk; b++) {
            blkreg[RG.blk0 + b] = r;
            grid.push_back(RG.blk0 + b);
        }
        if (RG.nBodies == 0) se.push_back(make_int4(r, -1, 0, 0));
        else if (RG.seq_contact) se.push_back(make_int4(r, -2, 0, 0));
        else
            for (int b = 0; b < RG.nBodies; b++) se.push_back(make_int4(r, b, RG.body0 + b, 0));
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
    if (!se.empty()) BBPD_HIP(hipMemcpy(s->d_se, se.data(), sizeof(int4) * se.size(), hipMemcpyHostToDevice));
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

// A problem checked and laid out for a batch on the host, in its own local slot and entry order: the
// admission proper (bbpd_batch_admit_prepared) only finds it a region and a slot range, offsets
// the local indices and uploads.
struct bbpd_admission {
    const bbpd_solver *batch = nullptr; // the batch it was prepared for
    int nN = 0, nB = 0, need = 0;       // nodes, bonds, blocks of BLOCK slots
    int noff = 0;
    std::vector<int> slot;              // node -> local slot
    std::vector<long long> ei, ej;      // bond -> local entry at its two ends
    std::vector<NodeRec> hpos;
    std::vector<double> hX, hrho, hm;
    std::vector<unsigned char> hc;
    std::vector<uint4> hmask;
    std::vector<unsigned long long> hmt; // lane masks per (slice, offset)
    bool any_failed = false;            // some bond starts failed: history and words uploaded
    std::vector<unsigned char> h0, hs0, hsc; // entries in the device width (FP32 mixed, FP64)
    std::vector<unsigned long long> hw; // history words (any_failed)
    std::vector<int4> off;
    std::vector<int2> rows;
    std::vector<StOff> hst;
    std::vector<double> tab;
    std::vector<float> tabf;
    std::vector<int> fam;               // members of the bodies in order, local slots
    std::vector<Body> hb;               // bodies (fam pointers set at admission)
    std::vector<int> nfam;
    Region RG{};                        // the descriptor, less what depends on the region and slots
    int ref = -1, cmod1 = -1, cmod2 = -1; // local observation slots
};

namespace bbpd_sv {

// The room for a problem of nN nodes: BBPD_ERR_NOMEM when it needs more node slots than the batch
// holds (it can never be admitted), BBPD_ERR_FULL when no region or slot range is free now.
int batch_room(bbpd_solver *s, int nN, int *rid, int *blk0)
{
    if ((long long)(nN + BLOCK - 1) / BLOCK > (long long)s->freeblk.size()) {
        bbpd::set_error("the problem needs " + std::to_string(((long long)nN + BLOCK - 1) / BLOCK * BLOCK) +
                        " node slots, more than the batch holds (" +
                        std::to_string((long long)s->freeblk.size() * BLOCK) + ")");
        return BBPD_ERR_NOMEM;
    }
    *rid = -1;
    for (int r = 0; r < s->maxRegions && *rid < 0; r++)
        if (!s->active[r]) *rid = r;
    const int need = (nN + BLOCK - 1) / BLOCK;
    *blk0 = -1;
    const int nblk = (int)s->freeblk.size();
    int run = 0;
    for (int b = 0; b < nblk; b++) {
        run = s->freeblk[b] ? run + 1 : 0;
        if (run == need) {
            *blk0 = b - need + 1;
            break;
        }
    }
    if (*rid < 0 || *blk0 < 0) {
        bbpd::set_error(*rid < 0 ? "batch full: no free region" : "batch full: no free slot range");
        return BBPD_ERR_FULL;
    }
    return BBPD_OK;
}

} // namespace bbpd_sv

extern "C" int bbpd_batch_prepare(bbpd_solver *s, const bbpd_problem *pr, bbpd_admission **out)
{
    BBPD_REQUIRE(s && s->batch && pr && out, "bad arguments");
    *out = nullptr;
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
    if ((long long)(nN + BLOCK - 1) / BLOCK > (long long)s->freeblk.size()) {
        int rid, blk0;
        return batch_room(s, nN, &rid, &blk0); // the never-fits error
    }
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

    std::unique_ptr<bbpd_admission> A(new bbpd_admission());
    A->batch = s;
    A->nN = nN;
    A->nB = nB;
    A->need = (nN + BLOCK - 1) / BLOCK;
    const int nSl = A->need * BLOCK;
    const int noff = sp.noff;
    A->noff = noff;
    A->slot = sp.slot;

    // node data in local slot order
    A->hpos.assign(nSl, NodeRec{});
    std::memset(A->hpos.data(), 0, sizeof(NodeRec) * nSl);
    A->hX.assign(3 * (size_t)nSl, 0.0);
    A->hrho.assign(nSl, 1.0);
    A->hm.assign(nSl, 1.0);
    A->hc.assign(nSl, 0);
    A->hmask.assign(nSl, make_uint4(0, 0, 0, 0));
    for (int n = 0; n < nN; n++) {
        const int q = sp.slot[n];
        double xv[3];
        for (int d = 0; d < 3; d++) {
            xv[d] = pr->undeformedCoordinates[n + (size_t)nN * d];
            A->hX[q + (size_t)nSl * d] = xv[d];
            if (pr->CONSTRAINTFLAG && pr->CONSTRAINTFLAG[n + (size_t)nN * d] == 1) A->hc[q] |= (unsigned char)(1u << d);
        }
        A->hpos[q].x = mixed ? 0.0 : xv[0];
        A->hpos[q].y = mixed ? 0.0 : xv[1];
        A->hpos[q].z = mixed ? 0.0 : xv[2];
        A->hpos[q].q = (uint64_t)sp.k[3 * (size_t)n] | ((uint64_t)sp.k[3 * (size_t)n + 1] << LAT_BITS) |
                       ((uint64_t)sp.k[3 * (size_t)n + 2] << (2 * LAT_BITS)) | ((uint64_t)sp.gidx[n] << G_SHIFT);
        A->hrho[q] = pr->DENSITY[n];
        if (pr->massVector) A->hm[q] = pr->massVector[n];
    }
    // entries: masks, history (sign bit set for bonds that start failed; all +0 when none does, which
    // is set on the device), streamed parameters, in the device width
    const long long nEr = (long long)nSl * noff;
    const size_t w = mixed ? 4 : 8;
    A->any_failed = pr->fail && std::any_of(pr->fail, pr->fail + nB, [](double f) { return f == 0; });
    std::vector<double> h0(A->any_failed ? nEr : 0, 0.0), hs0(pr->s0 ? nEr : 0, 0.0), hsc(pr->sc ? nEr : 0, 0.0);
    A->ei.assign(nB, 0);
    A->ej.assign(nB, 0);
    auto dense_local = [&](int q, int o) { return ((long long)(q / SLICE) * noff + o) * SLICE + (q % SLICE); };
    for (int k = 0; k < nB; k++) {
        const int qi = sp.slot[bi[k]], qj = sp.slot[bj[k]];
        const long long ei = dense_local(qi, sp.oi[k]), ej = dense_local(qj, sp.oj[k]);
        A->ei[k] = ei;
        A->ej[k] = ej;
        reinterpret_cast<uint32_t *>(&A->hmask[qi])[sp.oi[k] >> 5] |= 1u << (sp.oi[k] & 31);
        reinterpret_cast<uint32_t *>(&A->hmask[qj])[sp.oj[k] >> 5] |= 1u << (sp.oj[k] & 31);
        if (A->any_failed && pr->fail[k] == 0) h0[ei] = h0[ej] = -0.0;
        if (pr->s0) hs0[ei] = hs0[ej] = pr->s0[k];
        if (pr->sc) hsc[ei] = hsc[ej] = pr->sc[k];
    }
    A->hmt = lane_masks(A->hmask, nSl, noff);
    auto to_width = [&](const std::vector<double> &v, std::vector<unsigned char> &dst) {
        dst.resize(v.size() * w);
        if (mixed) {
            float *f = reinterpret_cast<float *>(dst.data());
            for (size_t i = 0; i < v.size(); i++) f[i] = (float)v[i];
        } else {
            std::memcpy(dst.data(), v.data(), v.size() * sizeof(double));
        }
    };
    if (A->any_failed) {
        A->hw = history_words(h0.data(), nSl / SLICE, noff);
        to_width(h0, A->h0);
    }
    if (pr->s0) to_width(hs0, A->hs0);
    if (pr->sc) to_width(hsc, A->hsc);
    // contact bodies (families as local slots), contact-member bits
    A->hb.assign(BBPD_MAX_BODIES, Body{});
    A->nfam.assign(BBPD_MAX_BODIES, 0);
    std::vector<unsigned char> in_family(nN, 0);
    std::vector<int> owner(nN, -1);
    int seq_contact = 0; // families that share members are applied in order by one block
    for (int bdy = 0; bdy < pr->nBodies; bdy++) {
        const int nf = pr->nFamily[bdy];
        BBPD_REQUIRE(nf >= 0 && (nf == 0 || pr->family[bdy]), "contact family missing");
        Body &B = A->hb[bdy];
        A->nfam[bdy] = nf;
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
            A->fam.push_back(sp.slot[id - 1]);
        }
        B.cx = pr->centreX[bdy];
        B.cz = pr->centreZ[bdy];
        B.R = pr->radius[bdy];
        B.moves = pr->moves[bdy];
        B.index = bdy;
    }
    if (pr->integrator == BBPD_INTEGRATOR_ADR)
        for (int n = 0; n < nN; n++)
            if (in_family[n]) A->hc[sp.slot[n]] |= CONTACT_BIT;
    // per-offset data and key tables
    const double hmr = sh.horizon - sh.RADIJ, hpr = sh.horizon + sh.RADIJ, two_r = 2 * sh.RADIJ;
    const double hdx[3] = {s->DX, s->DX, s->DX};
    int fast64 = 0;
    A->hst = offset_records(sp.off, hdx, hmr, hpr, sh.horizon, two_r, mixed, sp.kmin, sp.kmax, &fast64);
    A->off = sp.off;
    A->rows = sp.rows;
    A->tab = sp.tc;
    A->tab.insert(A->tab.end(), sp.ts0.begin(), sp.ts0.end());
    A->tab.insert(A->tab.end(), sp.tsc.begin(), sp.tsc.end());
    BBPD_REQUIRE((int)A->tab.size() <= RB_TAB, "key tables too large for a batch region");
    A->tabf.resize(A->tab.size());
    for (size_t i = 0; i < sp.tc.size(); i++) A->tabf[i] = (float)(sp.tc[i] * sh.cellVolume);
    for (size_t i = sp.tc.size(); i < A->tab.size(); i++) A->tabf[i] = (float)A->tab[i];

    // descriptor, less what depends on the region and its slots
    Region &RG = A->RG;
    RG.nN = nN;
    RG.nSlots = nSl;
    RG.nblk = A->need;
    RG.noff = noff;
    RG.nrow = sp.nrow;
    RG.rx = sp.rx;
    RG.fast64 = fast64;
    RG.kg = sp.G;
    RG.ntab = (int)A->tab.size();
    RG.ts0 = (int)sp.tc.size();
    RG.tsc = (int)(sp.tc.size() + sp.ts0.size());
    // key-tabled s0 / sc are not used in a batch: per-bond values are streamed
    RG.s0 = pr->s0 ? 0.0 : pr->s0_scalar;
    RG.sc = pr->sc ? 0.0 : pr->sc_scalar;
    RG.s1 = pr->s1;
    RG.s0lo = pr->s0 && nB > 0 ? *std::min_element(pr->s0, pr->s0 + nB) : 0.0;
    RG.sclo = pr->sc && nB > 0 ? *std::min_element(pr->sc, pr->sc + nB) : 0.0;
    RG.nBodies = pr->nBodies;
    RG.seq_contact = seq_contact;
    RG.schedule = pr->schedule;
    RG.nTimeSteps = pr->nTimeSteps;
    RG.ipi = pr->itersPerIncrement > 0 ? pr->itersPerIncrement : 1;
    RG.applied = pr->appliedDisplacement;
    RG.increment = pr->increment;
    RG.eq_tol = pr->eqTol;
    RG.ipi_min = pr->itersPerIncrementMin;
    auto local_or_none = [&](int id1) { return (id1 >= 1 && id1 <= nN) ? sp.slot[id1 - 1] : -1; };
    A->ref = local_or_none(pr->referenceNode);
    A->cmod1 = local_or_none(pr->cmodNode1);
    A->cmod2 = local_or_none(pr->cmodNode2);
    *out = A.release();
    return BBPD_OK;
}

extern "C" void bbpd_admission_free(bbpd_admission *a) { delete a; }

extern "C" int bbpd_batch_admit_prepared(bbpd_solver *s, bbpd_admission *A, double stop_frac, double stop_disp,
                                         int *region)
{
    BBPD_REQUIRE(s && s->batch && A && region, "bad arguments");
    BBPD_REQUIRE(A->batch == s, "the admission was prepared for another batch");
    *region = -1;
    const bool mixed = s->precision == BBPD_PRECISION_MIXED;
    int rid, blk0, rc;
    if ((rc = batch_room(s, A->nN, &rid, &blk0))) return rc;
    BBPD_HIP(hipSetDevice(s->device));
    const int need = A->need, nSl = need * BLOCK, noff = A->noff;
    const int base = blk0 * BLOCK;
    const long long e0 = (long long)base * ST_MAX;
    const long long nEr = (long long)nSl * noff;
    BBPD_HIP(hipStreamSynchronize(s->stream));

    auto &RH = s->rh[rid];
    RH.nN = A->nN;
    RH.nB = A->nB;
    RH.slot_of.resize(A->nN);
    for (int n = 0; n < A->nN; n++) RH.slot_of[n] = base + A->slot[n];
    RH.entry_i.resize(A->nB);
    RH.entry_j.resize(A->nB);
    for (int k = 0; k < A->nB; k++) {
        RH.entry_i[k] = e0 + A->ei[k];
        RH.entry_j[k] = e0 + A->ej[k];
    }
    int *fambase = s->d_fampool + (size_t)rid * RB_FAM;
    std::vector<int> fam(A->fam.size());
    for (size_t t = 0; t < fam.size(); t++) fam[t] = base + A->fam[t];
    std::vector<Body> hb(A->hb);
    {
        size_t f0 = 0;
        for (int bdy = 0; bdy < A->RG.nBodies; bdy++) {
            hb[bdy].fam = fambase + f0;
            hb[bdy].region = rid;
            f0 += (size_t)A->nfam[bdy];
        }
    }
    Region RG = A->RG;
    RG.base = base;
    RG.blk0 = blk0;
    RG.e0 = e0;
    RG.off0 = rid * ST_MAX;
    RG.row0 = rid * (ST_MAX + ST_PAD);
    RG.sst0 = rid * (ST_MAX + ST_PAD);
    RG.tab0 = rid * RB_TAB;
    RG.body0 = rid * BBPD_MAX_BODIES;
    RG.ref_slot = A->ref >= 0 ? base + A->ref : -1;
    RG.cmod1 = A->cmod1 >= 0 ? base + A->cmod1 : -1;
    RG.cmod2 = A->cmod2 >= 0 ? base + A->cmod2 : -1;
    RG.hist0 = (long long)rid * s->histRows;
    RG.hist_cap = s->histRows;
    RG.stop_frac = stop_frac;
    RG.stop_disp = stop_disp;

    // upload
    Params &P = s->P;
    for (int q = 0; q < 2; q++)
        BBPD_HIP(hipMemcpy(P.pos[q] + base, A->hpos.data(), sizeof(NodeRec) * nSl, hipMemcpyHostToDevice));
    double *Xc[3] = {(double *)P.X0, (double *)P.X1, (double *)P.X2};
    double *uc[3] = {P.u0, P.u1, P.u2}, *vc[3] = {P.v0, P.v1, P.v2}, *fc[3] = {P.F0, P.F1, P.F2};
    double *fpc[3] = {P.Fp0, P.Fp1, P.Fp2};
    for (int d = 0; d < 3; d++) {
        BBPD_HIP(hipMemcpy(Xc[d] + base, A->hX.data() + (size_t)nSl * d, sizeof(double) * nSl, hipMemcpyHostToDevice));
        // zero fills on the batch's stream (outside any captured graph), ordered before its next steps
        if (uc[d]) BBPD_HIP(hipMemsetAsync(uc[d] + base, 0, sizeof(double) * nSl, s->stream));
        BBPD_HIP(hipMemsetAsync(vc[d] + base, 0, sizeof(double) * nSl, s->stream));
        BBPD_HIP(hipMemsetAsync(fc[d] + base, 0, sizeof(double) * nSl, s->stream));
        if (fpc[d]) BBPD_HIP(hipMemsetAsync(fpc[d] + base, 0, sizeof(double) * nSl, s->stream));
    }
    BBPD_HIP(hipMemcpy((double *)P.rho + base, A->hrho.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
    if (P.mass) BBPD_HIP(hipMemcpy((double *)P.mass + base, A->hm.data(), sizeof(double) * nSl, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((unsigned char *)P.cons + base, A->hc.data(), nSl, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((uint4 *)P.smask + base, A->hmask.data(), sizeof(uint4) * nSl, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((unsigned long long *)P.smaskT + e0 / SLICE, A->hmt.data(),
                       sizeof(unsigned long long) * A->hmt.size(), hipMemcpyHostToDevice));
    const size_t w = mixed ? 4 : 8;
    if (A->any_failed) {
        BBPD_HIP(hipMemcpy((char *)P.hs + (size_t)e0 * w, A->h0.data(), A->h0.size(), hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(P.snz + 2 * (base / SLICE), A->hw.data(), sizeof(unsigned long long) * A->hw.size(),
                           hipMemcpyHostToDevice));
    } else {
        BBPD_HIP(hipMemsetAsync((char *)P.hs + (size_t)e0 * w, 0, (size_t)nEr * w, s->stream));
        BBPD_HIP(hipMemsetAsync(P.snz + 2 * (base / SLICE), 0, sizeof(unsigned long long) * 2 * (nSl / SLICE),
                                s->stream));
    }
    if (!A->hs0.empty())
        BBPD_HIP(hipMemcpy((char *)P.s0str + (size_t)e0 * w, A->hs0.data(), A->hs0.size(), hipMemcpyHostToDevice));
    if (!A->hsc.empty())
        BBPD_HIP(hipMemcpy((char *)P.scstr + (size_t)e0 * w, A->hsc.data(), A->hsc.size(), hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((int4 *)P.soff + RG.off0, A->off.data(), sizeof(int4) * noff, hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((int2 *)P.srow + RG.row0, A->rows.data(), sizeof(int2) * A->rows.size(), hipMemcpyHostToDevice));
    BBPD_HIP(hipMemcpy((StOff *)P.sst + RG.sst0, A->hst.data(), sizeof(StOff) * A->hst.size(), hipMemcpyHostToDevice));
    if (!A->tab.empty()) {
        BBPD_HIP(hipMemcpy((double *)P.tab + RG.tab0, A->tab.data(), sizeof(double) * A->tab.size(),
                           hipMemcpyHostToDevice));
        if (P.tabf)
            BBPD_HIP(hipMemcpy((float *)P.tabf + RG.tab0, A->tabf.data(), sizeof(float) * A->tabf.size(),
                               hipMemcpyHostToDevice));
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

extern "C" int bbpd_batch_admit(bbpd_solver *s, const bbpd_problem *pr, double stop_frac, double stop_disp,
                                int *region)
{
    BBPD_REQUIRE(s && s->batch && pr && region, "bad arguments");
    *region = -1;
    // room first (a region and a free slot range), before the set-up checks and the plan: a caller
    // that retries while the batch is full pays only this test
    int rid, blk0, rc;
    BBPD_REQUIRE(pr->nNodes > 0, "need nNodes > 0 and nBonds >= 0");
    if ((rc = batch_room(s, pr->nNodes, &rid, &blk0))) return rc;
    bbpd_admission *A = nullptr;
    if ((rc = bbpd_batch_prepare(s, pr, &A))) return rc;
    rc = bbpd_batch_admit_prepared(s, A, stop_frac, stop_disp, region);
    bbpd_admission_free(A);
    return rc;
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

