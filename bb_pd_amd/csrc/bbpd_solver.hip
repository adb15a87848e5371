// bbpd_solver.hip -- the fused, device-resident explicit step: the single-problem C ABI
// (bbpd_solver_*) over the kernels of bbpd_kernels.hip.
#include "bbpd_solver_impl.hpp"


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
    // FP64 form: UNDEFORMEDLENGTH or VOLUMECORRECTIONFACTORS other than their formulas' values are
    // streamed per entry on the record layout (both layouts otherwise recompute them from X); the mixed
    // form, whose stretch is formed from the coordinates, refuses them
    bool geom_differs = false;
    {
        const int rc0 = check_bonds(pr, bi, bj, mixed ? nullptr : &geom_differs);
        if (rc0) return fail(rc0);
    }
    double xmin[3], xmax[3], hq, hlat[3];
    // given lengths and factors run as off-lattice inputs (neighbour coordinates from records): the
    // bond kernel's variant for those carries the two streams
    const bool lattice = detect_lattice(pr, hlat, xmin, xmax, hq) && !geom_differs;
    s->lattice = lattice;

    // ---- layout.  Both forms take the stencil layout where it applies: it moves a fraction of the
    // record layout's bytes per entry and loads each neighbour record once per stencil row.
    // BBPD_LAYOUT=records asks for the record layout.
    StencilPlan sp;
    {
        const char *lay = std::getenv("BBPD_LAYOUT");
        const bool want_records = lay && std::strcmp(lay, "records") == 0;
        if (geom_differs) sp.why = "UNDEFORMEDLENGTH or VOLUMECORRECTIONFACTORS differ from their formulas";
        else if (!lattice) sp.why = "not a lattice";
        else if (want_records) sp.why = "record layout requested";
        else plan_stencil(pr, X, bi, bj, hlat, sp);
    }
    s->stencil = sp.ok;
    s->layout = sp.ok ? "stencil" : "records (" + sp.why + ")";
    if (s->stencil && env_launch_variant() >= 0) s->occ = env_launch_variant();
    // tiles of TILE_X x TILE_L nodes (mixed form), asked for with BBPD_TILES=1: the box's length along x
    // a multiple of TILE_X, its lines in groups of TILE_L, offsets within TILE_RX along x, the entry
    // offsets within 31 bits.  Slices are the default: on the 400 x 40 x 80 beam the tiled kernel stages
    // 46% fewer sites but spends more on its line bookkeeping, and runs slower (206 against 181-192 us).
    const int tnx = sp.ok ? sp.kmax[0] - sp.kmin[0] + 1 : 0;
    {
        const char *te = std::getenv("BBPD_TILES");
        const bool on = te && std::string(te) == "1";
        const long long nS = ((long long)nN + SLOT_ALIGN - 1) / SLOT_ALIGN * SLOT_ALIGN;
        s->tiles = sp.ok && mixed && on && tnx % TILE_X == 0 && (nN / tnx) % TILE_L == 0 && sp.rx <= TILE_RX &&
                   (nS * sp.noff + 2 * SLICE) * (long long)sizeof(float) < (1LL << 31);
        if (s->tiles) s->layout = "stencil, 16x4 tiles";
    }

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
    std::vector<double> h_L0, h_vcf;
    if (geom_differs) {
        h_L0.assign(nE, 0.0);
        h_vcf.assign(nE, 0.0);
    }
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
                if (geom_differs) {
                    h_L0[e] = pr->UNDEFORMEDLENGTH[k];
                    h_vcf[e] = pr->VOLUMECORRECTIONFACTORS[k];
                }
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
    double *L0d = nullptr, *vcfd = nullptr;
    if (geom_differs && ((rc = dalloc(&L0d, nE, s->owned)) || (rc = dalloc(&vcfd, nE, s->owned)))) return fail(rc);
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
        const std::vector<unsigned long long> hmt =
            s->tiles ? tile_lane_masks(hm, nN / SLICE, sp.noff, tnx) : lane_masks(hm, nSlots, sp.noff);
        BBPD_HIP(hipMemcpy(maskTd, hmt.data(), sizeof(unsigned long long) * hmt.size(), hipMemcpyHostToDevice));
        if (mixed) {
            std::vector<float> h0(nE, 0.0f);
            for (int k = 0; k < nB; k++)
                if (pr->fail && pr->fail[k] == 0) h0[s->entry_i[k]] = h0[s->entry_j[k]] = -0.0f;
            BBPD_HIP(hipMemcpy(hsd, h0.data(), sizeof(float) * nE, hipMemcpyHostToDevice));
            const std::vector<unsigned long long> hw = s->tiles ? tile_history_words(h0.data(), nN / SLICE, sp.noff, tnx)
                                                                : history_words(h0.data(), s->nSlices, sp.noff);
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
    if (L0d) {
        BBPD_HIP(hipMemcpy(L0d, h_L0.data(), 8 * nE, hipMemcpyHostToDevice));
        BBPD_HIP(hipMemcpy(vcfd, h_vcf.data(), 8 * nE, hipMemcpyHostToDevice));
    }
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
    P.L0str = L0d;
    P.vcfstr = vcfd;
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
    R0.tnx = s->tiles ? tnx : 0;
    R0.ntile = s->tiles ? nN / SLICE : 0;
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
        std::vector<int4> se;
        if (R0.seq_contact) se.push_back(make_int4(0, -2, 0, 0));
        else
            for (int bdy = 0; bdy < std::max(pr->nBodies, 1); bdy++)
                se.push_back(pr->nBodies > 0 ? make_int4(0, bdy, R0.body0 + bdy, 0) : make_int4(0, -1, 0, 0));
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
        BBPD_HIP(hipMemcpy(s->d_se, se.data(), sizeof(int4) * se.size(), hipMemcpyHostToDevice));
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
                      (cstrd ? 8 * nE : 0) + (s0d ? ew * nE : 0) + (scd ? ew * nE : 0) + (L0d ? 16 * nE : 0);

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

namespace bbpd_sv {

int region_observables(bbpd_solver *s, int ri, bbpd_observables *o)
{
    BBPD_HIP(hipSetDevice(s->device));
    const Region &RG = s->reg[ri];
    BBPD_HIP(hipMemsetAsync(s->counts, 0, sizeof(unsigned long long) * 4, s->stream));
    launch_count(s, ri);
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

namespace bbpd_sv {

// damage field (reference node order) and total damage energy of one region
// damage field, damage energy and (when asked) the strain energy of one region, in reference order
int region_damage(bbpd_solver *s, int ri, const std::vector<int> &slot_of, double *damage, double *energy,
                  double *density, double *strain)
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
    launch_damage(s, ri, dd, ep, sd);
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

namespace bbpd_sv {

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
    const int noff = (int)(nE / s->nSlices / SLICE);
    const std::vector<unsigned long long> hw = s->tiles ? tile_history_words(hv.data(), s->reg[0].ntile, noff, s->reg[0].tnx)
                                                        : history_words(hv.data(), s->nSlices, noff);
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
long long current_step(bbpd_solver *s, int *rc, int ri)
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

namespace bbpd_sv {

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
    int64_t per_entry = (s->stencil ? w : 16) + (s->P.cmode == PM_STREAM && !mixed ? 8 : 0) + (s->P.L0str ? 16 : 0) +
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
        // words per slice, or per tile (a region's tiles take its first words)
        const int64_t sl0 = RG.base / SLICE,
                      sl1 = RG.ntile ? sl0 + RG.ntile : (RG.base + ((int64_t)RG.nN + SLICE - 1) / SLICE * SLICE) / SLICE;
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

