// bbpd_plan.hip -- set-up checks of a problem and the stencil-layout plan (host code).
#include <thread>

#include "bbpd_solver_impl.hpp"

namespace bbpd_sv {

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

namespace {

// The checks of bonds [k0, k1) (see check_bonds): BBPD_OK, or the error code of the first that fails,
// with the message set when `report`.  With `geom`, a length or volume correction factor that differs
// from its formula sets *geom instead of failing.
int check_bond_range(const bbpd_problem *pr, int k0, int k1, int *bi, int *bj, bool report, int *geom)
{
    const int nN = pr->nNodes, nB = pr->nBonds;
    const double *X = pr->undeformedCoordinates, *BL = pr->BONDLIST;
    const double horizon = pr->horizon, RADIJ = pr->RADIJ;
    const double hmr = horizon - RADIJ, hpr = horizon + RADIJ, two_r = 2 * RADIJ;
    for (int k = k0; k < k1; k++) {
        double di = BL[k], dj = BL[(size_t)nB + k];
        int i = (int)di, j = (int)dj;
        if (di != (double)i || dj != (double)j || i < 1 || j < 1 || i > nN || j > nN || i == j) {
            if (report)
                bbpd::set_error("BONDLIST row " + std::to_string(k + 1) + " is not a pair of distinct ids in 1..nNodes");
            return BBPD_ERR_ARG;
        }
        bi[k] = i - 1;
        bj[k] = j - 1;
        double ex = X[i - 1] - X[j - 1], ey = X[i - 1 + nN] - X[j - 1 + nN],
               ez = X[i - 1 + 2 * (size_t)nN] - X[j - 1 + 2 * (size_t)nN];
        double L0 = std::sqrt((ex * ex + ey * ey) + ez * ez);
        if (!bits_equal(L0, pr->UNDEFORMEDLENGTH[k]) && geom) {
            *geom = 1;
        } else if (!bits_equal(L0, pr->UNDEFORMEDLENGTH[k])) {
            if (report)
                bbpd::set_error("UNDEFORMEDLENGTH(" + std::to_string(k + 1) +
                                ") differs from the length recomputed from undeformedCoordinates (buildhorizons.m:93-97)");
            return BBPD_ERR_UNSUPPORTED;
        }
        double vcf;
        if (L0 <= hmr) vcf = 1;
        else if ((L0 > hmr) && (L0 <= horizon)) vcf = (hpr - L0) / two_r;
        else vcf = 0;
        if (!bits_equal(vcf, pr->VOLUMECORRECTIONFACTORS[k]) && geom) {
            *geom = 1;
        } else if (!bits_equal(vcf, pr->VOLUMECORRECTIONFACTORS[k])) {
            if (report)
                bbpd::set_error("VOLUMECORRECTIONFACTORS(" + std::to_string(k + 1) +
                                ") differs from calculatevolumecorrectionfactors(L0, horizon, RADIJ)");
            return BBPD_ERR_UNSUPPORTED;
        }
        if (pr->fail && !(pr->fail[k] == 0 || pr->fail[k] == 1)) {
            if (report) bbpd::set_error("fail must hold 0 or 1");
            return BBPD_ERR_ARG;
        }
        if (pr->BONDTYPE && !(pr->BONDTYPE[k] == 0 || pr->BONDTYPE[k] == 1 || pr->BONDTYPE[k] == 2)) {
            if (report) bbpd::set_error("BONDTYPE must hold 0, 1 or 2");
            return BBPD_ERR_ARG;
        }
    }
    return BBPD_OK;
}

} // namespace

// Checks of a problem's bond arrays: ids, UNDEFORMEDLENGTH and VOLUMECORRECTIONFACTORS against their
// formulas (buildhorizons.m:93-97, calculatevolumecorrectionfactors.m:20-32: the device recomputes
// both from the coordinates), fail and BONDTYPE values.  Fills the 0-based ends.  A length or factor
// that differs from its formula is an error (BBPD_ERR_UNSUPPORTED), or, with geom_differs, sets
// *geom_differs (the caller then streams the given arrays).  Large problems are checked in chunks by
// several host threads; the first failing bond in bond order is the one reported.
int check_bonds(const bbpd_problem *pr, std::vector<int> &bi, std::vector<int> &bj, bool *geom_differs)
{
    const int nB = pr->nBonds;
    bi.assign(nB, 0);
    bj.assign(nB, 0);
    const int T = nB < (1 << 20) ? 1 : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    const int chunk = (nB + T - 1) / T;
    std::vector<int> rc(T, BBPD_OK), geom(T, 0);
    int *const g = geom_differs ? geom.data() : nullptr;
    std::vector<std::thread> th;
    for (int t = 1; t < T; t++)
        th.emplace_back([&, t] {
            rc[t] = check_bond_range(pr, std::min(nB, t * chunk), std::min(nB, (t + 1) * chunk), bi.data(), bj.data(),
                                     false, g ? g + t : nullptr);
        });
    rc[0] = check_bond_range(pr, 0, std::min(nB, chunk), bi.data(), bj.data(), true, g);
    for (auto &x : th) x.join();
    for (int t = 0; t < T; t++)
        if (rc[t] != BBPD_OK) {
            // the first failing chunk again, on this thread, for its message
            return check_bond_range(pr, std::min(nB, t * chunk), std::min(nB, (t + 1) * chunk), bi.data(), bj.data(),
                                    true, g ? g + t : nullptr);
        }
    if (geom_differs) *geom_differs = std::find(geom.begin(), geom.end(), 1) != geom.end();
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

// The same per tile (TILE_X x TILE_L slots, tile_slot): bit l of (tile, offset) set when the slot of
// lane l has the bond at that offset.
std::vector<unsigned long long> tile_lane_masks(const std::vector<uint4> &m, long long ntiles, int noff, int nx)
{
    std::vector<unsigned long long> r((size_t)ntiles * noff, 0ull);
    for (long long t = 0; t < ntiles; t++)
        for (int l = 0; l < SLICE; l++) {
            const uint32_t *w = reinterpret_cast<const uint32_t *>(&m[tile_slot((int)t, l, nx)]);
            for (int o = 0; o < noff; o++)
                if ((w[o >> 5] >> (o & 31)) & 1u) r[(size_t)t * noff + o] |= 1ull << l;
        }
    return r;
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


} // namespace bbpd_sv
