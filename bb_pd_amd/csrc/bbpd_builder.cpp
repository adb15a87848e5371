// bbpd_builder.cpp -- host-side input generator for regular-lattice beams.
//
// Produces the reference input arrays in the reference data model (FP64, column-major,
// BONDLIST as nBonds x 2 doubles of 1-based node ids) so that they can be handed to the
// drop-in entry points or to the fused solver unchanged.  Each step restates one input builder:
//   coordinates      input/functions/buildmaterialpointcoordinates.m:39-59 (x = DX*k, k from 1,
//                    x fastest, then y, then z)
//   bond list        input/functions/buildhorizons.m:54-83 (pairs i<j with |xi| <= delta).  The
//                    reference orders a node's partners by rangesearch's distance order; this
//                    builder uses ascending j, the canonical order all parity checks use.
//   undeformed length input/functions/buildhorizons.m:89-97 (sum of squared differences, sqrt)
//   notch            input/functions/buildnotch.m:13-67 + determineintersection.m:59-150: a bond is
//                    removed when its segment meets the rectangle x = xloc*DX, y in [ymin,ymax],
//                    z in [zmin, depth*DX] (closed).  Evaluated in exact rational arithmetic.
//   family counts    input/functions/buildnotch.m:73-85 (recount after removal)
//   vcf              input/functions/calculatevolumecorrectionfactors.m:16-34
//   stiffness        input/functions/buildbonddata.m:48-51 with c0 = 12E/(pi delta^4)
//                    (input/scripts/Gregoire_beam_tests.m:206-208); all nodes concrete.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <string>

#include "bbpd.h"

namespace bbpd {
void set_error(const std::string &msg);
}

namespace {

struct Offset {
    int dx, dy, dz;
};

// Integer lattice offsets inside the horizon.  Lattice distances are sqrt(integer), so the test
// |xi| <= hf*DX is evaluated as dx^2+dy^2+dz^2 <= hf^2.  It is rejected when hf^2 sits within
// 1e-9 of an integer, where the reference's floating-point range search could go either way.
bool horizon_offsets(double hf, std::vector<Offset> &out, std::string &err)
{
    double h2 = hf * hf;
    double fr = h2 - std::floor(h2);
    if (fr < 1e-9 || fr > 1 - 1e-9) {
        err = "horizon factor squared is (nearly) an integer: neighbour set ambiguous";
        return false;
    }
    int r = (int)std::ceil(hf);
    out.clear();
    for (int dz = -r; dz <= r; dz++)
        for (int dy = -r; dy <= r; dy++)
            for (int dx = -r; dx <= r; dx++) {
                if (!dx && !dy && !dz) continue;
                if ((double)(dx * dx + dy * dy + dz * dz) <= h2) out.push_back({dx, dy, dz});
            }
    return true;
}

struct Lattice {
    int nx, ny, nz;
    int64_t n() const { return (int64_t)nx * ny * nz; }
    int64_t id(int ix, int iy, int iz) const { return ix + (int64_t)nx * (iy + (int64_t)ny * iz); }
};

// Exact notch test.  Nodes sit at integer lattice coordinates k = 1..n (units of DX).  The plane
// is x = xloc (xloc*2 must be odd so no node lies on it).  The rectangle spans all y and
// z in [1, depth].  Returns true when the closed segment a-b meets the rectangle.
bool notch_cuts(int ax, int az, int bx, int bz, int P /* 2*xloc */, double depth)
{
    // crossing: (2ax - P) and (2bx - P) have opposite signs
    long sa = 2L * ax - P, sb = 2L * bx - P;
    if ((sa < 0) == (sb < 0)) return false;
    long D = (long)bx - ax; // nonzero
    // z_I = az + (P/2 - ax) (bz - az) / D = num / den with den = 2D
    long num = 2L * az * D + (long)(P - 2L * ax) * (bz - az);
    long den = 2L * D;
    if (den < 0) {
        num = -num;
        den = -den;
    }
    // z_I >= 1 always holds for nodes with z >= 1; test z_I <= depth exactly (den*depth is exact
    // for the short binary fractions used as notch depths)
    return (double)num <= depth * (double)den;
}

} // namespace

extern "C" int bbpd_beam_count(const bbpd_beam_spec *sp, int64_t *nNodes, int64_t *nBonds)
{
    if (!sp || !nNodes || !nBonds || sp->nx <= 0 || sp->ny <= 0 || sp->nz <= 0 || !(sp->DX > 0)) {
        bbpd::set_error("bbpd_beam_count: invalid lattice");
        return BBPD_ERR_ARG;
    }
    std::vector<Offset> off;
    std::string err;
    if (!horizon_offsets(sp->horizon_factor, off, err)) {
        bbpd::set_error(err);
        return BBPD_ERR_ARG;
    }
    int P = 0;
    if (sp->notch) {
        double p2 = sp->notch_x * 2;
        P = (int)std::lround(p2);
        if (std::fabs(p2 - P) > 1e-12 || (P % 2) == 0) {
            bbpd::set_error("bbpd_beam_count: notch_x must be a half-integer (plane between lattice planes)");
            return BBPD_ERR_ARG;
        }
    }
    Lattice L{sp->nx, sp->ny, sp->nz};
    int64_t cnt = 0;
    for (int iz = 0; iz < L.nz; iz++)
        for (int iy = 0; iy < L.ny; iy++)
            for (int ix = 0; ix < L.nx; ix++) {
                for (const Offset &o : off) {
                    int jx = ix + o.dx, jy = iy + o.dy, jz = iz + o.dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= L.nx || jy >= L.ny || jz >= L.nz) continue;
                    if (L.id(jx, jy, jz) <= L.id(ix, iy, iz)) continue;
                    if (sp->notch && notch_cuts(ix + 1, iz + 1, jx + 1, jz + 1, P, sp->notch_depth)) continue;
                    cnt++;
                }
            }
    *nNodes = L.n();
    *nBonds = cnt;
    return BBPD_OK;
}

extern "C" int bbpd_beam_build(const bbpd_beam_spec *sp, int64_t nBondsCap, double *X, double *BONDLIST,
                               double *UNDEFORMEDLENGTH, double *VOLUMECORRECTIONFACTORS, double *BONDSTIFFNESS,
                               double *nFAMILYMEMBERS)
{
    int64_t nN = 0, nB = 0;
    int rc = bbpd_beam_count(sp, &nN, &nB);
    if (rc) return rc;
    if (nB > nBondsCap || !X || !BONDLIST || !UNDEFORMEDLENGTH || !VOLUMECORRECTIONFACTORS || !BONDSTIFFNESS ||
        !nFAMILYMEMBERS) {
        bbpd::set_error("bbpd_beam_build: output arrays missing or too small");
        return BBPD_ERR_ARG;
    }
    std::vector<Offset> off;
    std::string err;
    horizon_offsets(sp->horizon_factor, off, err);
    // partners of a node in ascending id order: sort offsets by their id delta.  The delta order
    // equals the id order whenever both endpoints are inside the lattice.
    Lattice L{sp->nx, sp->ny, sp->nz};
    std::sort(off.begin(), off.end(), [&](const Offset &a, const Offset &b) {
        int64_t da = a.dx + (int64_t)L.nx * (a.dy + (int64_t)L.ny * a.dz);
        int64_t db = b.dx + (int64_t)L.nx * (b.dy + (int64_t)L.ny * b.dz);
        return da < db;
    });
    const double DX = sp->DX;
    // coordinates (buildmaterialpointcoordinates.m:46-52)
    for (int iz = 0; iz < L.nz; iz++)
        for (int iy = 0; iy < L.ny; iy++)
            for (int ix = 0; ix < L.nx; ix++) {
                int64_t n = L.id(ix, iy, iz);
                X[n] = DX * (ix + 1);
                X[n + nN] = DX * (iy + 1);
                X[n + 2 * nN] = DX * (iz + 1);
            }
    int P = sp->notch ? (int)std::lround(sp->notch_x * 2) : 0;
    std::memset(nFAMILYMEMBERS, 0, sizeof(double) * nN);
    int64_t k = 0;
    for (int iz = 0; iz < L.nz; iz++)
        for (int iy = 0; iy < L.ny; iy++)
            for (int ix = 0; ix < L.nx; ix++) {
                int64_t i = L.id(ix, iy, iz);
                for (const Offset &o : off) {
                    int jx = ix + o.dx, jy = iy + o.dy, jz = iz + o.dz;
                    if (jx < 0 || jy < 0 || jz < 0 || jx >= L.nx || jy >= L.ny || jz >= L.nz) continue;
                    int64_t j = L.id(jx, jy, jz);
                    if (j <= i) continue;
                    if (sp->notch && notch_cuts(ix + 1, iz + 1, jx + 1, jz + 1, P, sp->notch_depth)) continue;
                    BONDLIST[k] = (double)(i + 1);
                    BONDLIST[nB + k] = (double)(j + 1);
                    nFAMILYMEMBERS[i] += 1;
                    nFAMILYMEMBERS[j] += 1;
                    k++;
                }
            }
    // undeformed length (buildhorizons.m:93,97): MATLAB sum over a 1x3 row is ((a+b)+c)
    for (int64_t b = 0; b < nB; b++) {
        int64_t i = (int64_t)BONDLIST[b] - 1, j = (int64_t)BONDLIST[nB + b] - 1;
        double ex = X[i] - X[j], ey = X[i + nN] - X[j + nN], ez = X[i + 2 * nN] - X[j + 2 * nN];
        UNDEFORMEDLENGTH[b] = std::sqrt((ex * ex + ey * ey) + ez * ez);
    }
    // volume correction (calculatevolumecorrectionfactors.m:16-34), RADIJ = DX/2
    const double horizon = sp->horizon_factor * DX;
    const double RADIJ = DX / 2;
    for (int64_t b = 0; b < nB; b++) {
        double UL = UNDEFORMEDLENGTH[b];
        double v;
        if (UL <= (horizon - RADIJ)) v = 1;
        else if ((UL > (horizon - RADIJ)) && (UL <= horizon)) v = (horizon + RADIJ - UL) / (2 * RADIJ);
        else v = 0;
        VOLUMECORRECTIONFACTORS[b] = v;
    }
    // surface-corrected stiffness (buildbonddata.m:48-51); neighbourhoodVolume (Gregoire_beam_tests.m:206)
    const double cellVolume = DX * DX * DX;
    const double nbhdV = (4.0 / 3.0) * M_PI * (horizon * horizon * horizon);
    const double c0 = (12 * sp->E) / (M_PI * (horizon * horizon * horizon * horizon));
    for (int64_t b = 0; b < nB; b++) {
        int64_t i = (int64_t)BONDLIST[b] - 1, j = (int64_t)BONDLIST[nB + b] - 1;
        double vi = nFAMILYMEMBERS[i] * cellVolume, vj = nFAMILYMEMBERS[j] * cellVolume;
        double sf = (2 * nbhdV) / ((vi + vj));
        BONDSTIFFNESS[b] = sf * c0;
    }
    return BBPD_OK;
}
