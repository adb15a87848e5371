// bbpd_dropin.hip -- layer 1: drop-in device routines with the reference's argument lists.
//
// Each kernel restates one reference routine over the reference data model (FP64, column-major
// nodal arrays, BONDLIST as nBonds x 2 doubles of 1-based ids).  The arithmetic keeps the
// reference's operation order and is compiled with -ffp-contract=off; division and sqrt are the
// correctly rounded IEEE operations, so results match the reference C bit for bit except where
// exp() is involved (a different libm).
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>
#include <memory>

#include "bbpd_internal.hpp"

namespace {

constexpr int BLOCK = 256;

// ---------------------------------------------------------------------------------------------
// calculatedeformedlengthMex.c:27-47
// ---------------------------------------------------------------------------------------------
__global__ void k_deformedlength(const double *__restrict__ x, const double *__restrict__ BL,
                                 const double *__restrict__ L0, int nB, int nN, double *__restrict__ l,
                                 double *__restrict__ dX, double *__restrict__ dY, double *__restrict__ dZ,
                                 double *__restrict__ s)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nB) return;
    int i = (int)BL[k];
    int j = (int)BL[(size_t)nB + k];
    double ex = x[j - 1] - x[i - 1];
    double ey = x[(size_t)j + nN - 1] - x[(size_t)i + nN - 1];
    double ez = x[(size_t)j + 2 * (size_t)nN - 1] - x[(size_t)i + 2 * (size_t)nN - 1];
    double len = sqrt(ex * ex + ey * ey + ez * ez);
    dX[k] = ex;
    dY[k] = ey;
    dZ[k] = ez;
    l[k] = len;
    s[k] = (len - L0[k]) / L0[k];
}

// ---------------------------------------------------------------------------------------------
// calculatebsfexponentialMex.c:26-66
// ---------------------------------------------------------------------------------------------
__global__ void k_bsfexponential(const double *__restrict__ s, const double *__restrict__ s0,
                                 const double *__restrict__ sc, double k, double alpha, double *__restrict__ bsf,
                                 double *__restrict__ H, const double *__restrict__ BT, double *__restrict__ flag,
                                 int nB)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    if (BT[b] != 0) return;
    double st = s[b], a0 = s0[b], ac = sc[b];
    double t = bsf[b];
    if (a0 < st && st <= ac) {
        flag[b] = 1;
        double numerator = 1 - exp(-k * (st - a0) / (ac - a0));
        double denominator = 1 - exp(-k);
        double residual = alpha * (1 - (st - a0) / (ac - a0));
        t = 1 - (a0 / st) * ((1 - (numerator / denominator)) + residual) / (1 + alpha);
        bsf[b] = t;
    } else if (st > ac) {
        t = 1;
        bsf[b] = t;
    }
    double h = H[b];
    if (t > h) h = t;
    if (h > 1) h = 1;
    if (h != h) h = 0.0;
    H[b] = h;
}

// ---------------------------------------------------------------------------------------------
// calculatebondforcesMex.c:26-45 (left-to-right product, NaN -> 0)
// ---------------------------------------------------------------------------------------------
__global__ void k_bondforces(double *__restrict__ fX, double *__restrict__ fY, double *__restrict__ fZ,
                             const double *__restrict__ fail, const double *__restrict__ dX,
                             const double *__restrict__ dY, const double *__restrict__ dZ,
                             const double *__restrict__ l, const double *__restrict__ s,
                             const double *__restrict__ sp, const double *__restrict__ c, double V,
                             const double *__restrict__ vcf, const double *__restrict__ H, int nB)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    double p = fail[b] * c[b] * (1 - H[b]) * (s[b] - sp[b]) * V * vcf[b];
    double len = l[b];
    double gx = p * (dX[b] / len), gy = p * (dY[b] / len), gz = p * (dZ[b] / len);
    fX[b] = (gx != gx) ? 0.0 : gx;
    fY[b] = (gy != gy) ? 0.0 : gy;
    fZ[b] = (gz != gz) ? 0.0 : gz;
}

// ---------------------------------------------------------------------------------------------
// Nodal reduction as a deterministic gather.  inc[] lists, for every node, the bonds touching it
// in ascending bond order, coded (k << 1) | (node is the j end).  Serial order
// (calculatenodalforcesMex.c:27-44): F accumulates every contribution in bond order.  OpenMP
// order (calculatenodalforcesopenmpMex.c:50-81): bonds are split into T static blocks, each block
// sums into its scratch column, then F adds the T columns in thread order.
// ---------------------------------------------------------------------------------------------
__device__ inline int static_owner(long k, long nB, int T)
{
    long q = nB / T, r = nB % T;
    long cut = r * (q + 1);
    if (k < cut) return (int)(k / (q + 1));
    return (int)(r + (k - cut) / q);
}

__global__ void k_nodal_gather(int nN, long nB, const long *__restrict__ rowptr, const long *__restrict__ inc,
                               double *__restrict__ F, const double *__restrict__ fX, const double *__restrict__ fY,
                               const double *__restrict__ fZ, double *__restrict__ sX, double *__restrict__ sY,
                               double *__restrict__ sZ, int T)
{
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nN) return;
    long e = rowptr[n], e1 = rowptr[n + 1];
    if (T <= 0) {
        double ax = F[n], ay = F[(size_t)n + nN], az = F[(size_t)n + 2 * (size_t)nN];
        for (; e < e1; e++) {
            long w = inc[e];
            long k = w >> 1;
            if (w & 1) {
                ax -= fX[k];
                ay -= fY[k];
                az -= fZ[k];
            } else {
                ax += fX[k];
                ay += fY[k];
                az += fZ[k];
            }
        }
        F[n] = ax;
        F[(size_t)n + nN] = ay;
        F[(size_t)n + 2 * (size_t)nN] = az;
        return;
    }
    double Fx = F[n], Fy = F[(size_t)n + nN], Fz = F[(size_t)n + 2 * (size_t)nN];
    for (int t = 0; t < T; t++) {
        size_t q = (size_t)n + (size_t)nN * t;
        double px = sX ? sX[q] : 0.0, py = sY ? sY[q] : 0.0, pz = sZ ? sZ[q] : 0.0;
        while (e < e1) {
            long w = inc[e];
            long k = w >> 1;
            if (static_owner(k, nB, T) != t) break;
            if (w & 1) {
                px -= fX[k];
                py -= fY[k];
                pz -= fZ[k];
            } else {
                px += fX[k];
                py += fY[k];
                pz += fZ[k];
            }
            e++;
        }
        if (sX) sX[q] = px;
        if (sY) sY[q] = py;
        if (sZ) sZ[q] = pz;
        Fx += px;
        Fy += py;
        Fz += pz;
    }
    F[n] = Fx;
    F[(size_t)n + nN] = Fy;
    F[(size_t)n + 2 * (size_t)nN] = Fz;
}

// integer fingerprint of a bond list (order-independent sum of per-row hashes; exact in uint64)
__global__ void k_bl_hash(const double *__restrict__ BL, int nB, unsigned long long *out)
{
    int k = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long h = 0;
    if (k < nB) {
        unsigned long long i = (unsigned long long)(long long)BL[k];
        unsigned long long j = (unsigned long long)(long long)BL[(size_t)nB + k];
        unsigned long long z = (unsigned long long)k * 0x9E3779B97F4A7C15ull ^ (i << 32) ^ (j * 0xBF58476D1CE4E5B9ull);
        z ^= z >> 31;
        z *= 0x94D049BB133111EBull;
        z ^= z >> 29;
        h = z;
    }
    if (k < nB) atomicAdd(out, h);
}

// ---------------------------------------------------------------------------------------------
// timeintegrationeulercromerMex.c:26-53
// ---------------------------------------------------------------------------------------------
__global__ void k_eulercromer(double *__restrict__ a, const double *__restrict__ F, double D,
                              double *__restrict__ v, const double *__restrict__ rho,
                              const double *__restrict__ cons, double DT, double *__restrict__ u,
                              double *__restrict__ x, const double *__restrict__ X, int nN, int NOD)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= (long)nN * NOD) return;
    int n = (int)(q % nN);
    double acc = (F[q] - D * v[q]) / rho[n];
    if (cons[q] == 1) acc = 0;
    double vv = v[q] + (acc * DT);
    double uu = u[q] + (vv * DT);
    a[q] = acc;
    v[q] = vv;
    u[q] = uu;
    x[q] = X[q] + uu;
}

// ---------------------------------------------------------------------------------------------
// calculatebondfailureMex.c:26-44
// ---------------------------------------------------------------------------------------------
__global__ void k_bondfailure(double *__restrict__ fail, double ff, const double *__restrict__ BT,
                              const double *__restrict__ s, double csc, double css, int nB)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    double f = fail[b], t = BT[b], st = s[b];
    if (f == 1 && t == 0 && st > csc) f = ff;
    if (f == 1 && t == 1 && st > 3 * csc) f = ff;
    if (f == 1 && t == 2 && st > css) f = ff;
    fail[b] = f;
}

// ---------------------------------------------------------------------------------------------
// calculatebondsofteningfactor.m:42-48 (MATLAB max ignores a NaN operand)
// ---------------------------------------------------------------------------------------------
__global__ void k_bilinear(const double *__restrict__ s, const double *__restrict__ s0, int s0s,
                           const double *__restrict__ sc, int scs, double *__restrict__ flag,
                           double *__restrict__ H, const double *__restrict__ BT, int nB)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    double a = s0[(size_t)b * s0s], c = sc[(size_t)b * scs], st = s[b];
    double fl = flag[b];
    if (BT[b] == 0 && st > a) {
        fl = 1;
        flag[b] = 1;
    }
    double cand = (((st - a) / st) * (c / (c - a))) * fl;
    double h = H[b];
    if (cand == cand && cand > h) h = cand;
    if (h > 1) h = 1;
    if (h != h) h = 0;
    H[b] = h;
}

// ---------------------------------------------------------------------------------------------
// calculatebsftrilinear.m:40-112
// ---------------------------------------------------------------------------------------------
__global__ void k_trilinear(const double *__restrict__ s, double s0, double s1, double sc, double *__restrict__ H,
                            const double *__restrict__ BT, double *__restrict__ flag, int nB)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nB) return;
    const double beta = 0.25;
    const double eta = s1 / s0;
    double h = H[b];
    if (BT[b] == 0) {
        double st = s[b];
        double bsf = 0;
        if ((st > s0) && (st <= s1)) {
            bsf = (1 - ((eta - beta) / (eta - 1) * (s0 / st))) + ((1 - beta) / (eta - 1));
            flag[b] = 1;
        } else if ((st > s1) && (st <= sc)) {
            bsf = 1 - ((s0 * beta) / st) * ((sc - st) / (sc - s1));
        } else if (st > sc) {
            bsf = 1;
        }
        if (bsf > h) h = bsf;
    }
    if (h != h) h = 0;
    H[b] = h;
}

// ---------------------------------------------------------------------------------------------
// calculatedampingcoefficient.m:59-112: per-block sums in a fixed tree, then one block
// ---------------------------------------------------------------------------------------------
__global__ void k_damping_partials(const double *__restrict__ F, const double *__restrict__ m,
                                   const double *__restrict__ Fp, double DT, const double *__restrict__ vp,
                                   const double *__restrict__ u, int nN, double *__restrict__ part)
{
    __shared__ double sh[BLOCK];
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    double c1 = 0, c2 = 0;
    if (n < nN) {
        double mi = m[n];
        for (int d = 0; d < 3; d++) {
            size_t q = (size_t)n + (size_t)nN * d;
            double uq = u[q];
            if (vp[q] != 0) c1 = c1 - (uq * uq) * ((F[q] / mi - Fp[q] / mi) / (DT * vp[q]));
            c2 = c2 + uq * uq;
        }
    }
    double s1 = bbpd::block_sum_fixed<BLOCK>(c1, sh);
    double s2 = bbpd::block_sum_fixed<BLOCK>(c2, sh);
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s1;
        part[2 * blockIdx.x + 1] = s2;
    }
}

__global__ void k_damping_final(const double *__restrict__ part, int nPart, double *__restrict__ cn_out)
{
    __shared__ double sh[BLOCK];
    double c1 = 0, c2 = 0;
    for (int p = threadIdx.x; p < nPart; p += BLOCK) {
        c1 = c1 + part[2 * p];
        c2 = c2 + part[2 * p + 1];
    }
    double s1 = bbpd::block_sum_fixed<BLOCK>(c1, sh);
    double s2 = bbpd::block_sum_fixed<BLOCK>(c2, sh);
    if (threadIdx.x == 0) {
        double cn = 0;
        if (s2 != 0) cn = ((s1 / s2) > 0) ? 2 * sqrt(s1 / s2) : 0;
        if (cn > 2) cn = 1.9;
        cn_out[0] = cn;
        cn_out[1] = s1;
        cn_out[2] = s2;
    }
}

// ---------------------------------------------------------------------------------------------
// timeintegrationADR.m:51-80
// ---------------------------------------------------------------------------------------------
__global__ void k_adr(int it, double DT, double *__restrict__ vf, const double *__restrict__ F,
                      const double *__restrict__ m, double cn, const double *__restrict__ vp, double *__restrict__ u,
                      const double *__restrict__ cons, int nN)
{
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= 3L * nN) return;
    int n = (int)(q % nN);
    double minv = 1 / m[n];
    double nv;
    if (it == 1) nv = 0.5 * DT * minv * F[q];
    else nv = (((2 - cn * DT) * vp[q]) + (2 * DT * minv * F[q])) / (2 + cn * DT);
    double nu = u[q] + (DT * nv);
    if (cons[q] == 1) {
        nu = 0;
        nv = 0;
    }
    vf[q] = nv;
    u[q] = nu;
}

// ---------------------------------------------------------------------------------------------
// calculatecontactforce.m:62-121 -- one block; members are projected in parallel, the reaction is
// summed in family order by one thread (the order of MATLAB's sum over dpenfz).
// ---------------------------------------------------------------------------------------------
constexpr int CONTACT_TILE = 4096;

__global__ void k_contact(const double *__restrict__ family, int nFam, double cx, double cz, double R, double disp,
                          const double *__restrict__ X, double *__restrict__ x, double *__restrict__ u,
                          double *__restrict__ v, double DT, double V, const double *__restrict__ rho, int nN,
                          double *__restrict__ fz_out)
{
    __shared__ double dfz[CONTACT_TILE];
    __shared__ unsigned char hit[CONTACT_TILE];
    const double czm = cz + disp;
    double total = 0;
    for (int base = 0; base < nFam; base += CONTACT_TILE) {
        int cnt = min(CONTACT_TILE, nFam - base);
        for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
            int i = (int)family[base + t] - 1;
            double ddx = x[i] - cx;
            double ddz = x[(size_t)i + 2 * (size_t)nN] - czm;
            double dist = sqrt((ddx * ddx) + (ddz * ddz));
            hit[t] = 0;
            if (dist < R) {
                hit[t] = 1;
                x[i] = cx + (ddx / dist) * R;
                x[(size_t)i + 2 * (size_t)nN] = czm + (ddz / dist) * R;
                double d_z = 0;
                for (int d = 0; d < 3; d++) {
                    size_t q = (size_t)i + (size_t)nN * d;
                    double uo = u[q], vo = v[q];
                    double nu = x[q] - X[q];
                    double nv = (nu - uo) / DT;
                    u[q] = nu;
                    v[q] = nv;
                    if (d == 2) d_z = -1 * rho[i] * (nv - vo) / DT * V;
                }
                dfz[t] = d_z;
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            for (int t = 0; t < cnt; t++)
                if (hit[t]) total = total + dfz[t];
        __syncthreads();
    }
    if (threadIdx.x == 0) fz_out[0] = total;
}

// ---------------------------------------------------------------------------------------------
// Cached node incidence lists for the nodal reduction, keyed by the bond list's device address,
// sizes and content fingerprint.
// ---------------------------------------------------------------------------------------------
struct IncPlan {
    const double *bl = nullptr;
    int nB = 0, nN = 0;
    unsigned long long hash = 0;
    int device = -1;
    long *rowptr = nullptr;
    long *inc = nullptr;
    ~IncPlan()
    {
        if (rowptr) (void)hipFree(rowptr);
        if (inc) (void)hipFree(inc);
    }
};

std::mutex g_plan_mu;
std::unique_ptr<IncPlan> g_plan;
unsigned long long *g_hash_buf[64] = {nullptr};
double *g_small_buf[64] = {nullptr}; // per-device small scratch for scalar results and partials
size_t g_small_cap[64] = {0};

int device_scratch(int dev, size_t bytes, double **out)
{
    if (dev < 0 || dev >= 64) {
        bbpd::set_error("device ordinal out of range");
        return BBPD_ERR_ARG;
    }
    if (g_small_cap[dev] < bytes) {
        if (g_small_buf[dev]) BBPD_HIP(hipFree(g_small_buf[dev]));
        g_small_buf[dev] = nullptr;
        g_small_cap[dev] = 0;
        BBPD_HIP(hipMalloc(&g_small_buf[dev], bytes));
        g_small_cap[dev] = bytes;
    }
    *out = g_small_buf[dev];
    return BBPD_OK;
}

int get_plan(const double *BL, int nB, int nN, hipStream_t st, IncPlan **out)
{
    int dev = 0;
    BBPD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        bbpd::set_error("device ordinal out of range");
        return BBPD_ERR_ARG;
    }
    if (!g_hash_buf[dev]) BBPD_HIP(hipMalloc(&g_hash_buf[dev], sizeof(unsigned long long)));
    unsigned long long h = 0;
    BBPD_HIP(hipMemsetAsync(g_hash_buf[dev], 0, sizeof(unsigned long long), st));
    if (nB > 0) k_bl_hash<<<bbpd::grid_for(nB, BLOCK), BLOCK, 0, st>>>(BL, nB, g_hash_buf[dev]);
    BBPD_HIP(hipGetLastError());
    BBPD_HIP(hipMemcpyAsync(&h, g_hash_buf[dev], sizeof(h), hipMemcpyDeviceToHost, st));
    BBPD_HIP(hipStreamSynchronize(st));
    if (g_plan && g_plan->bl == BL && g_plan->nB == nB && g_plan->nN == nN && g_plan->hash == h &&
        g_plan->device == dev) {
        *out = g_plan.get();
        return BBPD_OK;
    }
    std::vector<double> hb((size_t)nB * 2);
    if (nB > 0) BBPD_HIP(hipMemcpy(hb.data(), BL, sizeof(double) * 2 * (size_t)nB, hipMemcpyDeviceToHost));
    std::vector<long> rowptr((size_t)nN + 1, 0);
    for (int k = 0; k < nB; k++) {
        long i = (long)hb[k], j = (long)hb[(size_t)nB + k];
        if (i < 1 || i > nN || j < 1 || j > nN) {
            bbpd::set_error("BONDLIST holds a node id outside 1..nNodes");
            return BBPD_ERR_ARG;
        }
        rowptr[i]++;
        rowptr[j]++;
    }
    for (int n = 0; n < nN; n++) rowptr[n + 1] += rowptr[n];
    std::vector<long> fill(rowptr.begin(), rowptr.end() - 1);
    std::vector<long> inc((size_t)rowptr[nN]);
    for (int k = 0; k < nB; k++) {
        long i = (long)hb[k] - 1, j = (long)hb[(size_t)nB + k] - 1;
        inc[fill[i]++] = ((long)k << 1);
        inc[fill[j]++] = ((long)k << 1) | 1;
    }
    auto p = std::make_unique<IncPlan>();
    p->bl = BL;
    p->nB = nB;
    p->nN = nN;
    p->hash = h;
    p->device = dev;
    BBPD_HIP(hipMalloc(&p->rowptr, sizeof(long) * ((size_t)nN + 1)));
    BBPD_HIP(hipMalloc(&p->inc, sizeof(long) * (inc.size() ? inc.size() : 1)));
    BBPD_HIP(hipMemcpy(p->rowptr, rowptr.data(), sizeof(long) * ((size_t)nN + 1), hipMemcpyHostToDevice));
    if (!inc.empty()) BBPD_HIP(hipMemcpy(p->inc, inc.data(), sizeof(long) * inc.size(), hipMemcpyHostToDevice));
    g_plan = std::move(p);
    *out = g_plan.get();
    return BBPD_OK;
}

} // namespace

extern "C" int bbpd_calculatedeformedlength(const double *x, const double *BL, const double *L0, int nBonds,
                                            int nNodes, double *l, double *dX, double *dY, double *dZ, double *s,
                                            void *stream)
{
    BBPD_REQUIRE(nBonds >= 0 && nNodes >= 0, "negative size");
    BBPD_REQUIRE(nBonds == 0 || (x && BL && L0 && l && dX && dY && dZ && s), "null array");
    if (!nBonds) return BBPD_OK;
    k_deformedlength<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(x, BL, L0, nBonds, nNodes,
                                                                                         l, dX, dY, dZ, s);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatebsfexponential(const double *s, const double *s0, const double *sc, double k,
                                            double alpha, double *bsf, double *H, const double *BT, double *flag,
                                            int nBonds, void *stream)
{
    BBPD_REQUIRE(nBonds >= 0, "negative size");
    BBPD_REQUIRE(nBonds == 0 || (s && s0 && sc && bsf && H && BT && flag), "null array");
    if (!nBonds) return BBPD_OK;
    k_bsfexponential<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(s, s0, sc, k, alpha, bsf,
                                                                                         H, BT, flag, nBonds);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatebondforces(double *fX, double *fY, double *fZ, const double *fail, const double *dX,
                                        const double *dY, const double *dZ, const double *l, const double *s,
                                        const double *sp, const double *c, double V, const double *vcf,
                                        const double *H, int nBonds, void *stream)
{
    BBPD_REQUIRE(nBonds >= 0, "negative size");
    BBPD_REQUIRE(nBonds == 0 || (fX && fY && fZ && fail && dX && dY && dZ && l && s && sp && c && vcf && H),
                 "null array");
    if (!nBonds) return BBPD_OK;
    k_bondforces<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(
        fX, fY, fZ, fail, dX, dY, dZ, l, s, sp, c, V, vcf, H, nBonds);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatenodalforces(int nBonds, int nNodes, const double *BL, double *F, const double *fX,
                                         const double *fY, const double *fZ, double *sX, double *sY, double *sZ,
                                         int nThreads, void *stream)
{
    BBPD_REQUIRE(nBonds >= 0 && nNodes >= 0, "negative size");
    BBPD_REQUIRE(nThreads >= 1 && nThreads <= 4096, "nThreads must be in 1..4096");
    BBPD_REQUIRE(nBonds == 0 || (BL && F && fX && fY && fZ), "null array");
    BBPD_REQUIRE((sX && sY && sZ) || (!sX && !sY && !sZ), "scratch arrays must be all given or all NULL");
    if (!nNodes) return BBPD_OK;
    hipStream_t st = bbpd::as_stream(stream);
    std::lock_guard<std::mutex> lk(g_plan_mu);
    IncPlan *p = nullptr;
    int rc = get_plan(BL, nBonds, nNodes, st, &p);
    if (rc) return rc;
    // serial semantics when no scratch is given and one thread is asked for
    int T = (!sX && nThreads == 1) ? 0 : nThreads;
    k_nodal_gather<<<bbpd::grid_for(nNodes, BLOCK), BLOCK, 0, st>>>(nNodes, (long)nBonds, p->rowptr, p->inc, F, fX,
                                                                    fY, fZ, sX, sY, sZ, T);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_timeintegrationeulercromer(double *a, const double *F, double D, double *v, const double *rho,
                                               const double *cons, double DT, double *u, double *x, const double *X,
                                               int nNodes, int NOD, void *stream)
{
    BBPD_REQUIRE(nNodes >= 0 && NOD >= 1 && NOD <= 3, "bad nNodes/NOD");
    BBPD_REQUIRE(nNodes == 0 || (a && F && v && rho && cons && u && x && X), "null array");
    long n = (long)nNodes * NOD;
    if (!n) return BBPD_OK;
    k_eulercromer<<<bbpd::grid_for(n, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(a, F, D, v, rho, cons, DT, u, x, X,
                                                                                 nNodes, NOD);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatebondfailure(double *fail, double ff, const double *BT, const double *s, double csc,
                                         double css, int nBonds, void *stream)
{
    BBPD_REQUIRE(nBonds >= 0, "negative size");
    BBPD_REQUIRE(nBonds == 0 || (fail && BT && s), "null array");
    if (!nBonds) return BBPD_OK;
    k_bondfailure<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(fail, ff, BT, s, csc, css,
                                                                                      nBonds);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatebondsofteningfactor(const double *s, const double *s0, int s0s, const double *sc,
                                                 int scs, double *flag, double *H, const double *BT, int nBonds,
                                                 void *stream)
{
    BBPD_REQUIRE(nBonds >= 0, "negative size");
    BBPD_REQUIRE((s0s == 0 || s0s == 1) && (scs == 0 || scs == 1), "strides must be 0 (scalar) or 1 (array)");
    BBPD_REQUIRE(nBonds == 0 || (s && s0 && sc && flag && H && BT), "null array");
    if (!nBonds) return BBPD_OK;
    k_bilinear<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(s, s0, s0s, sc, scs, flag, H, BT,
                                                                                   nBonds);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatebsftrilinear(const double *s, double s0, double s1, double sc, double *H,
                                          const double *BT, double *flag, int nBonds, void *stream)
{
    BBPD_REQUIRE(nBonds >= 0, "negative size");
    BBPD_REQUIRE(nBonds == 0 || (s && H && BT && flag), "null array");
    if (!nBonds) return BBPD_OK;
    k_trilinear<<<bbpd::grid_for(nBonds, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(s, s0, s1, sc, H, BT, flag,
                                                                                    nBonds);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatedampingcoefficient(const double *F, const double *m, const double *Fp, double DT,
                                                const double *vp, const double *u, int nNodes, double *cn_host,
                                                void *stream)
{
    BBPD_REQUIRE(nNodes >= 0 && cn_host, "bad arguments");
    BBPD_REQUIRE(nNodes == 0 || (F && m && Fp && vp && u), "null array");
    hipStream_t st = bbpd::as_stream(stream);
    int dev = 0;
    BBPD_HIP(hipGetDevice(&dev));
    unsigned nb = bbpd::grid_for(nNodes, BLOCK);
    double *buf = nullptr;
    int rc = device_scratch(dev, sizeof(double) * (2 * (size_t)nb + 4), &buf);
    if (rc) return rc;
    k_damping_partials<<<nb, BLOCK, 0, st>>>(F, m, Fp, DT, vp, u, nNodes, buf + 4);
    k_damping_final<<<1, BLOCK, 0, st>>>(buf + 4, (int)nb, buf);
    BBPD_HIP(hipGetLastError());
    double r[3];
    BBPD_HIP(hipMemcpyAsync(r, buf, sizeof(r), hipMemcpyDeviceToHost, st));
    BBPD_HIP(hipStreamSynchronize(st));
    *cn_host = r[0];
    return BBPD_OK;
}

// ---------------------------------------------------------------------------------------------
// Allocation-free variant of bbpd_calculatedampingcoefficient.  Its block partials and result live
// in `scratch`, a file-scope device array that each GPU holds for the life of the process.  In
// host code the array's name stands for the symbol, not a device address: the address is looked
// up once per device with hipGetSymbolAddress and only that address is handed to kernels and
// copies.  Every call on a GPU shares that GPU's array, so calls that overlap on one GPU race.
// ---------------------------------------------------------------------------------------------
namespace {
constexpr int NOALLOC_BLOCKS = (BBPD_NOALLOC_MAX_NODES + BLOCK - 1) / BLOCK;
constexpr int NOALLOC_RESULT = 4; // doubles at the front: cn, cn1, cn2 (and one spare)
} // namespace

static __device__ __align__(256) double scratch[NOALLOC_RESULT + 2 * NOALLOC_BLOCKS];

namespace {

std::mutex g_scratch_mu;
double *g_scratch_ptr[64] = {nullptr};

int scratch_ptr(double **out)
{
    int dev = 0;
    BBPD_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) {
        bbpd::set_error("device ordinal out of range");
        return BBPD_ERR_ARG;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return bbpd::hip_fail(e, "hipGetSymbolAddress(scratch)");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return BBPD_OK;
}

} // namespace

extern "C" int bbpd_calculatedampingcoefficient_noalloc(const double *F, const double *m, const double *Fp,
                                                        double DT, const double *vp, const double *u, int nNodes,
                                                        double *cn_host, void *stream)
{
    BBPD_REQUIRE(nNodes >= 0 && cn_host, "bad arguments");
    BBPD_REQUIRE(nNodes <= BBPD_NOALLOC_MAX_NODES,
                 "nNodes exceeds BBPD_NOALLOC_MAX_NODES (16777216) of the allocation-free entry point");
    BBPD_REQUIRE(nNodes == 0 || (F && m && Fp && vp && u), "null array");
    hipStream_t st = bbpd::as_stream(stream);
    double *buf = nullptr;
    int rc = scratch_ptr(&buf);
    if (rc) return rc;
    // every partial the final sum reads is written by the first kernel: nothing to clear
    const unsigned nb = bbpd::grid_for(nNodes, BLOCK);
    k_damping_partials<<<nb, BLOCK, 0, st>>>(F, m, Fp, DT, vp, u, nNodes, buf + NOALLOC_RESULT);
    k_damping_final<<<1, BLOCK, 0, st>>>(buf + NOALLOC_RESULT, (int)nb, buf);
    BBPD_HIP(hipGetLastError());
    double r[3];
    BBPD_HIP(hipMemcpyAsync(r, buf, sizeof(r), hipMemcpyDeviceToHost, st));
    BBPD_HIP(hipStreamSynchronize(st));
    *cn_host = r[0];
    return BBPD_OK;
}

extern "C" int bbpd_timeintegrationADR(int it, double DT, double *vf, const double *F, const double *m, double cn,
                                       const double *vp, double *u, const double *cons, int nNodes, void *stream)
{
    BBPD_REQUIRE(nNodes >= 0, "negative size");
    BBPD_REQUIRE(nNodes == 0 || (vf && F && m && vp && u && cons), "null array");
    BBPD_REQUIRE(vf == vp || vf + 3L * nNodes <= vp || vp + 3L * nNodes <= vf,
                 "forward and previous half-step velocities overlap partially");
    long n = 3L * nNodes;
    if (!n) return BBPD_OK;
    k_adr<<<bbpd::grid_for(n, BLOCK), BLOCK, 0, bbpd::as_stream(stream)>>>(it, DT, vf, F, m, cn, vp, u, cons, nNodes);
    BBPD_HIP(hipGetLastError());
    return BBPD_OK;
}

extern "C" int bbpd_calculatecontactforce(const double *family, int nFam, double cx, double cz, double R, double disp,
                                          const double *X, double *x, double *u, double *v, double DT, double V,
                                          const double *rho, int nNodes, double *fz_host, void *stream)
{
    BBPD_REQUIRE(nFam >= 0 && nNodes >= 0 && fz_host, "bad arguments");
    BBPD_REQUIRE(nFam == 0 || (family && X && x && u && v && rho), "null array");
    hipStream_t st = bbpd::as_stream(stream);
    if (!nFam) {
        *fz_host = 0;
        return BBPD_OK;
    }
    int dev = 0;
    BBPD_HIP(hipGetDevice(&dev));
    double *buf = nullptr;
    int rc = device_scratch(dev, sizeof(double) * 4, &buf);
    if (rc) return rc;
    k_contact<<<1, 1024, 0, st>>>(family, nFam, cx, cz, R, disp, X, x, u, v, DT, V, rho, nNodes, buf);
    BBPD_HIP(hipGetLastError());
    double r = 0;
    BBPD_HIP(hipMemcpyAsync(&r, buf, sizeof(r), hipMemcpyDeviceToHost, st));
    BBPD_HIP(hipStreamSynchronize(st));
    *fz_host = r;
    return BBPD_OK;
}
