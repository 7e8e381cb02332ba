// kem_zh.hip -- device evaluation of the zonal-harmonic expansion (one lane per point).
//
// KZonalHarmonicFieldSolver<KElectrostaticBasis>::ElectricFieldAndPotential
// (FieldSolvers/ZonalHarmonic/Solver/src/KElectrostaticZonalHarmonicFieldSolver.cc:98-124):
// at each node of the container tree the central expansion is tried, then the remote one,
// then the node's subsets (their results summed in order), and a node without subsets
// integrates its elements directly.  Each lane walks the tree with a small explicit stack;
// the source point of an expansion is the one with the smallest convergence ratio over the
// node's list (KZonalHarmonicComputer.hh:84-230 on its first call), computed in FP32 as there.
// The series are the reference's (central :356-495, remote :719-857) in FP64 without FMA
// contraction, so the convergence tests see the reference's terms.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "kem_device.hpp"
#include "kem_policy.hpp"
#include "kem_zh.h"

namespace kem {
namespace {

constexpr int kBlock = 256;
constexpr int kMaxDepth = 24;

#pragma clang fp contract(off)

struct ZhOut {
    double ex, ey, ez, phi;
};

// the source point of the node's list [b, b + n) with the smallest (rho / rho_sp)^2 (central)
// or (rho_sp / rho)^2 (remote); false when it exceeds the convergence ratio squared
__device__ bool pick(const ZHDevice& z, int b, int n, bool cen, double px, double py, double pz, int& best)
{
    if (n <= 0) return false;
    const float p0 = (float)px, p1 = (float)py, fz = (float)pz;
    const float r2 = p0 * p0 + p1 * p1;
    const float crs = (float)z.conv_ratio * (float)z.conv_ratio;
    float bmin = 1.e20f;
    best = b;
    for (int i = b; i < b + n; i++) {
        const double* s = z.sp + 4 * i;
        const float rhof = (float)s[1];
        const double rho2 = (double)(rhof * rhof);
        const float delz = fz - (float)s[0];
        const float v = cen ? (float)((r2 + delz * delz) * (1. / rho2)) : (float)(rho2 / (r2 + delz * delz));
        if (v < bmin) {
            bmin = v;
            best = i;
        }
    }
    return !(bmin > crs);
}

__device__ bool central(const ZHDevice& z, int i, double px, double py, double pz, ZhOut& o)
{
    const double* s = z.sp + 4 * i;
    const double z0 = s[0], sprho = s[1];
    const double* c = z.coeff + (int64_t)s[2];
    const int nc = (int)s[3];
    const double* l0 = z.leg;
    const double* l1 = z.leg + z.nleg;
    const double* l2 = z.leg + 2 * z.nleg;
    const double* l3 = z.leg + 3 * z.nleg;
    const double r = sqrt(px * px + py * py);
    if (r < z.prox && fabs(pz - z0) < z.prox) {
        o.ez = -c[1] / sprho;
        o.ex = o.ey = 0.;
        o.phi = c[0];
        return true;
    }
    const double delz = pz - z0, rho = sqrt(r * r + delz * delz);
    const double u = delz / rho, sv = r / rho, rc = rho / sprho;
    double rcn = rc;
    double Ez = c[1] + c[2] * 2. * rc * u, Er = c[2] * rc, Phi = c[0] + c[1] * rc * u;
    bool cp = false, cz = false, cr = false;
    double lp = 1.e30, lz = 1.e30, lr = 1.e30;
    double p1m1 = u, p1m2 = 1., pp1 = 1., pp2 = 0.;
    for (int n = 2; n < nc - 1; n++) {
        const double p1 = l0[n] * u * p1m1 - l1[n] * p1m2;
        const double p1p = l2[n] * u * pp1 - l3[n] * pp2;
        rcn *= rc;
        const double ezp = c[n + 1] * (n + 1.) * rcn * p1, erp = c[n + 1] * rcn * p1p, php = c[n] * rcn * p1;
        Phi += php;
        Ez += ezp;
        Er += erp;
        const double rp = z.conv_param * fabs(Phi), rz = z.conv_param * fabs(Ez), rr = z.conv_param * fabs(Er);
        if (fabs(php) < rp && fabs(lp) < rp) cp = true;
        if (fabs(ezp) < rz && fabs(lz) < rz) cz = true;
        if ((fabs(erp) < rr && fabs(lr) < rr) || r < z.prox) cr = true;
        if (cp && cz && cr) break;
        lp = php;
        lz = ezp;
        lr = erp;
        p1m2 = p1m1;
        p1m1 = p1;
        pp2 = pp1;
        pp1 = p1p;
    }
    if (!(cp && cz && cr)) return false;
    Ez *= -1. / sprho;
    Er *= sv / sprho;
    o.phi = Phi;
    o.ez = Ez;
    if (r < z.prox) {
        o.ex = o.ey = 0.;
    } else {
        o.ex = px / r * Er;
        o.ey = py / r * Er;
    }
    return true;
}

__device__ bool remote(const ZHDevice& z, int i, double px, double py, double pz, ZhOut& o)
{
    const double* s = z.sp + 4 * i;
    const double z0 = s[0], sprho = s[1];
    const double* c = z.coeff + (int64_t)s[2];
    const int nc = (int)s[3];
    const double* l0 = z.leg;
    const double* l1 = z.leg + z.nleg;
    const double* l2 = z.leg + 2 * z.nleg;
    const double* l3 = z.leg + 3 * z.nleg;
    const double r = sqrt(px * px + py * py);
    const double delz = pz - z0;
    double rho = sqrt(r * r + delz * delz);
    if (rho < 1.e-9) rho = 1.e-9;
    const double u = delz / rho, sv = r / rho, rr = sprho / rho;
    double pn1 = c[0] * rr, zn1 = 0., rn1 = 0.;
    double rrn = rr * rr;
    double pn = c[1] * rrn * u, zn = c[0] * rrn * u, rn = c[0] * rrn;
    double Phi = pn1 + pn, Ez = zn1 + zn, Er = rn1 + rn;
    bool cp = false, cz = false, cr = false;
    double p1m1 = u, p1m2 = 1., pp1 = 1., pp2 = 0.;
    for (int n = 2; n < nc; n++) {
        const double p1 = l0[n] * u * p1m1 - l1[n] * p1m2;
        const double p1p = l2[n] * u * pp1 - l3[n] * pp2;
        rrn *= rr;
        pn = c[n] * rrn * p1;
        zn = c[n - 1] * n * rrn * p1;
        rn = c[n - 1] * rrn * p1p;
        Phi += pn;
        Ez += zn;
        Er += rn;
        const double rp = z.conv_param * fabs(Phi), rz = z.conv_param * fabs(Ez), rq = z.conv_param * fabs(Er);
        if (fabs(pn) < rp && fabs(pn1) < rp) cp = true;
        if (fabs(zn) < rz && fabs(zn1) < rz) cz = true;
        if ((fabs(rn) < rq && fabs(rn1) < rq) || r < z.prox) cr = true;
        if (cp && cz && cr) break;
        pn1 = pn;
        zn1 = zn;
        rn1 = rn;
        p1m2 = p1m1;
        p1m1 = p1;
        pp2 = pp1;
        pp1 = p1p;
    }
    if (!(cp && cz && cr)) return false;
    Ez *= 1. / sprho;
    Er *= sv / sprho;
    o.phi = Phi;
    o.ez = Ez;
    if (r < z.prox) {
        o.ex = o.ey = 0.;
    } else {
        o.ex = px / r * Er;
        o.ey = py / r * Er;
    }
    return true;
}

__global__ __launch_bounds__(kBlock) void k_zh_field(ZHDevice z, const double* __restrict__ geom,
                                                     const double* __restrict__ nodes,
                                                     const double* __restrict__ recs,
                                                     const double* __restrict__ sigma, int policy,
                                                     const double* __restrict__ px, const double* __restrict__ py,
                                                     const double* __restrict__ pz, int64_t npts,
                                                     double* __restrict__ phi, double* __restrict__ ex,
                                                     double* __restrict__ ey, double* __restrict__ ez,
                                                     unsigned long long* __restrict__ counts)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npts) return;
    const double x = px[i], y = py[i], w = pz[i];
    // explicit stack of the nodes being summed over their subsets
    int fnode[kMaxDepth], fnext[kMaxDepth];
    double fa[kMaxDepth][4];
    int d = 0;
    fnode[0] = 0;
    fnext[0] = -1;
    double res[4] = {0., 0., 0., 0.};
    unsigned long long nc = 0, nr = 0, nd = 0;
    while (d >= 0) {
        const int* nd8 = z.node + 8 * fnode[d];
        bool done = false;
        if (fnext[d] < 0) {
            ZhOut o;
            int s;
            if (pick(z, nd8[0], nd8[1], true, x, y, w, s) && central(z, s, x, y, w, o)) {
                nc++;
                done = true;
            } else if (pick(z, nd8[2], nd8[3], false, x, y, w, s) && remote(z, s, x, y, w, o)) {
                nr++;
                done = true;
            } else if (nd8[5] == 0 || d + 1 >= kMaxDepth) {
                // the node's integrating field solver: Kahan sums in container order
                nd++;
                double sm[4] = {0., 0., 0., 0.}, cm[4] = {0., 0., 0., 0.};
                for (int k = nd8[6]; k < nd8[6] + nd8[7]; k++) {
                    const int64_t j = z.elem[k];
                    const dev::F4 v = dev::element_any<true>(geom + j * KG_STRIDE, nodes + j * KEM_NODE_STRIDE,
                                                        recs + j * KEM_REC, policy, x, y, w);
                    dev::kahan_add(v.x, sigma[j], sm[0], cm[0]);
                    dev::kahan_add(v.y, sigma[j], sm[1], cm[1]);
                    dev::kahan_add(v.z, sigma[j], sm[2], cm[2]);
                    dev::kahan_add(v.p, sigma[j], sm[3], cm[3]);
                }
                o.ex = sm[0];
                o.ey = sm[1];
                o.ez = sm[2];
                o.phi = sm[3];
                done = true;
            }
            if (done) {
                res[0] = o.ex;
                res[1] = o.ey;
                res[2] = o.ez;
                res[3] = o.phi;
            } else {
                fnext[d] = 0;
                fa[d][0] = fa[d][1] = fa[d][2] = fa[d][3] = 0.;
            }
        } else if (fnext[d] >= nd8[5]) {
            // every subset summed: the node's result
            for (int q = 0; q < 4; q++) res[q] = fa[d][q];
            done = true;
        }
        if (done) {
            d--;
            if (d >= 0)
                for (int q = 0; q < 4; q++) fa[d][q] = fa[d][q] + res[q];
            continue;
        }
        // descend into the next subset
        const int ch = z.child[nd8[4] + fnext[d]];
        fnext[d]++;
        d++;
        fnode[d] = ch;
        fnext[d] = -1;
    }
    phi[i] = res[3];
    ex[i] = res[0];
    ey[i] = res[1];
    ez[i] = res[2];
    if (counts) {
        if (nc) atomicAdd(&counts[0], nc);
        if (nr) atomicAdd(&counts[1], nr);
        if (nd) atomicAdd(&counts[2], nd);
    }
}

}  // namespace

hipError_t launch_zh_field(hipStream_t st, const ZHDevice& z, const double* geom, const double* nodes,
                           const double* recs, const double* sigma, int policy, const double* px, const double* py,
                           const double* pz, int64_t npts, double* phi, double* ex, double* ey, double* ez,
                           unsigned long long* counts)
{
    if (npts <= 0) return hipSuccess;
    if (z.depth > kMaxDepth) return hipErrorInvalidValue;
    k_zh_field<<<(unsigned)((npts + kBlock - 1) / kBlock), kBlock, 0, st>>>(z, geom, nodes, recs, sigma, policy, px,
                                                                           py, pz, npts, phi, ex, ey, ez, counts);
    return hipGetLastError();
}

}  // namespace kem
