// kem_device.hpp -- FP64 device integrals of the electrostatic BEM path (gfx950).
//
// Per (element, point) arithmetic of KEMField's integrators, reading the precomputed element
// record (kem_layout.h).  Element data is wave-uniform in every kernel here (one element at a
// time for 64 points), so record reads compile to scalar loads and the only divergence is the
// per-lane choice of quadrature rule.
//
// Reference (KEMField/Source/BoundaryIntegrals/Electrostatic/...):
//   Numeric/src/KElectrostaticCubatureTriangleIntegrator.cc:387-665   cubature + dispatch
//   Numeric/src/KElectrostaticCubatureRectangleIntegrator.cc:292-673  rectangle
//   RWG/src/KElectrostaticRWGTriangleIntegrator.cc:8-1240              closed form
//   RWG/src/KElectrostaticRWGRectangleIntegrator.cc:8-1430
//   Surfaces/include/KSolidAngle.hh:58-305                             solid angle
//   Numeric/src/KElectrostaticQuadratureLineSegmentIntegrator.cc       line segment
//   Analytic/src/KElectrostaticAnalyticLineSegmentIntegrator.cc:15-83
#ifndef KEM_DEVICE_HPP
#define KEM_DEVICE_HPP

#include <hip/hip_runtime.h>

#include "kem_layout.h"

namespace kem {
namespace dev {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEps0 = 8.8541878128E-12;
constexpr double kK = 1. / (4. * kPi * kEps0);

// distance-ratio cut-offs: ...CubatureTriangleIntegrator.hh:51-53, ...CubatureRectangleIntegrator.hh:52-54,
// ...QuadratureLineSegmentIntegrator.hh:43-44
constexpr double kTriCub12 = 131.4, kTriCub33 = 30.4, kTriRWG = 3.42;
constexpr double kRectCub12 = 196.2, kRectCub33 = 44.6, kRectRWG = 4.6;
constexpr double kLineQuad16 = 27.7, kLineAna = 2.;

// closed-form thresholds: ...RWGTriangleIntegrator.hh:49-53
constexpr double kMinDistToLine = 1.E-14, kShiftN3 = 1.E-7, kLogArgQuotient = 1.E-6, kTolLambda = 1.E-15;

struct F4 {
    double x, y, z, p;  // Ex, Ey, Ez, phi
};

}  // namespace dev
}  // namespace kem

#include "kem_axisym.hpp"

namespace kem {
namespace dev {

__device__ __forceinline__ double rnorm(double x, double y, double z) { return sqrt(x * x + y * y + z * z); }

// 1/sqrt(a) for a > 0: the hardware estimate (v_rsq_f64, |rel. error| <= 5.2e-8 measured on
// gfx950) plus one third-order correction y' = y + y h (1/2 + 3/8 h), h = 1 - a y^2, which
// lands within 1.4e-16 relative (measured max over 2^20 log-uniform arguments; the library
// 1/sqrt gives 1.65e-16) at 6 instructions instead of a correctly rounded sqrt and an IEEE
// division (~29 instructions).
__device__ __forceinline__ double rsqrt_acc(double a)
{
    const double y = __builtin_amdgcn_rsq(a);
    const double h = __builtin_fma(-a * y, y, 1.0);
    return __builtin_fma(y * h, __builtin_fma(0.375, h, 0.5), y);
}

// ---------------------------------------------------------------- cubature
// pref * sum_g w_g (P - Q_g)/|P - Q_g|^3 and pref * sum_g w_g/|P - Q_g| (..._TriNP / _RectNP)
template <int NQ, bool FIELD>
__device__ __forceinline__ F4 cubature(const double* __restrict__ nd, double pref, double px, double py, double pz)
{
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
#pragma unroll
    for (int g = 0; g < NQ; g++) {
        const double dx = px - nd[4 * g + 0];
        const double dy = py - nd[4 * g + 1];
        const double dz = pz - nd[4 * g + 2];
        const double w = nd[4 * g + 3];
        const double ir = rsqrt_acc(dx * dx + dy * dy + dz * dz);
        if (FIELD) {
            const double f = ir * ir * ir * w;
            s0 += dx * f;
            s1 += dy * f;
            s2 += dz * f;
        }
        s3 += w * ir;
    }
    return F4{pref * s0, pref * s1, pref * s2, pref * s3};
}

// line segment Gauss-Legendre over symmetric node pairs (ElectricFieldAndPotential_nNodes)
template <int HALF, bool FIELD>
__device__ __forceinline__ F4 line_quadrature(const double* __restrict__ nd, double pref, double px, double py,
                                              double pz)
{
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
#pragma unroll
    for (int i = 0; i < HALF; i++) {
        const double* qp = nd + 8 * i;
        const double* qm = qp + 4;
        const double wf = qp[3];
        const double ax = px - qp[0], ay = py - qp[1], az = pz - qp[2];
        const double bx = px - qm[0], by = py - qm[1], bz = pz - qm[2];
        double ia = rsqrt_acc(ax * ax + ay * ay + az * az);
        double ib = rsqrt_acc(bx * bx + by * by + bz * bz);
        s3 += wf * (ia + ib);
        if (FIELD) {
            ia = ia * ia * ia;
            ib = ib * ib * ib;
            s0 += wf * (ia * ax + ib * bx);
            s1 += wf * (ia * ay + ib * by);
            s2 += wf * (ia * az + ib * bz);
        }
    }
    return F4{pref * s0, pref * s1, pref * s2, pref * s3};
}

// thin-wire closed form with the inside-the-wire substitution
// (KElectrostaticAnalyticLineSegmentIntegrator::Potential/ElectricField)
__device__ __forceinline__ F4 line_analytic(const double* __restrict__ g, double px, double py, double pz)
{
    const double* p0 = g + KG_V;
    const double* p1 = g + KG_V + 3;
    const double L = g[KG_A], dia = g[KG_B];
    double Da = rnorm(p0[0] - px, p0[1] - py, p0[2] - pz);
    double Db = rnorm(p1[0] - px, p1[1] - py, p1[2] - pz);
    double DaP = Da, DbP = Db, DaE = Da, DbE = Db;
    if (!((Da + Db) > (L + dia))) {
        const double* u = g + KG_U;
        const double p = (px - p0[0]) * u[0] + (py - p0[1]) * u[1] + (pz - p0[2]) * u[2];
        if (!(p < (-dia * .5) || p > (L + dia * .5))) {
            const double fx = p0[0] + p * u[0], fy = p0[1] + p * u[1], fz = p0[2] + p * u[2];
            if (!(rnorm(px - fx, py - fy, pz - fz) >= dia * .5)) {
                const double ax = p0[0] - fx, ay = p0[1] - fy, az = p0[2] - fz;
                const double bx = p1[0] - fx, by = p1[1] - fy, bz = p1[2] - fz;
                const double a2 = ax * ax + ay * ay + az * az, b2 = bx * bx + by * by + bz * bz;
                DaP = sqrt(a2 + dia * dia * .25);
                DbP = sqrt(b2 + dia * dia * .25);
                DaE = sqrt(a2);
                DbE = sqrt(b2);
            }
        }
    }
    F4 r;
    r.p = dia / (4. * kEps0) * log((DaP + DbP + L) / (DaP + DbP - L));
    double den = (DaE * (DaE + DbE + L) * (DaE + DbE - L) * DbE) * 4. * kEps0;
    den = -1. / den;
    const double s = 2. * L * dia * den;
    r.x = (p1[0] * DaE - px * (DaE + DbE) + p0[0] * DbE) * s;
    r.y = (p1[1] * DaE - py * (DaE + DbE) + p0[1] * DbE) * s;
    r.z = (p1[2] * DaE - pz * (DaE + DbE) + p0[2] * DbE) * s;
    return r;
}

// ---------------------------------------------------------------- solid angle
// Van Oosterom-Strackee: Omega = 2 atan2(|u0.(u1 x u2)|, 1 + u0.u1 + u0.u2 + u1.u2), signed by h
__device__ __forceinline__ double tri_omega(const double* __restrict__ V, double px, double py, double pz)
{
    double u[3][3];
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const double dx = V[3 * v] - px, dy = V[3 * v + 1] - py, dz = V[3 * v + 2] - pz;
        const double im = rsqrt_acc(dx * dx + dy * dy + dz * dz);
        u[v][0] = im * dx;
        u[v][1] = im * dy;
        u[v][2] = im * dz;
    }
    const double x = 1. + (u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2]) +
                     (u[0][0] * u[2][0] + u[0][1] * u[2][1] + u[0][2] * u[2][2]) +
                     (u[1][0] * u[2][0] + u[1][1] * u[2][1] + u[1][2] * u[2][2]);
    const double a0 = u[1][1] * u[2][2] - u[1][2] * u[2][1];
    const double a1 = u[1][2] * u[2][0] - u[1][0] * u[2][2];
    const double a2 = u[1][0] * u[2][1] - u[1][1] * u[2][0];
    const double y = fabs(u[0][0] * a0 + u[0][1] * a1 + u[0][2] * a2);
    return fabs(2. * atan2(y, x));
}

// KSolidAngle::SolidAngleTriangleAsArray (KSolidAngle.hh:58-144)
__device__ __forceinline__ double solid_angle_tri(const double* __restrict__ g, double px, double py, double pz)
{
    const double cx = px - g[KG_HX], cy = py - g[KG_HY], cz = pz - g[KG_HZ];
    const double h = g[KG_NX] * cx + g[KG_NY] * cy + g[KG_NZ] * cz;
    double res = (sqrt(cx * cx + cy * cy + cz * cz) <= 1.e-12) ? 2. * kPi : tri_omega(g + KG_V, px, py, pz);
    return h < 0. ? -res : res;
}

// KSolidAngle::SolidAngleRectangleAsArray (KSolidAngle.hh:166-305): two triangles P0P1P2 and P2P3P0;
// in the plane: 2 pi inside, 0 outside
__device__ __forceinline__ double solid_angle_rect(const double* __restrict__ g, double px, double py, double pz)
{
    const double h = g[KG_NX] * (px - g[KG_HX]) + g[KG_NY] * (py - g[KG_HY]) + g[KG_NZ] * (pz - g[KG_HZ]);
    double res;
    if (fabs(h) < 1.e-12) {
        const double* V = g + KG_V;
        const double* s1 = g + KG_SV;
        const double* s2 = g + KG_SV + 3;
        const double dx = V[0] - px, dy = V[1] - py, dz = V[2] - pz;
        const double A = g[KG_A], B = g[KG_B];
        const double l1 = (-1) * (dx * s1[0] + dy * s1[1] + dz * s1[2]) / (A * A);
        const double l2 = (-1) * (dx * s2[0] + dy * s2[1] + dz * s2[2]) / (B * B);
        res = (l1 > 0. && l1 < 1. && l2 > 0. && l2 < 1.) ? 2. * kPi : 0.;
    } else {
        const double* V = g + KG_V;
        const double t1[9] = {V[0], V[1], V[2], V[3], V[4], V[5], V[6], V[7], V[8]};
        const double t2[9] = {V[6], V[7], V[8], V[9], V[10], V[11], V[0], V[1], V[2]};
        res = tri_omega(t1, px, py, pz) + tri_omega(t2, px, py, pz);
    }
    return h < 0. ? -res : res;
}

// ---------------------------------------------------------------- closed form (RWG)
// Line-integral term IqLFieldAndPotential: returns (field part, potential part).  Edges are
// visited in a rolled loop that re-reads the per-edge geometry (scalar loads) and recomputes
// the two vertex distances of each edge, keeping register pressure low.
template <int NE, bool FIELD>
__device__ __forceinline__ F4 edge_logs(const double* __restrict__ g, double px, double py, double pz, int count,
                                        int line, double dist)
{
    F4 acc{0., 0., 0., 0.};
#pragma unroll 1
    for (int e = 0; e < NE; e++) {
        const int e1 = (e + 1 == NE) ? 0 : e + 1;
        const double* u = g + KG_U + 3 * e;
        const double* m = g + KG_M + 3 * e;
        const double* mid = g + KG_MID + 3 * e;
        const double* va = g + KG_V + 3 * e;
        const double* vb = g + KG_V + 3 * e1;
        const double ax = va[0] - px, ay = va[1] - py, az = va[2] - pz;
        const double bx = vb[0] - px, by = vb[1] - py, bz = vb[2] - pz;
        const double rM = sqrt(ax * ax + ay * ay + az * az);
        const double rP = sqrt(bx * bx + by * by + bz * bz);
        const double sM = ax * u[0] + ay * u[1] + az * u[2];
        const double sP = bx * u[0] + by * u[1] + bz * u[2];
        bool have = false;
        double val = 0.;
        if (line != e) {
            double nom, den;
            if ((rM + sM) > (rP - sP)) {
                nom = rP + sP;
                den = rM + sM;
            } else {
                nom = rM - sM;
                den = rP - sP;
            }
            val = log(nom) - log(den);
            have = true;
        } else if (count == 1 && fabs(dist / sM) < kLogArgQuotient) {
            // point on the edge line: second-order Taylor argument (LogArgTaylor)
            double q = fabs(dist / sM);
            q = q < 1.e-14 ? 1.e-14 : q;
            val = log(rP + sP) - log(0.5 * fabs(sM) * q * q);
            have = true;
        }
        if (have) {
            const double t = m[0] * (px - mid[0]) + m[1] * (py - mid[1]) + m[2] * (pz - mid[2]);
            if (FIELD) {
                acc.x += m[0] * val;
                acc.y += m[1] * val;
                acc.z += m[2] * val;
            }
            acc.p += t * val;
        }
    }
    return acc;
}

template <int NE>
__device__ __forceinline__ double solid_angle(const double* __restrict__ g, double px, double py, double pz)
{
    return NE == 3 ? solid_angle_tri(g, px, py, pz) : solid_angle_rect(g, px, py, pz);
}

template <int NE, bool FIELD>
__device__ __forceinline__ F4 closed_form_at(const double* __restrict__ g, double px, double py, double pz, int count,
                                             int line, double dist)
{
    const double nx = g[KG_NX], ny = g[KG_NY], nz = g[KG_NZ];
    const double h = nx * (px - g[KG_HX]) + ny * (py - g[KG_HY]) + nz * (pz - g[KG_HZ]);
    const double om = solid_angle<NE>(g, px, py, pz);
    const F4 iL = edge_logs<NE, FIELD>(g, px, py, pz, count, line, dist);
    return F4{nx * om + iL.x, ny * om + iL.y, nz * om + iL.z, (-h * om) - iL.p};
}

// KElectrostaticRWG{Triangle,Rectangle}Integrator::ElectricFieldAndPotential
template <int NE, bool FIELD>
__device__ __forceinline__ F4 closed_form(const double* __restrict__ g, double px, double py, double pz)
{
    int count = 0, line = 99;
    double dmin = 0.;
#pragma unroll 1
    for (int e = 0; e < NE; e++) {
        const double dx = g[KG_V + 3 * e] - px, dy = g[KG_V + 3 * e + 1] - py, dz = g[KG_V + 3 * e + 2] - pz;
        const double* s = g + KG_SV + 3 * e;
        const double il = g[KG_IL + e];
        const double cx = s[1] * dz - s[2] * dy, cy = s[2] * dx - s[0] * dz, cz = s[0] * dy - s[1] * dx;
        const double dl = sqrt(cx * cx + cy * cy + cz * cz) * il;
        const double lam = ((-dx * s[0]) + (-dy * s[1]) + (-dz * s[2])) * (il * il);
        if (dl < kMinDistToLine && lam >= -kTolLambda && lam <= (1. + kTolLambda)) {
            dmin = dl;
            count++;
            line = e;
        }
    }
    // on a vertex (count == 2): mean of the points moved by +-1e-7 along the normal, without
    // edge correction; otherwise one evaluation at P
    const int nrep = count == 2 ? 2 : 1;
    const double sh = count == 2 ? kShiftN3 : 0.;
    if (count == 2) {
        count = 9;
        line = 9;
        dmin = 9.;
    }
    F4 acc{0., 0., 0., 0.};
#pragma unroll 1
    for (int r = 0; r < nrep; r++) {
        const double sg = r == 0 ? sh : -sh;
        const F4 a = closed_form_at<NE, FIELD>(g, px + sg * g[KG_NX], py + sg * g[KG_NY], pz + sg * g[KG_NZ], count,
                                               line, dmin);
        acc.x += a.x;
        acc.y += a.y;
        acc.z += a.z;
        acc.p += a.p;
    }
    const double f = nrep == 2 ? kK * 0.5 : kK;
    return F4{f * acc.x, f * acc.y, f * acc.z, f * acc.p};
}

// ---------------------------------------------------------------- dispatch
// Distance ratio |centre - P| * (1/size), as KElectrostaticCubature*Integrator computes it.
__device__ __forceinline__ double distance_ratio(const double* __restrict__ g, double px, double py, double pz)
{
    const double dx = g[KG_CX] - px, dy = g[KG_CY] - py, dz = g[KG_CZ] - pz;
    // explicit FMAs: every kernel classifies a pair with identical bits, so the far and near
    // passes partition the pairs exactly
    return sqrt(__builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx))) * g[KG_INVSIZE];
}

// One (element, point) pair under the "numeric" policy (default) or the closed form everywhere
// ("rwg").  Returns the unit-charge (E, phi).
template <bool FIELD>
__device__ __forceinline__ F4 element_eval(const double* __restrict__ g, const double* __restrict__ nd, int policy,
                                           double px, double py, double pz)
{
    const int type = (int)g[KG_TYPE];
    const double dr = distance_ratio(g, px, py, pz);
    const double pref = g[KG_PREF];
    if (type == KEM_TRIANGLE) {
        if (policy == 0) {
            if (dr > kTriCub12) return cubature<7, FIELD>(nd + 4 * KEM_N7, pref, px, py, pz);
            if (dr > kTriCub33) return cubature<12, FIELD>(nd + 4 * KEM_N12, pref, px, py, pz);
            if (dr > kTriRWG) return cubature<33, FIELD>(nd + 4 * KEM_N33, pref, px, py, pz);
        }
        return closed_form<3, FIELD>(g, px, py, pz);
    }
    if (type == KEM_RECTANGLE) {
        if (policy == 0) {
            if (dr > kRectCub12) return cubature<7, FIELD>(nd + 4 * KEM_N7, pref, px, py, pz);
            if (dr > kRectCub33) return cubature<12, FIELD>(nd + 4 * KEM_N12, pref, px, py, pz);
            if (dr > kRectRWG) return cubature<33, FIELD>(nd + 4 * KEM_N33, pref, px, py, pz);
        }
        return closed_form<4, FIELD>(g, px, py, pz);
    }
    if (policy == 0) {
        if (dr > kLineQuad16) return line_quadrature<2, FIELD>(nd + 4 * KEM_L4, pref, px, py, pz);
        if (dr > kLineAna) return line_quadrature<8, FIELD>(nd + 4 * KEM_L16, pref, px, py, pz);
    }
    return line_analytic(g, px, py, pz);
}

// Squared distance |centre - P|^2 with explicit FMAs: every field kernel classifies a pair
// from these identical bits, so the far and near passes partition the pairs exactly.
__device__ __forceinline__ double dist2(const double* __restrict__ g, double px, double py, double pz)
{
    const double dx = g[KG_CX] - px, dy = g[KG_CY] - py, dz = g[KG_CZ] - pz;
    return __builtin_fma(dz, dz, __builtin_fma(dy, dy, dx * dx));
}

// Far/near split of the numeric policy for the field kernels: a pair is "near" when the
// dispatch sends it to a closed form (dr <= 3.42 triangle, 4.6 rectangle, 2 segment; every
// pair under the rwg policy).  The thresholds are compared as squared distances against
// (dr_cut * size)^2 precomputed per element (kem_layout.h), so no square root is taken per
// pair; this agrees with the reference's `dr > cut` except for distances within an ulp of a
// cut-off.  The far pass evaluates only cubature / Gauss-Legendre pairs and stays
// register-lean; the near pass evaluates only closed-form pairs.
__device__ __forceinline__ bool is_near2(const double* __restrict__ g, double r2, int policy)
{
    return policy != 0 || !(r2 > g[KG_NEAR2]);
}

// Far pass of one pair (reference ladder, FP64).  Called only for pairs outside the near set.
// FIELD = false: the potential only (collocation rows of Dirichlet elements).
template <bool FIELD = true>
__device__ __forceinline__ F4 element_far(const double* __restrict__ g, const double* __restrict__ nd, int type,
                                          double r2, double px, double py, double pz)
{
    const double pref = g[KG_PREF];
    if (type == KEM_LINESEGMENT) {
        if (r2 > g[KG_REF7_2]) return line_quadrature<2, FIELD>(nd + 4 * KEM_L4, pref, px, py, pz);
        return line_quadrature<8, FIELD>(nd + 4 * KEM_L16, pref, px, py, pz);
    }
    if (r2 > g[KG_REF7_2]) return cubature<7, FIELD>(nd + 4 * KEM_N7, pref, px, py, pz);
    if (r2 > g[KG_REF12_2]) return cubature<12, FIELD>(nd + 4 * KEM_N12, pref, px, py, pz);
    return cubature<33, FIELD>(nd + 4 * KEM_N33, pref, px, py, pz);
}

// element_far with one rule for the whole wavefront: the most nodes the reference's ladder
// picks for any active lane (the lanes of near pairs are inactive here).  The lanes of a
// wavefront straddling a cut-off then run one rule instead of both; a lane given more nodes than
// its own rule gets the same integral to within the rules' truncation, below 1e-15 relative
// beyond their cut-offs.  Used by the assembly, whose census counts 16.6 node evaluations per
// lane and pair when each lane runs its own rule, 14.9 with one rule per wavefront, 13.6 useful.
template <bool FIELD = true>
__device__ __forceinline__ F4 element_far_wave(const double* __restrict__ g, const double* __restrict__ nd, int type,
                                               double r2, double px, double py, double pz)
{
    const double pref = g[KG_PREF];
    if (type == KEM_LINESEGMENT) {
        if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_REF7_2])) == 0)
            return line_quadrature<2, FIELD>(nd + 4 * KEM_L4, pref, px, py, pz);
        return line_quadrature<8, FIELD>(nd + 4 * KEM_L16, pref, px, py, pz);
    }
    if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_REF12_2])) != 0)
        return cubature<33, FIELD>(nd + 4 * KEM_N33, pref, px, py, pz);
    if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_REF7_2])) != 0)
        return cubature<12, FIELD>(nd + 4 * KEM_N12, pref, px, py, pz);
    return cubature<7, FIELD>(nd + 4 * KEM_N7, pref, px, py, pz);
}

// The potential of cubature<NQ, false> at T points per lane from one pass over the nodes of the
// assembly's node table (KEM_ANODE_STRIDE layout): (dx, dy, dz)[t] = P_t - c, d2[t] = |P_t - c|^2,
// r^2 = d2 + |q'|^2 - 2 (P - c) . q'.  Beyond the near cut-offs |P - c| is at least 2 element
// sizes and |q'| about one, so the cancellation costs at most a few ulps of r^2.
template <int NQ, int T>
__device__ __forceinline__ void cubature_pot_n(const double* __restrict__ an, double pref, const double (&dx)[T],
                                               const double (&dy)[T], const double (&dz)[T], const double (&d2)[T],
                                               double (&v)[T])
{
    double s[T];
#pragma unroll
    for (int t = 0; t < T; t++) s[t] = 0.;
#pragma unroll
    for (int g = 0; g < NQ; g++) {
        const double* q = an + 5 * g;
#pragma unroll
        for (int t = 0; t < T; t++) {
            const double r2 = __builtin_fma(dx[t], q[0], __builtin_fma(dy[t], q[1], __builtin_fma(dz[t], q[2], d2[t] + q[3])));
            s[t] = __builtin_fma(q[4], rsqrt_acc(r2), s[t]);
        }
    }
#pragma unroll
    for (int t = 0; t < T; t++) v[t] = pref * s[t];
}

// element_far_wave's potential at T points per lane (squared distances d2[t] from the element
// centre, offsets (dx, dy, dz)[t] = P_t - c; a point takes part in the rule choice only if its
// pair is far, far[t]); segments from the cubature nodes nd, surfaces from the table an
template <int T>
__device__ __forceinline__ void element_far_wave_pot_n(const double* __restrict__ g, const double* __restrict__ nd,
                                                       const double* __restrict__ an, int type,
                                                       const double (&d2)[T], const bool (&far)[T],
                                                       const double (&dx)[T], const double (&dy)[T],
                                                       const double (&dz)[T], const double (&px)[T],
                                                       const double (&py)[T], const double (&pz)[T], double (&v)[T])
{
    const double pref = g[KG_PREF];
    auto needs = [&](double cut) {
        bool b = false;
#pragma unroll
        for (int t = 0; t < T; t++) b = b || (far[t] && !(d2[t] > cut));
        return __builtin_amdgcn_ballot_w64(b) != 0;
    };
    if (type == KEM_LINESEGMENT) {
        const bool l16 = needs(g[KG_REF7_2]);
#pragma unroll
        for (int t = 0; t < T; t++)
            v[t] = (l16 ? line_quadrature<8, false>(nd + 4 * KEM_L16, pref, px[t], py[t], pz[t])
                        : line_quadrature<2, false>(nd + 4 * KEM_L4, pref, px[t], py[t], pz[t]))
                       .p;
        return;
    }
    if (needs(g[KG_REF12_2])) return cubature_pot_n<33, T>(an + 5 * KEM_N33, pref, dx, dy, dz, d2, v);
    if (needs(g[KG_REF7_2])) return cubature_pot_n<12, T>(an + 5 * KEM_N12, pref, dx, dy, dz, d2, v);
    return cubature_pot_n<7, T>(an + 5 * KEM_N7, pref, dx, dy, dz, d2, v);
}

// Near pass of one pair (closed forms).
__device__ __forceinline__ F4 element_near(const double* __restrict__ g, int type, double px, double py, double pz)
{
    if (type == KEM_TRIANGLE) return closed_form<3, true>(g, px, py, pz);
    if (type == KEM_RECTANGLE) return closed_form<4, true>(g, px, py, pz);
    return line_analytic(g, px, py, pz);
}

// Near pass of one pair when the mesh holds conic sections or rings (kernels instantiated for
// such meshes only, so the triangle kernels keep their register budget).
__device__ __forceinline__ F4 element_near_any(const double* __restrict__ g, int type, double px, double py,
                                               double pz)
{
    if (type >= KEM_CONICSECTION) return axisym_eval<true>(g, type, px, py, pz);
    return element_near(g, type, px, py, pz);
}

// ---------------------------------------------------------------- mixed precision (FP32 far field)
// Rule ladder for the FP32 far field (thresholds KG_MIX7_2 / KG_MIX12_2 set in kem_prep.cpp).
// The reference's cut-offs keep every pair at machine precision; at FP32 accuracy the
// truncation error only has to stay well below FP32 rounding (~1e-7).  Measured max relative
// truncation error (potential and field, 23 triangles / 20 rectangles of aspect <= 4, random
// directions, against the closed form): triangle 6-point (degree 4) 1.5e-8 at dr = 12,
// 7-point 1.1e-8 at dr = 8, 12-point 8.4e-9 at dr = 4, 33-point < 1e-10 at dr >= 3.42;
// rectangle 7-point 1e-8 at dr = 12, 12-point 4.4e-9 at dr = 6.  Near pairs (the same
// closed-form set as the FP64 path) stay FP64.

// FP32 node sums relative to the element centre: (dx, dy, dz) = P - c in FP32 (computed in
// FP64 and rounded once), nodes q = Q - c in FP32, so P - Q = (P - c) - q carries the relative
// accuracy of FP32 in |P - Q| instead of losing digits to the absolute coordinates.
template <int NQ>
__device__ __forceinline__ void cubature32(const float* __restrict__ nd, float dx, float dy, float dz, float& s0,
                                           float& s1, float& s2, float& s3)
{
    s0 = s1 = s2 = s3 = 0.f;
#pragma unroll
    for (int g = 0; g < NQ; g++) {
        const float ex = dx - nd[4 * g + 0];
        const float ey = dy - nd[4 * g + 1];
        const float ez = dz - nd[4 * g + 2];
        const float w = nd[4 * g + 3];
        const float ir = __builtin_amdgcn_rsqf(ex * ex + ey * ey + ez * ez);
        const float wi = w * ir;
        const float f = wi * ir * ir;
        s0 += ex * f;
        s1 += ey * f;
        s2 += ez * f;
        s3 += wi;
    }
}

// Far pass of one pair in FP32 without the prefactor (the caller scales by pref * sigma in
// FP64).  Called only for pairs outside the near set; r2 = dist2 (FP64).
__device__ __forceinline__ void element_far32(const double* __restrict__ g, const float* __restrict__ nd, int type,
                                              double r2, float dx, float dy, float dz, float& s0, float& s1, float& s2,
                                              float& s3)
{
    if (type == KEM_LINESEGMENT) {
        // node pairs stored consecutively, weights already scaled by the half length
        if (r2 > g[KG_MIX7_2]) return cubature32<4>(nd + 4 * KEM_L4, dx, dy, dz, s0, s1, s2, s3);
        return cubature32<16>(nd + 4 * KEM_L16, dx, dy, dz, s0, s1, s2, s3);
    }
    if (r2 > g[KG_MIX6_2]) return cubature32<6>(nd + 4 * KEM_N6, dx, dy, dz, s0, s1, s2, s3);
    if (r2 > g[KG_MIX7_2]) return cubature32<7>(nd + 4 * KEM_N7, dx, dy, dz, s0, s1, s2, s3);
    if (r2 > g[KG_MIX12_2]) return cubature32<12>(nd + 4 * KEM_N12, dx, dy, dz, s0, s1, s2, s3);
    return cubature32<33>(nd + 4 * KEM_N33, dx, dy, dz, s0, s1, s2, s3);
}

// element_far32 with one rule for the whole wavefront (all active lanes evaluate the same
// element): the highest-degree rule the per-pair ladder picks for any active lane.  Lanes whose
// ladder would pick a lower rule get one of higher degree, whose error at their distance is
// below the lower rule's; divergent lanes would instead run every picked rule in turn.
__device__ __forceinline__ void element_far32_wave(const double* __restrict__ g, const float* __restrict__ nd,
                                                   int type, double r2, float dx, float dy, float dz, float& s0,
                                                   float& s1, float& s2, float& s3)
{
    if (type == KEM_LINESEGMENT) {
        if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_MIX7_2])) == 0)
            return cubature32<4>(nd + 4 * KEM_L4, dx, dy, dz, s0, s1, s2, s3);
        return cubature32<16>(nd + 4 * KEM_L16, dx, dy, dz, s0, s1, s2, s3);
    }
    if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_MIX12_2])) != 0)
        return cubature32<33>(nd + 4 * KEM_N33, dx, dy, dz, s0, s1, s2, s3);
    if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_MIX7_2])) != 0)
        return cubature32<12>(nd + 4 * KEM_N12, dx, dy, dz, s0, s1, s2, s3);
    if (__builtin_amdgcn_ballot_w64(!(r2 > g[KG_MIX6_2])) != 0)
        return cubature32<7>(nd + 4 * KEM_N7, dx, dy, dz, s0, s1, s2, s3);
    return cubature32<6>(nd + 4 * KEM_N6, dx, dy, dz, s0, s1, s2, s3);
}

// Kahan accumulation of sigma * v into (s, c), component-wise, as the reference sums
// (KElectrostaticIntegratingFieldSolver.hh:325-333)
__device__ __forceinline__ void kahan_add(double v, double sj, double& s, double& c)
{
    const double y = v * sj - c;
    const double t = s + y;
    c = (t - s) - y;
    s = t;
}

}  // namespace dev
}  // namespace kem

#endif
