// kem_axisym.hpp -- FP64 device integrals of the axisymmetric elements: charged rings and conic
// sections (the surface of revolution of a segment about the z axis).  Every integrator policy
// of the reference uses the same analytic integrators for them
// (Core/src/KElectrostaticBoundaryIntegratorFactory.cc:39-73), so they are policy-free here.
//
// Reference (KEMField/Source/...):
//   BoundaryIntegrals/Electrostatic/Analytic/src/KElectrostaticAnalyticRingIntegrator.cc:26-205
//   BoundaryIntegrals/Electrostatic/Analytic/src/KElectrostaticAnalyticConicSectionIntegrator.cc:28-221
//   Math/src/KGaussianQuadrature.cc:18-88        composite rule with end corrections
//   Math/src/KEllipticIntegrals.cc:14-125        K, E, (E - K)/k^2
//
// Element record (kem_layout.h): KG_V .. KG_V+5 = P0, P1 of a conic section (radius = x,
// axial position = z of each end), or P of a ring; KG_A = |P1 - P0|.  Their pairs are always
// "near" (KG_NEAR2 = infinity), so the far passes skip them and the near pass evaluates them.
#ifndef KEM_AXISYM_HPP
#define KEM_AXISYM_HPP

#include <hip/hip_runtime.h>

#include "kem_layout.h"

namespace kem {
namespace dev {

// No FMA contraction anywhere in this file: the conic-section integral near its generating
// line differences nearly equal radii, and fused roundings there move the result by 4e-11
// (measured on the CPU with and without contraction); unfused, it follows the reference's
// roundings to 4e-16.
#define KEM_AX_EXACT _Pragma("clang fp contract(off)")

// K(k) and E(k) with eta = 1 - k^2 by the arithmetic-geometric mean: K = pi / (2 M(1, sqrt eta)),
// E = K (1 - sum_n 2^(n-1) c_n^2), c_0 = k.  The reference evaluates Cody's fits, whose log
// term is log(eta + 1e-19) (KEllipticIntegrals.cc:54); ax_K reproduces that regularisation,
// -log1p(1e-19/eta)/2 away from the exact K (finite at eta = 0).
struct EllKE {
    double K, E;
};

__device__ __forceinline__ EllKE ax_agm(double eta)
{
    KEM_AX_EXACT
    double a = 1., b = sqrt(eta), s = 0.5 * (1. - eta), p = 0.5;
#pragma unroll 1
    for (int it = 0; it < 40 && fabs(a - b) > 1e-16 * a; it++) {
        const double c = 0.5 * (a - b);
        const double an = 0.5 * (a + b);
        b = sqrt(a * b);
        a = an;
        p *= 2.;
        s += p * c * c;
    }
    const double K = (0.5 * 3.141592653589793238462643383279502884) / a;
    return EllKE{K, K * (1. - s)};
}

constexpr double kLn4Ref = 1.386294361119890;  // the reference's ln 4 (KEllipticIntegrals.cc:25)

// K and E at modulus k as the reference computes them (eta = 1 - k*k, KEllipticIntegrals.cc:24,66)
__device__ __forceinline__ EllKE ax_KE(double k)
{
    KEM_AX_EXACT
    const double eta = 1. - k * k;
    if (!(eta > 0.)) return EllKE{kLn4Ref - 0.5 * log(fabs(eta) + 1e-19), 1.};
    EllKE r = ax_agm(eta);
    r.K -= 0.5 * log1p(1e-19 / eta);
    return r;
}

// (E - K)/k^2: the difference above k^2 = 0.8, the power series
// sum_n [(2n-1)!!/(2n)!!]^2 (-1/(2n-1) - 1) (pi/2) k^(2n-2) below (KEllipticIntegrals.cc:101-125)
__device__ __forceinline__ double ax_EK(double k, const EllKE& ke)
{
    KEM_AX_EXACT
    const double k2 = k * k;
    if (k2 > 0.8) return (ke.E - ke.K) / k2;
    double a = 0.5 * 3.141592653589793238462643383279502884, k2n = 1., ek = 0.;
#pragma unroll 1
    for (int n = 1; n <= 900; n++) {
        const double cn = (2. * n - 1.) / (2. * n);
        a *= cn * cn;
        ek += (-a / (2. * n - 1.) - a) * k2n;
        k2n *= k2;
        if (k2n < 1e-16) break;
    }
    return ek;
}

// Ring integrands at distance p along a generating line (par: z, r, zA, rA, zB, rB, L):
// f0 = R K/S (potential), f1 = R/S^3 (-2 R (E-K)/k^2 + dr E/eta) (radial field),
// f2 = dz/S^3 E/eta R (axial field); KElectrostaticAnalyticRingIntegrator.cc:93-205.
template <bool FIELD>
__device__ __forceinline__ void ax_ring_f(double p, const double* par, double f[3])
{
    KEM_AX_EXACT
    const double Z = par[2] + p / par[6] * (par[4] - par[2]);
    const double R = par[3] + p / par[6] * (par[5] - par[3]);
    const double dz = par[0] - Z, dr = par[1] - R, sumr = R + par[1];
    const double eta = (dr * dr + dz * dz) / (sumr * sumr + dz * dz);
    const double S = sqrt(sumr * sumr + dz * dz);
    const double kp = eta > 1. ? 0. : sqrt(1. - eta);
    f[0] = R * ax_KE(kp).K / S;
    if (FIELD) {
        const double k = sqrt(1. - eta);
        const EllKE ke = ax_KE(k);
        const double s3 = S * S * S;
        f[1] = R / s3 * (-2. * R * ax_EK(k, ke) + dr / eta * ke.E);
        f[2] = dz / s3 * ke.E / eta * R;
    } else {
        f[1] = f[2] = 0.;
    }
}

// Glueck's composite rule on n+1 equidistant nodes of [a, b] with end corrections
// (KGaussianQuadrature.cc:18-88).  The end weights are the reference's: they are not fixed by
// polynomial exactness (other end weights are exact to the same degree for every n) and the
// conic-section results depend on the rule's error term.
__device__ __forceinline__ double ax_end_weight(int i, int n)
{
    KEM_AX_EXACT
    const int j = i < n - i ? i : n - i;  // distance to the nearer end
    if (n < 20) {
        switch (j) {
        case 0: return 51. / 160.;
        case 1: return 991. / 720.;
        case 2: return 59. / 90.;
        case 3: return 97. / 80.;
        case 4: return 1333. / 1440.;
        case 5: return 91. / 90.;
        default: return 1.;
        }
    }
    switch (j) {
    case 0: return 0.2803440531305107e0;
    case 1: return 0.1648702325837748e1;
    case 2: return -0.2027449845679092e0;
    case 3: return 0.2797927414021179e1;
    case 4: return -0.9761199294532843e0;
    case 5: return 0.2556499393738999e1;
    case 6: return 0.1451083002645404e0;
    case 7: return 0.1311227127425048e1;
    case 8: return 0.9324249063051143e0;
    case 9: return 0.1006631393298060e1;
    default: return 1.;
    }
}

template <bool FIELD>
__device__ __forceinline__ void ax_quad(double a, double b, const double* par, int n, double q[3])
{
    KEM_AX_EXACT
    if (n < 12) n = 12;
    const double h = (b - a) / n;
    double s0 = 0., s1 = 0., s2 = 0.;
#pragma unroll 1
    for (int i = 0; i <= n; i++) {
        double f[3];
        ax_ring_f<FIELD>(a + h * i, par, f);
        const double w = ax_end_weight(i, n);
        s0 += w * f[0];
        s1 += w * f[1];
        s2 += w * f[2];
    }
    q[0] = s0 * h;
    q[1] = s1 * h;
    q[2] = s2 * h;
}

// radial field Er at cylindrical radius r onto (Ex, Ey): zero on the axis (r < 1e-14)
__device__ __forceinline__ F4 ax_out(double px, double py, double r, double c, const double q[3])
{
    KEM_AX_EXACT
    const double er = c * q[1];
    F4 o;
    o.p = c * q[0];
    o.z = c * q[2];
    if (r < 1.e-14) {
        o.x = o.y = 0.;
    } else {
        o.x = px / r * er;
        o.y = py / r * er;
    }
    return o;
}

// unit charge on a ring of radius R at height Z: prefactor 1/(2 pi^2 eps0 R) (RingIntegrator.cc:26-67)
template <bool FIELD>
__device__ __forceinline__ F4 ring_eval(const double* __restrict__ g, double px, double py, double pz)
{
    KEM_AX_EXACT
    const double R = g[KG_V + 0], Z = g[KG_V + 2];
    const double r = sqrt(px * px + py * py);
    const double par[7] = {pz, r, Z, R, Z, R, 1.};
    double q[3];
    ax_ring_f<FIELD>(0., par, q);
    const double c = 1. / (R * 2. * kPi * kPi * kEps0);
    return ax_out(px, py, r, c, q);
}

// unit surface charge density on a conic section: ring integrands over the generating line,
// 20 / 100 / 500 nodes by |Da + Db - L|/L, and below 5e-4 a geometric split towards the point's
// projection with the logarithmic remainder in closed form (ConicSectionIntegrator.cc:28-221)
template <bool FIELD>
__device__ __forceinline__ F4 conic_eval(const double* __restrict__ g, double px, double py, double pz)
{
    KEM_AX_EXACT
    const double ra = g[KG_V + 0], za = g[KG_V + 2], rb = g[KG_V + 3], zb = g[KG_V + 5], L = g[KG_A];
    const double z = pz, r = sqrt(px * px + py * py);
    const double par[7] = {z, r, za, ra, zb, rb, L};
    const double Da = sqrt((za - z) * (za - z) + (ra - r) * (ra - r));
    const double Db = sqrt((zb - z) * (zb - z) + (rb - r) * (rb - r));
    const double D = fabs(Da + Db - L) / L;
    double q[3] = {0., 0., 0.};
    if (D >= 5.e-2) {
        ax_quad<FIELD>(0., L, par, 20, q);
    } else if (D >= 5.e-3) {
        ax_quad<FIELD>(0., L, par, 100, q);
    } else if (D >= 5.e-4) {
        ax_quad<FIELD>(0., L, par, 500, q);
    } else {
        const double u0 = (zb - za) / L, u1 = (rb - ra) / L;
        const double pp = (z - za) * u0 + (r - ra) * u1;
        const double lead = kLn4Ref + log(2. * r + 1.e-12) + 1.;
        if (pp < L) {
            const double pmin = pp < 0. ? 0. : pp;
            double b = L, a = pmin + (b - pmin) * 0.3, t[3];
#pragma unroll 1
            for (int it = 1; it <= 35; it++) {
                ax_quad<FIELD>(a, b, par, 50, t);
                q[0] += t[0];
                q[1] += t[1];
                q[2] += t[2];
                if (fabs(a - pmin) / L < 1.e-8) break;
                b = a;
                a = pmin + (b - pmin) * 0.3;
            }
            const double h = fabs(a - pmin);
            const double tail = h * (lead - log(h)) / 2.;
            q[0] += tail;
            q[1] += tail;
            q[2] += tail;
        }
        if (pp > 0.) {
            const double pmax = pp > L ? L : pp;
            double b = 0., a = pmax - (pmax - b) * 0.3, t[3];
#pragma unroll 1
            for (int it = 1; it <= 35; it++) {
                ax_quad<FIELD>(b, a, par, 50, t);
                q[0] += t[0];
                q[1] += t[1];
                q[2] += t[2];
                if (fabs(pmax - a) / L < 1.e-8) break;
                b = a;
                a = pmax - (pmax - b) * 0.3;
            }
            const double h = fabs(pmax - a);
            const double tail = h * (lead - log(h)) / 2.;
            q[0] += tail;
            q[1] += tail;
            q[2] += tail;
        }
    }
    return ax_out(px, py, r, 1. / (kPi * kEps0), q);
}

template <bool FIELD>
__device__ __forceinline__ F4 axisym_eval(const double* __restrict__ g, int type, double px, double py, double pz)
{
    KEM_AX_EXACT
    return type == KEM_RING ? ring_eval<FIELD>(g, px, py, pz) : conic_eval<FIELD>(g, px, py, pz);
}

#undef KEM_AX_EXACT

}  // namespace dev
}  // namespace kem

#endif
