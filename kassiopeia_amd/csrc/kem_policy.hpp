// kem_policy.hpp -- device integrals of the "analytic" and "reference" integrator policies.
//
// These policies are selectable by name (KElectrostaticBoundaryIntegratorFactory.cc:42-77) and
// serve tests and cross-checks rather than the production path, so they work from the raw
// 11-double shape record, recomputing per pair what the reference recomputes per call, with
// the reference's branch structure, thresholds and expression order:
//   analytic  -- Analytic/src/KElectrostaticAnalyticTriangleIntegrator.cc:13-506,
//                Analytic/src/KElectrostaticAnalyticRectangleIntegrator.cc:25-205,
//                segments: Analytic/src/KElectrostaticAnalyticLineSegmentIntegrator.cc:15-83
//   reference -- Reference/src/KElectrostaticBiQuadratureTriangleIntegrator.cc:24-150,
//                Reference/src/KElectrostaticBiQuadratureRectangleIntegrator.cc:18-104,
//                Reference/src/KElectrostatic256NodeQuadratureLineSegmentIntegrator.cc:192-270
// (paths under KEMField/Source/BoundaryIntegrals/Electrostatic/).
#ifndef KEM_POLICY_HPP
#define KEM_POLICY_HPP

#include "kem_device.hpp"

// Every function here is compiled without FMA contraction, as the reference is: the analytic
// primitives are ill-conditioned (differences of logarithms and arctangents of nearly equal
// arguments), and fusing a*b + c changes their results by up to ~1e-11 relative, which is
// the size of the rounding differences they amplify.
#define KEM_EXACT _Pragma("clang fp contract(off)")

namespace kem {
namespace dev {
namespace pol {

constexpr double kPiOverTwo = kPi / 2.;

__device__ __forceinline__ int sgn(double v) { return (v > 0.) - (v < 0.); }
__device__ __forceinline__ double dot(const double* a, const double* b)
{
    KEM_EXACT
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
__device__ __forceinline__ double mag(const double* a)
{
    KEM_EXACT
    return sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
}

// unit(N1 x N2) of a record (KTriangle.cc:16, KThreeVector::Unit divides by the magnitude)
__device__ __forceinline__ void normal(const double* d, double* n)
{
    KEM_EXACT
    const double c[3] = {d[6] * d[10] - d[7] * d[9], d[7] * d[8] - d[5] * d[10], d[5] * d[9] - d[6] * d[8]};
    const double m = mag(c);
    for (int k = 0; k < 3; k++) n[k] = c[k] / m;
}

// ------------------------------------------------------------ analytic triangle primitives
__device__ double F1(double a, double b, double u)
{
    KEM_EXACT
    return u * asinh((a + b * u) / sqrt(u * u + 1.));
}

__device__ void g12(double a, double b, double u1, double u2, double& g1, double& g2)
{
    KEM_EXACT
    g1 = (sqrt(b * b + 1.) * sqrt(a * a + 2 * a * b * u1 + (b * b + 1.) * u1 * u1 + 1.) + b * (a + b * u1) + u1);
    g2 = (sqrt(b * b + 1.) * sqrt(a * a + 2 * a * b * u2 + (b * b + 1.) * u2 * u2 + 1.) + b * (a + b * u2) + u2);
    if (g1 <= 0) g1 = -(1 + a * a + b * b) / (2 * (b * b + 1) * u1);
    if (g2 <= 0) g2 = -(1 + a * a + b * b) / (2 * (b * b + 1) * u2);
}

__device__ double I3(double a, double b, double u1, double u2)
{
    KEM_EXACT
    double g1, g2;
    g12(a, b, u1, u2, g1, g2);
    if (fabs(g1) < 1.e-12 && fabs(a) < 1.e-14) g1 = 1.e-12;
    if (fabs(g2) < 1.e-12 && fabs(a) < 1.e-14) g2 = 1.e-12;
    return a / sqrt(b * b + 1.) * log(fabs(g2 / g1));
}

__device__ double I3p(double a, double b, double u1, double u2)
{
    KEM_EXACT
    double g1, g2;
    g12(a, b, u1, u2, g1, g2);
    return 1. / sqrt(b * b + 1.) * log(fabs(g2 / g1));
}

__device__ double I4t(double alpha, double gamma, double q2, double prefac, double t1, double t2)
{
    KEM_EXACT
    const double q = sqrt(q2);
    const double g1 = sqrt(gamma * t1 * t1 + alpha);
    const double g2 = sqrt(gamma * t2 * t2 + alpha);
    if (t1 > 1.e15 || t2 > 1.e15) {
        if (t2 < 1.e15) return (prefac * 1. / q * (atan(g2 / q) - kPiOverTwo));
        if (t1 < 1.e15) return (prefac * 1. / q * (kPiOverTwo - atan(g1 / q)));
        return 0.;
    }
    return prefac * 1. / q * atan(q * (g2 - g1) / (q2 + g1 * g2));
}

// The reference nudges u1, u2 or a by 1e-14 and re-enters itself; the loop takes the same
// steps (each nudge leaves its 1e-14 window within two steps; the bound only guards the loop).
__device__ double I4(double a, double b, double u1, double u2)
{
    KEM_EXACT
    for (int it = 0; it < 64; it++) {
        if (fabs(u1 - b / a) < 1.e-14) {
            u1 = (u2 > b / a) ? u1 + 1.e-14 : u1 - 1.e-14;
            continue;
        }
        if (fabs(u2 - b / a) < 1.e-14) {
            u2 = (u1 > b / a) ? u2 + 1.e-14 : u2 - 1.e-14;
            continue;
        }
        if (fabs(a) < 1.e-14) {
            a = (a > 0) ? a + 1.e-14 : a - 1.e-14;
            continue;
        }
        break;
    }
    const double alpha = 1. + (a * a) / (b * b);
    const double gamma = (a * a + b * b) * (a * a + b * b + 1.) / (b * b);
    const double q2 = (a * a + b * b) * (a * a + b * b) / (b * b);
    const double prefac = (a * a / b + b);
    const double t1 = (a * u1 != b) ? (b * u1 + a) / (a * u1 - b) : 1.e15;
    const double t2 = (a * u2 != b) ? (b * u2 + a) / (a * u2 - b) : 1.e15;
    double sign = 1.;
    if (a < 0.) sign = -sign;
    if (b < 0.) sign = -sign;
    if (u1 > b / a) sign = -sign;
    if (sgn(u1 - b / a) != sgn(u2 - b / a))
        return sign * (I4t(alpha, gamma, q2, prefac, t1, 1.e16) + I4t(alpha, gamma, q2, prefac, t2, 1.e16));
    return sign * I4t(alpha, gamma, q2, prefac, t1, t2);
}

__device__ double I42t(double alpha, double gamma, double prefac, double t1, double t2)
{
    KEM_EXACT
    const double g1 = sqrt(alpha * t1 * t1 + gamma);
    const double g2 = sqrt(alpha * t2 * t2 + gamma);
    if (gamma - alpha <= 0) return prefac * (g2 - g1) / ((alpha - gamma) + g1 * g2);
    const double q = sqrt(gamma - alpha);
    return prefac * 1. / q * atanh(q * (g2 - g1) / ((alpha - gamma) + g1 * g2));
}

__device__ double I42(double a, double b, double u1, double u2)
{
    KEM_EXACT
    const double alpha = 1. + (a * a) / (b * b);
    const double gamma = (a * a + b * b) * (a * a + b * b + 1) / (b * b);
    const double lambda = -a / b;
    const double prefac = (a * a / b + b);
    const double t1 = (b * u1 != -a) ? (b - a * u1) / (a + b * u1) : 1.e15;
    const double t2 = (b * u2 != -a) ? (b - a * u2) / (a + b * u2) : 1.e15;
    if (sgn(u1 - lambda) != sgn(u2 - lambda)) {
        if (u1 > lambda) return (I42t(alpha, gamma, prefac, fabs(t1), 1.e15) + I42t(alpha, gamma, prefac, fabs(t2), 1.e15));
        return (I42t(alpha, gamma, prefac, 1.e15, fabs(t1)) + I42t(alpha, gamma, prefac, 1.e15, fabs(t2)));
    }
    if (u1 > lambda) return I42t(alpha, gamma, prefac, t1, t2);
    return I42t(alpha, gamma, prefac, t2, t1);
}

__device__ double I1(double a, double b, double u1, double u2)
{
    KEM_EXACT
    return F1(a, b, u2) - F1(a, b, u1) + I3(a, b, u1, u2) - I4(a, b, u1, u2);
}

__device__ double I6(double x, double u1, double u2)
{
    KEM_EXACT
    if (fabs(x) < 1.e-15) return 0;
    return x * log((sqrt(u2 * u2 + x * x + 1.) + u2) / (sqrt(u1 * u1 + x * x + 1.) + u1));
}

__device__ double I7(double x, double u1, double u2)
{
    KEM_EXACT
    const double t1 = (fabs(u1) > 1.e-16) ? 1. / u1 : 1.e16;
    const double t2 = (fabs(u2) > 1.e-16) ? 1. / u2 : 1.e16;
    const double g1 = sqrt(1. + t1 * t1 * (1. + x * x));
    const double g2 = sqrt(1. + t2 * t2 * (1. + x * x));
    return atan(x * (g2 - g1) / (x * x + g2 * g1));
}

__device__ double I2(double x, double u1, double u2)
{
    KEM_EXACT
    const double f = (F1(x, 0., u2) - F1(x, 0., u1));
    if (sgn(u1) != sgn(u2)) {
        if (u1 <= 0.) return f + I6(x, u1, u2) + I7(x, 0., fabs(u1)) + I7(x, 0., fabs(u2));
        return f + I6(x, u1, u2) + I7(x, fabs(u1), 0.) + I7(x, fabs(u2), 0.);
    }
    if (u1 <= 0.) return f + I6(x, u1, u2) + I7(x, u2, u1);
    return f + I6(x, u1, u2) + I7(x, u1, u2);
}

__device__ double J2(double a, double u1, double u2)
{
    KEM_EXACT
    if (a == 0.) return 0.;
    const double g1 = sqrt(u1 * u1 + a * a + 1.);
    const double g2 = sqrt(u2 * u2 + a * a + 1.);
    return a / (2. * fabs(a)) * log(((g2 - fabs(a)) * (g1 + fabs(a))) / ((g2 + fabs(a)) * (g1 - fabs(a))));
}

__device__ double ez_edge(double a, double b, double u0, double u1)
{
    KEM_EXACT
    if (fabs(b) > 1.e-14) return I4(a, b, u0, u1);
    if (sgn(u0) != sgn(u1)) {
        if (u0 <= 0.) return -(I7(a, 0., fabs(u0)) + I7(a, 0., fabs(u1)));
        return -(I7(a, fabs(u0), 0.) + I7(a, fabs(u1), 0.));
    }
    if (u0 <= 0.) return I7(a, u0, u1);
    return I7(a, u1, u0);
}

__device__ double ey_edge(double a, double b, double u0, double u1)
{
    KEM_EXACT
    if (fabs(b) > 1.e-14) return I42(a, b, u0, u1) + b * I3p(a, b, u0, u1);
    return J2(a, u0, u1);
}

__device__ double noz_edge(double a, double b, double y, double ylim)
{
    KEM_EXACT
    const double logArg =
        (1. + b * b) * y + a * b + sqrt(1. + b * b) * sqrt((1. + b * b) * y * y + 2 * a * b * y + a * a);
    if (logArg > 1.e-14) {
        if (fabs(y) > ylim) return y * asinh((a + b * y) / fabs(y)) + a / sqrt(1. + b * b) * log(logArg);
        return a / sqrt(1. + b * b) * log(logArg);
    }
    if (fabs(y) > ylim) return y * asinh(y * b / fabs(y));
    return 0.;
}

__device__ double noz(double a2, double b2, double a1, double b1, double y)
{
    KEM_EXACT
    return noz_edge(a2, b2, y, 1.e-14) - noz_edge(a1, b1, y, 5.e-14);
}

// KElectrostaticAnalyticTriangleIntegrator::Potential (:13-83) and ElectricField (:85-174)
__device__ F4 analytic_triangle(const double* d, double px, double py, double pz)
{
    KEM_EXACT
    const double* N1 = d + 5;
    const double* N2 = d + 8;
    double N3[3];
    normal(d, N3);
    const double p[3] = {d[2] - px, d[3] - py, d[4] - pz};
    const double n1dotn2 = dot(N1, N2);
    const double t[3] = {N2[0] - n1dotn2 * N1[0], N2[1] - n1dotn2 * N1[1], N2[2] - n1dotn2 * N1[2]};
    const double tm = mag(t);
    double N2p[3];
    for (int k = 0; k < 3; k++) N2p[k] = tm > 0. ? t[k] / tm : t[k];
    const double n2dotn2p = dot(N2, N2p);
    const double x0 = dot(p, N1);
    double y0 = dot(p, N2p);
    const double z0 = -dot(p, N3);
    const double x1 = x0 + d[0];
    const double x2 = x0 + d[1] * n1dotn2;
    double y1, zsign;
    if (z0 < 0) {
        y0 = -y0;
        y1 = y0 - d[1] * n2dotn2p;
        zsign = -1;
    } else {
        y1 = y0 + d[1] * n2dotn2p;
        zsign = 1;
    }
    const double z = fabs(z0);
    double u0, u1, a0, a1;
    if (z > 1.e-14) {
        u0 = y0 / z;
        if (u0 == 0.) u0 = 1.e-18;
        u1 = y1 / z;
        if (u1 == 0.) u1 = 1.e-18;
        a0 = (x0 * y1 - x2 * y0) / (z * (y1 - y0));
        a1 = (x1 * y1 - x2 * y0) / (z * (y1 - y0));
    } else {
        u0 = y0;
        if (u0 == 0.) u0 = 1.e-18;
        u1 = y1;
        if (u1 == 0.) u1 = 1.e-18;
        a0 = (x0 * y1 - x2 * y0) / (y1 - y0);
        a1 = (x1 * y1 - x2 * y0) / (y1 - y0);
    }
    const double b0 = (x2 - x0) / (y1 - y0);
    const double b1 = (x2 - x1) / (y1 - y0);

    double I;
    if (z > 1.e-14) {
        if (fabs(b0) < 1.e-13)
            I = z * (I1(a1, b1, u0, u1) - I2(a0, u0, u1));
        else if (fabs(b1) < 1.e-13)
            I = z * (I2(a1, u0, u1) - I1(a0, b0, u0, u1));
        else
            I = z * (I1(a1, b1, u0, u1) - I1(a0, b0, u0, u1));
    } else {
        I = (noz(a1, b1, a0, b0, u1) - noz(a1, b1, a0, b0, u0));
    }
    F4 r;
    r.p = fabs(I) / (4. * kPi * kEps0);

    double c[3];
    for (int k = 0; k < 3; k++) c[k] = d[2 + k] + (d[0] * d[5 + k] + d[1] * d[8 + k]) / 3.;
    const double dc[3] = {c[0] - px, c[1] - py, c[2] - pz};
    const double dist = mag(dc);
    const double prefac = 1. / (4. * kPi * kEps0);
    const double l0 = zsign * prefac * (I3p(a1, b1, u0, u1) - I3p(a0, b0, u0, u1));
    const double l1 = prefac * (ey_edge(a0, b0, u0, u1) - ey_edge(a1, b1, u0, u1));
    double l2;
    if (z > 1.e-14)
        l2 = prefac * (ez_edge(a1, b1, u0, u1) - ez_edge(a0, b0, u0, u1));
    else
        l2 = (dist < 1.e-12) ? 1. / (2. * kEps0) : 0.;
    r.x = N1[0] * l0 + N2p[0] * l1 + N3[0] * l2;
    r.y = N1[1] * l0 + N2p[1] * l1 + N3[1] * l2;
    r.z = N1[2] * l0 + N2p[2] * l1 + N3[2] * l2;
    return r;
}

// ------------------------------------------------------------ analytic rectangle
__device__ double rect_ln(double x, double y, double w)
{
    KEM_EXACT
    const double r = sqrt(fabs(x * x + y * y + w * w));
    const double r0 = sqrt(fabs(y * y + w * w));
    const double xa = fabs(x);
    const double c1 = (xa < 1.e-10) ? 0. : xa * log(fabs(y + r) + 1.e-12);
    const double c2 = (fabs(y) < 1.e-12) ? 0. : y * log(fabs((xa + r) / r0) + 1.e-12);
    const double c3 = (fabs(w) < 1.e-12) ? 0. : w * (atan(xa / w) + atan(y * w / (x * x + w * w + xa * r)) - atan(y / w));
    const double ret = c1 + c2 - xa + c3;
    return (x < 0.) ? -ret : ret;
}

__device__ double rect_exy(double x1, double x2, double y1, double y2, double z)
{
    KEM_EXACT
    double a1 = y2 + sqrt(y2 * y2 + x1 * x1 + z * z);
    double a2 = y2 + sqrt(y2 * y2 + x2 * x2 + z * z);
    double a3 = y1 + sqrt(y1 * y1 + x1 * x1 + z * z);
    double a4 = y1 + sqrt(y1 * y1 + x2 * x2 + z * z);
    if (fabs(z) < 1.e-14) {
        if (fabs(x1) < 1.e-14) {
            a1 = fabs(y1);
            a3 = fabs(y2);
        }
        if (fabs(x2) < 1.e-14) {
            a2 = fabs(y1);
            a4 = fabs(y2);
        }
    }
    if (fabs(a1 - a3) < 1.e-14) a1 = a3 = 1.;
    if (fabs(a2 - a4) < 1.e-14) a2 = a4 = 1.;
    return log((a2 * a3) / (a1 * a4));
}

__device__ double rect_ez(double x1, double x2, double y1, double y2, double z)
{
    KEM_EXACT
    const double t1 = (fabs(y1) > 1.e-15 ? z / y1 : 1.e15);
    const double t2 = (fabs(y2) > 1.e-15 ? z / y2 : 1.e15);
    double g1 = sqrt(((x2 * x2 + z * z) * t1 * t1 + z * z) / (x2 * x2));
    double g2 = sqrt(((x1 * x1 + z * z) * t1 * t1 + z * z) / (x1 * x1));
    double g3 = sqrt(((x2 * x2 + z * z) * t2 * t2 + z * z) / (x2 * x2));
    double g4 = sqrt(((x1 * x1 + z * z) * t2 * t2 + z * z) / (x1 * x1));
    if (x2 < 0.) g1 = -g1;
    if (x1 < 0.) g2 = -g2;
    if (x2 < 0.) g3 = -g3;
    if (x1 < 0.) g4 = -g4;
    return atan(g3) - atan(g4) - atan(g1) + atan(g2);
}

// KElectrostaticAnalyticRectangleIntegrator::Potential (:25-43) and ElectricField (:45-121)
__device__ F4 analytic_rectangle(const double* d, double px, double py, double pz)
{
    KEM_EXACT
    const double* N1 = d + 5;
    const double* N2 = d + 8;
    double N3[3];
    normal(d, N3);
    const double p[3] = {px - d[2], py - d[3], pz - d[4]};
    const double uP = dot(p, N1), vP = dot(p, N2), w = dot(p, N3);
    const double xmin = -uP, xmax = -uP + d[0], ymin = -vP, ymax = -vP + d[1];
    F4 r;
    r.p = (rect_ln(xmax, ymax, w) - rect_ln(xmin, ymax, w) - rect_ln(xmax, ymin, w) + rect_ln(xmin, ymin, w)) /
          (4. * kPi * kEps0);
    const double prefac = 1. / (4. * kPi * kEps0);
    const double l0 = prefac * rect_exy(xmin, xmax, ymin, ymax, w);
    const double l1 = prefac * rect_exy(ymin, ymax, xmin, xmax, w);
    const double tmin = w / ymin, tmax = w / ymax;
    const double sign_z = (w < 0) ? -1. : 1.;
    double l2;
    if (fabs(w) < 1.e-13)
        l2 = (xmin < 0. && xmax > 0. && ymin < 0. && ymax > 0.) ? 1.0 / (2. * kEps0) : 0.;
    else if (sgn(tmin) != sgn(tmax))
        l2 = prefac * sign_z * fabs(rect_ez(xmin, xmax, 0, fabs(ymin), w) + rect_ez(xmin, xmax, 0, fabs(ymax), w));
    else
        l2 = prefac * sign_z * fabs(rect_ez(xmin, xmax, ymax, ymin, w));
    r.x = N1[0] * l0 + N2[0] * l1 + N3[0] * l2;
    r.y = N1[1] * l0 + N2[1] * l1 + N3[1] * l2;
    r.z = N1[2] * l0 + N2[2] * l1 + N3[2] * l2;
    return r;
}

// ------------------------------------------------------------ reference policy
// Gauss-Legendre 32-point rule, positive nodes ascending (mirror nodes -x, same weights),
// applied as sum_j w_j f(B + A x_j) then sum_j w_j f(B - A x_j), times A (KGaussLegendreQuadrature.cc)
static __constant__ const double kGL32x[16] = {
    0.0483076656877383162348, 0.144471961582796493485, 0.239287362252137074545, 0.33186860228212764978,
    0.421351276130635345364,  0.506899908932229390024, 0.587715757240762329041, 0.663044266930215200975,
    0.732182118740289680387,  0.794483795967942406963, 0.849367613732569970134, 0.896321155766052123965,
    0.934906075937739689171,  0.964762255587506430774, 0.9856115115452683354,   0.997263861849481563545};
static __constant__ const double kGL32w[16] = {
    0.0965400885147278005668, 0.0956387200792748594191, 0.0938443990808045656392, 0.0911738786957638847129,
    0.0876520930044038111428, 0.0833119242269467552222, 0.0781938957870703064717, 0.0723457941088485062254,
    0.0658222227763618468377, 0.0586840934785355471453, 0.0509980592623761761962, 0.0428358980222266806569,
    0.0342738629130214331027, 0.0253920653092620594558, 0.0162743947309056706052, 0.00701861000947009660041};

// unit point charge at Q seen from P, times 1/(4 pi eps0): triangle form C = 1/R, C^3
// (BiQuadratureTriangleIntegrator.cc:51-70); rectangle form 1/R, R^3 (...RectangleIntegrator.cc:33-56)
template <bool RECT>
__device__ __forceinline__ void point_charge(double qx, double qy, double qz, double px, double py, double pz,
                                             double* s, double w)
{
    KEM_EXACT
    const double x = px - qx, y = py - qy, z = pz - qz;
    const double R = sqrt(x * x + y * y + z * z);
    if (RECT) {
        const double R3 = R * R * R;
        s[0] += w * (kK / R3 * x);
        s[1] += w * (kK / R3 * y);
        s[2] += w * (kK / R3 * z);
        s[3] += w * (kK / R);
    } else {
        const double C = 1. / R, C3 = C * C * C;
        s[0] += w * (kK * C3 * x);
        s[1] += w * (kK * C3 * y);
        s[2] += w * (kK * C3 * z);
        s[3] += w * (kK * C);
    }
}

// outer rule over the height y in [0, h] of vertex P1 above side P2-P0, inner rule along the
// chord from side P1-P2 parallel to P2->P0
__device__ F4 biquad_triangle(const double* d, double px, double py, double pz)
{
    KEM_EXACT
    double A[3], B[3], C[3], a[3], b[3], c[3], ua[3];
    for (int k = 0; k < 3; k++) {
        A[k] = d[2 + k];
        B[k] = d[2 + k] + d[5 + k] * d[0];
        C[k] = d[2 + k] + d[8 + k] * d[1];
    }
    for (int k = 0; k < 3; k++) {
        a[k] = A[k] - C[k];
        b[k] = B[k] - C[k];
        c[k] = B[k] - A[k];
    }
    const double la = mag(a), lb = mag(b);
    for (int k = 0; k < 3; k++) ua[k] = (1. / la) * a[k];
    const double cosg = dot(a, b) / (la * lb);
    double DB[3];
    for (int k = 0; k < 3; k++) DB[k] = (C[k] + (lb * cosg) * ua[k]) - B[k];
    const double h = mag(DB);
    double so[4] = {0., 0., 0., 0.};
    const double Ao = h / 2., Bo = h / 2.;
    for (int half = 0; half < 2; half++)
        for (int jo = 0; jo < 16; jo++) {
            const double y = half == 0 ? Bo + Ao * kGL32x[jo] : Bo - Ao * kGL32x[jo];
            double E[3], EF[3];
            for (int k = 0; k < 3; k++) {
                E[k] = B[k] - ((y / h) * b[k]);
                EF[k] = E[k] - (B[k] - ((y / h) * c[k]));
            }
            const double L = mag(EF);
            const double Ai = L / 2., Bi = L / 2.;
            double si[4] = {0., 0., 0., 0.};
            for (int hi = 0; hi < 2; hi++)
                for (int ji = 0; ji < 16; ji++) {
                    const double x = hi == 0 ? Bi + Ai * kGL32x[ji] : Bi - Ai * kGL32x[ji];
                    point_charge<false>(E[0] + x * ua[0], E[1] + x * ua[1], E[2] + x * ua[2], px, py, pz, si,
                                        kGL32w[ji]);
                }
            for (int k = 0; k < 4; k++) so[k] += kGL32w[jo] * (Ai * si[k]);
        }
    return F4{Ao * so[0], Ao * so[1], Ao * so[2], Ao * so[3]};
}

// outer rule over x in [0, A] along N1, inner over y in [0, B] along N2
__device__ F4 biquad_rectangle(const double* d, double px, double py, double pz)
{
    KEM_EXACT
    double so[4] = {0., 0., 0., 0.};
    const double Ao = d[0] / 2., Bo = d[0] / 2.;
    const double Ai = d[1] / 2., Bi = d[1] / 2.;
    for (int half = 0; half < 2; half++)
        for (int jo = 0; jo < 16; jo++) {
            const double x = half == 0 ? Bo + Ao * kGL32x[jo] : Bo - Ao * kGL32x[jo];
            double si[4] = {0., 0., 0., 0.};
            for (int hi = 0; hi < 2; hi++)
                for (int ji = 0; ji < 16; ji++) {
                    const double y = hi == 0 ? Bi + Ai * kGL32x[ji] : Bi - Ai * kGL32x[ji];
                    point_charge<true>(d[2] + x * d[5] + y * d[8], d[3] + x * d[6] + y * d[9],
                                       d[4] + x * d[7] + y * d[10], px, py, pz, si, kGL32w[ji]);
                }
            for (int k = 0; k < 4; k++) so[k] += kGL32w[jo] * (Ai * si[k]);
        }
    return F4{Ao * so[0], Ao * so[1], Ao * so[2], Ao * so[3]};
}

// the segment in 8 equal parts, the 32-node rule on each (ElectricFieldAndPotential :192-270)
__device__ F4 line_256(const double* d, double px, double py, double pz)
{
    KEM_EXACT
    const double P[3] = {px, py, pz};
    const double L = sqrt((d[3] - d[0]) * (d[3] - d[0]) + (d[4] - d[1]) * (d[4] - d[1]) + (d[5] - d[2]) * (d[5] - d[2]));
    const double pu = 1. / L;
    const double u[3] = {pu * (d[3] - d[0]), pu * (d[4] - d[1]), pu * (d[5] - d[2])};
    const double seg = L / 8, hs = 0.5 * seg;
    const double prefac = 0.25 * d[6] / kEps0;
    double sf[3] = {0., 0., 0.}, sp = 0.;
    for (int k = 0; k < 8; k++) {
        double ctr[3];
        for (int j = 0; j < 3; j++) ctr[j] = 0.5 * ((d[j] + (k * seg * u[j])) + (d[j] + ((k + 1) * seg * u[j])));
        for (int i = 0; i < 16; i++) {
            const double wf = kGL32w[i] * hs, nd = kGL32x[i] * hs;
            double pd[3], nv[3], pm = 0., nm = 0.;
            for (int j = 0; j < 3; j++) {
                pd[j] = P[j] - (ctr[j] + (u[j] * nd));
                pm += pd[j] * pd[j];
                nv[j] = P[j] - (ctr[j] - (u[j] * nd));
                nm += nv[j] * nv[j];
            }
            pm = 1. / sqrt(pm);
            nm = 1. / sqrt(nm);
            sp += wf * (pm + nm);
            pm = pm * pm * pm;
            nm = nm * nm * nm;
            for (int j = 0; j < 3; j++) sf[j] += wf * (pm * pd[j] + nm * nv[j]);
        }
    }
    return F4{prefac * sf[0], prefac * sf[1], prefac * sf[2], prefac * sp};
}

}  // namespace pol

// One pair under the "analytic" (2) or "reference" (3) policy from the raw shape record d;
// segments under "analytic" use the analytic wire integral of the element record g.
__device__ __noinline__ F4 element_policy(const double* __restrict__ g, const double* __restrict__ d, int policy,
                                          double px, double py, double pz)
{
    KEM_EXACT
    const int type = (int)g[KG_TYPE];
    if (policy == 2) {
        if (type == KEM_TRIANGLE) return pol::analytic_triangle(d, px, py, pz);
        if (type == KEM_RECTANGLE) return pol::analytic_rectangle(d, px, py, pz);
        return line_analytic(g, px, py, pz);
    }
    if (type == KEM_TRIANGLE) return pol::biquad_triangle(d, px, py, pz);
    if (type == KEM_RECTANGLE) return pol::biquad_rectangle(d, px, py, pz);
    return pol::line_256(d, px, py, pz);
}

// One pair of any shape under any policy, the axisymmetric shapes included (their analytic
// integrators serve every policy): the kernels for meshes with conic sections or rings.
template <bool FIELD>
__device__ __forceinline__ F4 element_any(const double* __restrict__ g, const double* __restrict__ nd,
                                          const double* __restrict__ d, int policy, double px, double py, double pz)
{
    const int type = (int)g[KG_TYPE];
    if (type >= KEM_CONICSECTION) return axisym_eval<FIELD>(g, type, px, py, pz);
    if (policy >= 2) return element_policy(g, d, policy, px, py, pz);
    return element_eval<FIELD>(g, nd, policy, px, py, pz);
}

}  // namespace dev
}  // namespace kem

#endif
