// kem_prep.cpp -- host expansion of KEMField shape records into device element records.
//
// Runs once per mesh upload.  Every quantity is computed with the expression (and order of
// operations) the reference integrators use for it, so that the per-pair device code
// only adds the point-dependent arithmetic.  Reference files (KEMField/Source/...):
//   Surfaces/src/KTriangle.cc:18-28, Surfaces/include/KTriangle.hh:27-34, KRectangle.hh:27-34
//   BoundaryIntegrals/Electrostatic/Numeric/{include,src}/KElectrostaticCubatureTriangleIntegrator.*
//   BoundaryIntegrals/Electrostatic/Numeric/{include,src}/KElectrostaticCubatureRectangleIntegrator.*
//   BoundaryIntegrals/Electrostatic/Numeric/{include,src}/KElectrostaticQuadratureLineSegmentIntegrator.*
//   BoundaryIntegrals/Electrostatic/RWG/src/KElectrostaticRWG{Triangle,Rectangle}Integrator.cc
#include "kem_prep.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <utility>
#include <vector>

namespace kem {

namespace {

inline double sq(double x) { return x * x; }

// Published cubature rules (normalised weights, sum 1):
//  triangle 7-point Radon / Stroud (degree 5), 12-point Gatermann 1988 (degree 7),
//  33-point Papanicolopulos 2011 (degree 12): KElectrostaticCubatureTriangleIntegrator.hh:56-258
//  rectangle Stroud 5-2 7-point, Tyler 12-point, 33-point (degree 13 on [-1,1]^2):
//  KElectrostaticCubatureRectangleIntegrator.hh:64-160, .cc:43-290
const double kT12a[4] = {0.6238226509439084e-1, 0.5522545665692000e-1, 0.3432430294509488e-1, 0.5158423343536001};
const double kT12b[4] = {0.6751786707392436e-1, 0.3215024938520156, 0.6609491961867980, 0.2777161669764050};
const double kT12c[4] = {0.8700998678316848, 0.6232720494910644, 0.3047265008681072, 0.2064414986699949};
const double kT12w[4] = {0.2651702815743450e-1, 0.4388140871444811e-1, 0.2877504278497528e-1,
                         0.6749318700980879e-1};
const double kT33a[8] = {4.570749859701478e-01, 1.197767026828138e-01, 2.359249810891690e-02, 7.814843446812914e-01,
                         9.507072731273288e-01, 1.162960196779266e-01, 2.303415635526714e-02, 2.138249025617059e-02};
const double kT33b[8] = {2.714625070149261e-01, 4.401116486585931e-01, 4.882037509455416e-01, 1.092578276593543e-01,
                         2.464636343633559e-02, 2.554542286385173e-01, 2.916556797383410e-01, 1.272797172335894e-01};
const double kT33w[8] = {6.254121319590276e-02, 4.991833492806094e-02, 2.426683808145203e-02, 2.848605206887754e-02,
                         7.931642509973638e-03, 4.322736365941421e-02, 2.178358503860756e-02, 1.508367757651144e-02};
const double kR33X[9] = {0.00000000000000000000, 0.77880971155441942252, 0.95729769978630736566,
                         0.13818345986246535375, 0.94132722587292523695, 0.47580862521827590507,
                         0.75580535657208143627, 0.69625007849174941396, 0.34271655604040678941};
const double kR33Y[9] = {0.00000000000000000000, 0.98348668243987226379, 0.85955600564163892859,
                         0.95892517028753485754, 0.39073621612946100068, 0.85007667369974857597,
                         0.64782163718701073204, 0.70741508996444936217e-1, 0.40930456169403884330};
const double kR33W[9] = {0.30038211543122536139 / 4.,   0.29991838864499131666e-1 / 4., 0.38174421317083669640e-1 / 4.,
                         0.60424923817749980681e-1 / 4., 0.77492738533105339358e-1 / 4., 0.11884466730059560108 / 4.,
                         0.12976355037000271129 / 4.,   0.21334158145718938943 / 4.,   0.25687074948196783651 / 4.};
// Gauss-Legendre half tables (positive nodes): ...QuadratureLineSegmentIntegrator.hh:59-83
const double kGL4x[2] = {0.339981043584856, 0.861136311594053};
const double kGL4w[2] = {0.652145154862546, 0.347854845137454};
const double kGL16x[8] = {0.09501250983763744, 0.28160355077925891, 0.45801677765722739, 0.61787624440264375,
                          0.75540440835500303, 0.86563120238783174, 0.94457502307323258, 0.98940093499164993};
const double kGL16w[8] = {0.189450610455068496, 0.182603415044923589, 0.169156519395002532, 0.149595988816576731,
                          0.124628971255533872, 0.095158511682492785, 0.062253523938647892, 0.027152459411754095};

inline void put_node(double* nodes, int slot, const double q[3], double w)
{
    double* p = nodes + 4 * slot;
    p[0] = q[0];
    p[1] = q[1];
    p[2] = q[2];
    p[3] = w;
}

// Q = a*A + b*B + c*C component-wise (GaussPoints_Tri*P)
inline void bary(const double A[3], const double B[3], const double C[3], double a, double b, double c, double Q[3])
{
    for (int k = 0; k < 3; k++) Q[k] = a * A[k] + b * B[k] + c * C[k];
}

}  // namespace
void triangle_rule_table(double* t);
namespace {

void triangle_nodes(const double* d, double* nodes)
{
    double A[3], B[3], C[3], Q[3], t[4 * KEM_NODE_SLOTS];
    for (int k = 0; k < 3; k++) {
        A[k] = d[2 + k];
        B[k] = d[2 + k] + (d[0] * d[5 + k]);
        C[k] = d[2 + k] + (d[1] * d[8 + k]);
    }
    triangle_rule_table(t);
    for (int slot = 0; slot < KEM_N33 + 33; slot++) {
        const double* r = t + 4 * slot;
        bary(A, B, C, r[0], r[1], r[2], Q);
        put_node(nodes, slot, Q, r[3]);
    }
}

void rectangle_rule(const double* d, int slot, int n, const double* x, const double* y, const double* w, double* nodes)
{
    const double a2 = 0.5 * d[0], b2 = 0.5 * d[1];
    double cen[3], Q[3];
    for (int k = 0; k < 3; k++) cen[k] = d[2 + k] + a2 * d[5 + k] + b2 * d[8 + k];
    for (int i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) Q[k] = cen[k] + (x[i] * a2) * d[5 + k] + (y[i] * b2) * d[8 + k];
        put_node(nodes, slot + i, Q, w[i]);
    }
}

void rectangle_nodes(const double* d, double* nodes)
{
    {
        const double t2 = std::sqrt((7. - std::sqrt(24.)) / 15.), s2 = std::sqrt((7. + std::sqrt(24.)) / 15.),
                     r2 = std::sqrt(7. / 15.);
        const double x[7] = {0., r2, -r2, s2, -s2, t2, -t2};
        const double y[7] = {0., r2, -r2, -t2, t2, -s2, s2};
        const double w[7] = {2. / 7., 25. / 168., 25. / 168., 5. / 48., 5. / 48., 5. / 48., 5. / 48.};
        rectangle_rule(d, KEM_N7, 7, x, y, w, nodes);
    }
    {
        const double r = std::sqrt(6. / 7.), s = std::sqrt((114. - 3. * std::sqrt(583.)) / 287.),
                     t = std::sqrt((114. + 3. * std::sqrt(583.)) / 287.);
        const double x[12] = {r, -r, 0., 0., s, s, -s, -s, t, t, -t, -t};
        const double y[12] = {0., 0., r, -r, s, -s, s, -s, t, -t, t, -t};
        const double B1 = 49 / 810., B2 = (178981. + 2769. * std::sqrt(583)) / 1888920.,
                     B3 = (178981. - 2769. * std::sqrt(583)) / 1888920.;
        const double w[12] = {B1, B1, B1, B1, B2, B2, B2, B2, B3, B3, B3, B3};
        rectangle_rule(d, KEM_N12, 12, x, y, w, nodes);
    }
    {
        double x[33], y[33], w[33];
        x[0] = kR33X[0];
        y[0] = kR33Y[0];
        w[0] = kR33W[0];
        int m = 1;
        for (int o = 1; o < 9; o++) {
            const double xo[4] = {kR33X[o], -kR33X[o], -kR33Y[o], kR33Y[o]};
            const double yo[4] = {kR33Y[o], -kR33Y[o], kR33X[o], -kR33X[o]};
            for (int q = 0; q < 4; q++, m++) {
                x[m] = xo[q];
                y[m] = yo[q];
                w[m] = kR33W[o];
            }
        }
        rectangle_rule(d, KEM_N33, 33, x, y, w, nodes);
    }
}

void line_nodes(const double* d, double* nodes)
{
    const double c[3] = {0.5 * (d[0] + d[3]), 0.5 * (d[1] + d[4]), 0.5 * (d[2] + d[5])};
    const double L = std::sqrt(sq(d[3] - d[0]) + sq(d[4] - d[1]) + sq(d[5] - d[2]));
    const double hl = 0.5 * L, iu = 1. / L;
    const double u[3] = {iu * (d[3] - d[0]), iu * (d[4] - d[1]), iu * (d[5] - d[2])};
    auto rule = [&](int slot, int half, const double* xs, const double* ws) {
        for (int i = 0; i < half; i++) {
            double x = xs[i] * hl, wf = ws[i] * hl, qp[3], qm[3];
            for (int k = 0; k < 3; k++) {
                qp[k] = c[k] + (u[k] * x);
                qm[k] = c[k] - (u[k] * x);
            }
            put_node(nodes, slot + 2 * i, qp, wf);
            put_node(nodes, slot + 2 * i + 1, qm, wf);
        }
    };
    rule(KEM_L4, 2, kGL4x, kGL4w);
    rule(KEM_L16, 8, kGL16x, kGL16w);
}

// KConicSection::Normal (Surfaces/src/KConicSection.cc:59-104): in the x-z plane, -z / +z for a
// flat section running out / in, +x / -x for a cylinder running up / down, else the unit
// normal of the generating line; KRing::Normal (KRing.cc:40-43) is +x
void axisym_normal(int type, const double* d, double n[3])
{
    n[1] = 0.;
    if (type == KEM_RING) {
        n[0] = 1.;
        n[2] = 0.;
        return;
    }
    const double r0 = d[0], z0 = d[2], r1 = d[3], z1 = d[5];
    if (std::fabs(z0 - z1) < 1.e-10) {
        n[0] = 0.;
        n[2] = r0 < r1 ? -1. : 1.;
    } else if (std::fabs(r0 - r1) < 1.e-10) {
        n[2] = 0.;
        n[0] = z0 < z1 ? 1. : -1.;
    } else {
        const double m = (z0 - z1) / (r1 - r0);
        n[0] = m / std::sqrt(1. + m * m);
        n[2] = 1. / std::sqrt(1. + m * m);
    }
}

bool finite_all(const double* d, int n)
{
    for (int k = 0; k < n; k++)
        if (!std::isfinite(d[k])) return false;
    return true;
}

// squared-distance thresholds of the rule ladders (see kem_layout.h); the distance-ratio
// cut-offs are the reference's (KElectrostaticCubature*Integrator.hh, ...LineSegmentIntegrator.hh)
// and the FP32 ladder's (kem_device.hpp)
void set_thresholds(int type, double* g)
{
    const double size = 1. / g[KG_INVSIZE];
    auto t2 = [&](double dr) { return (dr * size) * (dr * size); };
    if (type == KEM_TRIANGLE) {
        g[KG_NEAR2] = t2(3.42);
        g[KG_REF7_2] = t2(131.4);
        g[KG_REF12_2] = t2(30.4);
        g[KG_MIX7_2] = t2(8.0);
        g[KG_MIX12_2] = t2(4.0);
        g[KG_MIX6_2] = t2(12.0);
    } else if (type == KEM_RECTANGLE) {
        g[KG_NEAR2] = t2(4.6);
        g[KG_REF7_2] = t2(196.2);
        g[KG_REF12_2] = t2(44.6);
        g[KG_MIX7_2] = t2(12.0);
        g[KG_MIX12_2] = t2(6.0);
        g[KG_MIX6_2] = INFINITY;
    } else {
        g[KG_NEAR2] = t2(2.0);
        g[KG_REF7_2] = t2(27.7);
        g[KG_REF12_2] = t2(27.7);
        g[KG_MIX7_2] = t2(27.7);
        g[KG_MIX12_2] = t2(27.7);
        g[KG_MIX6_2] = INFINITY;
    }
}

}  // namespace

void triangle_rule_table(double* t)
{
    auto put = [&](int slot, double a, double b, double c, double w) {
        t[4 * slot + 0] = a;
        t[4 * slot + 1] = b;
        t[4 * slot + 2] = c;
        t[4 * slot + 3] = w;
    };
    // 7-point: centroid, then two orbits (alpha,beta,beta) in the order (a,b,c),(b,a,c),(c,b,a)
    const double th = 1. / 3., q15 = std::sqrt(15), r = (1. + q15) / 7., s = (1. - q15) / 7.;
    const double a7[3] = {th, th + (2. * th * s), th + (2. * th * r)};
    const double b7[3] = {th, th - (th * s), th - (th * r)};
    const double w7[3] = {9 / 40., (155. + q15) / 1200., (155. - q15) / 1200.};
    int slot = KEM_N7;
    put(slot++, a7[0], b7[0], b7[0], w7[0]);
    for (int o = 1; o < 3; o++) {
        put(slot++, a7[o], b7[o], b7[o], w7[o]);
        put(slot++, b7[o], a7[o], b7[o], w7[o]);
        put(slot++, b7[o], b7[o], a7[o], w7[o]);
    }
    // 12-point: cyclic orbits (a,b,c),(b,c,a),(c,a,b), weights doubled
    slot = KEM_N12;
    for (int o = 0; o < 4; o++) {
        const double a = kT12a[o], b = kT12b[o], c = kT12c[o];
        put(slot++, a, b, c, kT12w[o] * 2.);
        put(slot++, b, c, a, kT12w[o] * 2.);
        put(slot++, c, a, b, kT12w[o] * 2.);
    }
    // 33-point: five 3-orbits and three 6-orbits
    slot = KEM_N33;
    for (int o = 0; o < 8; o++) {
        const double a = kT33a[o], b = kT33b[o], c = 1. - kT33a[o] - kT33b[o];
        const double perm[6][3] = {{a, b, c}, {b, a, c}, {c, b, a}, {a, c, b}, {c, a, b}, {b, c, a}};
        const int np = o < 5 ? 3 : 6;
        for (int p = 0; p < np; p++) put(slot++, perm[p][0], perm[p][1], perm[p][2], kT33w[o]);
    }
}

int prep_element(int type, const double* d, double* g, double* nodes)
{
    std::memset(g, 0, sizeof(double) * KG_STRIDE);
    std::memset(nodes, 0, sizeof(double) * KEM_NODE_STRIDE);
    g[KG_TYPE] = type;
    if (type == KEM_LINESEGMENT) {
        if (!finite_all(d, 7)) return -1;
        const double L = std::sqrt(((d[3] - d[0]) * (d[3] - d[0])) + ((d[4] - d[1]) * (d[4] - d[1])) +
                                   ((d[5] - d[2]) * (d[5] - d[2])));
        if (!(L > 0.) || !(d[6] > 0.)) return -1;
        for (int k = 0; k < 3; k++) {
            g[KG_CX + k] = (d[k] + d[3 + k]) * 0.5;
            g[KG_V + k] = d[k];
            g[KG_V + 3 + k] = d[3 + k];
            g[KG_U + k] = (d[3 + k] - d[k]) / L;
        }
        g[KG_INVSIZE] = 1. / L;
        g[KG_PREF] = 0.25 * d[6] / kEps0;
        g[KG_A] = L;
        g[KG_B] = d[6];
        set_thresholds(type, g);
        line_nodes(d, nodes);
        return 0;
    }
    if (type == KEM_CONICSECTION || type == KEM_RING) {
        // axisymmetric shapes: evaluated by the analytic ring / conic-section integrators at every
        // distance (the near pass), so every threshold is infinite; radius = x, axial position = z
        const int nd = type == KEM_RING ? 3 : 6;
        if (!finite_all(d, nd)) return -1;
        double c[3], n[3];
        for (int k = 0; k < nd; k++) g[KG_V + k] = d[k];
        if (type == KEM_RING) {
            if (!(d[0] > 0.)) return -1;  // the ring prefactor divides by the radius
            for (int k = 0; k < 3; k++) c[k] = d[k];
            g[KG_INVSIZE] = 1. / d[0];
        } else {
            const double L = std::sqrt(sq(d[0] - d[3]) + sq(d[1] - d[4]) + sq(d[2] - d[5]));
            if (!(L > 0.)) return -1;
            for (int k = 0; k < 3; k++) c[k] = (d[k] + d[3 + k]) * .5;
            g[KG_A] = L;
            g[KG_INVSIZE] = 1. / L;
        }
        axisym_normal(type, d, n);
        for (int k = 0; k < 3; k++) {
            g[KG_CX + k] = c[k];
            g[KG_HX + k] = c[k];
            g[KG_NX + k] = n[k];
        }
        g[KG_PREF] = 1.;
        for (int k = KG_NEAR2; k <= KG_MIX6_2; k++) g[k] = INFINITY;
        return 0;
    }
    if (type != KEM_TRIANGLE && type != KEM_RECTANGLE) return -1;
    if (!finite_all(d, 11) || !(d[0] > 0.) || !(d[1] > 0.)) return -1;
    // unit normal n = unit(N1 x N2), normalised by multiplying with the inverse magnitude
    double n[3] = {d[6] * d[10] - d[7] * d[9], d[7] * d[8] - d[5] * d[10], d[5] * d[9] - d[6] * d[8]};
    const double nn = std::sqrt(sq(n[0]) + sq(n[1]) + sq(n[2]));
    if (!(nn > 1e-12)) return -1;
    const double inn = 1. / nn;
    for (int k = 0; k < 3; k++) g[KG_NX + k] = n[k] * inn;
    g[KG_A] = d[0];
    g[KG_B] = d[1];
    if (type == KEM_TRIANGLE) {
        double V[3][3];
        for (int k = 0; k < 3; k++) {
            V[0][k] = d[2 + k];
            V[1][k] = d[2 + k] + (d[0] * d[5 + k]);
            V[2][k] = d[2 + k] + (d[1] * d[8 + k]);
        }
        const double l12 = std::sqrt(sq(V[2][0] - V[1][0]) + sq(V[2][1] - V[1][1]) + sq(V[2][2] - V[1][2]));
        if (!(l12 > 0.)) return -1;
        for (int k = 0; k < 3; k++) {
            // centroid (KTriangle::Centroid; also the dispatch centre and the centre of h)
            g[KG_CX + k] = d[2 + k] + (d[0] * d[5 + k] + d[1] * d[8 + k]) / 3.;
            g[KG_HX + k] = g[KG_CX + k];
        }
        g[KG_INVSIZE] = 1. / ((d[0] + d[1] + l12) / 3.);
        g[KG_PREF] = 0.5 * d[0] * d[1] * nn * kOneOverFourPiEps0;
        g[KG_NE] = 3;
        const double e12[3] = {V[2][0] - V[1][0], V[2][1] - V[1][1], V[2][2] - V[1][2]};
        const double i12 = 1. / l12;
        const double hl[3] = {0.5 * d[0], 0.5 * l12, 0.5 * d[1]};
        for (int k = 0; k < 3; k++) {
            for (int v = 0; v < 3; v++) g[KG_V + 3 * v + k] = V[v][k];
            g[KG_U + 0 + k] = d[5 + k];
            g[KG_U + 3 + k] = i12 * e12[k];
            g[KG_U + 6 + k] = -d[8 + k];
            g[KG_SV + 0 + k] = d[0] * d[5 + k];
            g[KG_SV + 3 + k] = e12[k];
            g[KG_SV + 6 + k] = (-1) * d[1] * d[8 + k];
        }
        g[KG_IL + 0] = 1. / d[0];
        g[KG_IL + 1] = 1. / l12;
        g[KG_IL + 2] = 1. / d[1];
        for (int e = 0; e < 3; e++)
            for (int k = 0; k < 3; k++) g[KG_MID + 3 * e + k] = hl[e] * g[KG_U + 3 * e + k] + V[e][k];
        triangle_nodes(d, nodes);
    } else {
        double V[4][3];
        for (int k = 0; k < 3; k++) {
            V[0][k] = d[2 + k];
            V[1][k] = d[2 + k] + (d[0] * d[5 + k]);
            V[2][k] = d[2 + k] + (d[0] * d[5 + k]) + (d[1] * d[8 + k]);
            V[3][k] = d[2 + k] + (d[1] * d[8 + k]);
            // cubature / dispatch centre (...RectangleIntegrator.cc:470-473)
            g[KG_CX + k] = d[2 + k] + 0.5 * (d[0] * d[5 + k] + d[1] * d[8 + k]);
            // closed-form centre (RWG rectangle .cc:1212-1214, KSolidAngle.hh:188-190)
            g[KG_HX + k] = d[2 + k] + 0.5 * (d[0] * d[5 + k]) + 0.5 * (d[1] * d[8 + k]);
        }
        g[KG_INVSIZE] = 1. / (0.5 * (d[0] + d[1]));
        g[KG_PREF] = d[0] * d[1] * kOneOverFourPiEps0;
        g[KG_NE] = 4;
        for (int k = 0; k < 3; k++) {
            for (int v = 0; v < 4; v++) g[KG_V + 3 * v + k] = V[v][k];
            g[KG_U + 0 + k] = d[5 + k];
            g[KG_U + 3 + k] = d[8 + k];
            g[KG_U + 6 + k] = -d[5 + k];
            g[KG_U + 9 + k] = -d[8 + k];
            g[KG_MID + 0 + k] = (0.5 * d[0] * d[5 + k]) + V[0][k];
            g[KG_MID + 3 + k] = (0.5 * d[1] * d[8 + k]) + V[1][k];
            g[KG_MID + 6 + k] = (-0.5 * d[0] * d[5 + k]) + V[2][k];
            g[KG_MID + 9 + k] = (-0.5 * d[1] * d[8 + k]) + V[3][k];
            g[KG_SV + 0 + k] = d[0] * d[5 + k];
            g[KG_SV + 3 + k] = d[1] * d[8 + k];
            g[KG_SV + 6 + k] = -d[0] * d[5 + k];
            g[KG_SV + 9 + k] = -d[1] * d[8 + k];
        }
        g[KG_IL + 0] = 1. / d[0];
        g[KG_IL + 1] = 1. / d[1];
        g[KG_IL + 2] = 1. / d[0];
        g[KG_IL + 3] = 1. / d[1];
        rectangle_nodes(d, nodes);
    }
    set_thresholds(type, g);
    // in-plane outward edge normals m_e = u_e x n
    const int ne = (int)g[KG_NE];
    for (int e = 0; e < ne; e++) {
        const double* u = g + KG_U + 3 * e;
        const double* nv = g + KG_NX;
        double* m = g + KG_M + 3 * e;
        m[0] = (u[1] * nv[2]) - (u[2] * nv[1]);
        m[1] = (u[2] * nv[0]) - (u[0] * nv[2]);
        m[2] = (u[0] * nv[1]) - (u[1] * nv[0]);
    }
    return 0;
}

void prep_nodes32(const double* g, const double* nodes, float* nodes32)
{
    for (int s = 0; s < KEM_NODE_SLOTS; s++) {
        const double* q = nodes + 4 * s;
        float* o = nodes32 + 4 * s;
        if (q[3] == 0.) {
            o[0] = o[1] = o[2] = o[3] = 0.f;
            continue;
        }
        o[0] = (float)(q[0] - g[KG_CX]);
        o[1] = (float)(q[1] - g[KG_CY]);
        o[2] = (float)(q[2] - g[KG_CZ]);
        o[3] = (float)q[3];
    }
    // degree-4 6-point rule (Dunavant 1985, positive weights) for the far end of the mixed
    // ladder: orbits (1-2a, a, a), a = 0.445948490915965 (w 0.223381589678011) and
    // a = 0.091576213509771 (w 0.109951743655322)
    for (int s = KEM_N6; s < KEM_NODE32_SLOTS; s++)
        for (int k = 0; k < 4; k++) nodes32[4 * s + k] = 0.f;
    if ((int)g[KG_TYPE] != KEM_TRIANGLE) return;
    const double a[2] = {0.445948490915965, 0.091576213509771};
    const double w[2] = {0.223381589678011, 0.109951743655322};
    int s = KEM_N6;
    for (int o = 0; o < 2; o++) {
        const double b[3][3] = {{1. - 2. * a[o], a[o], a[o]}, {a[o], 1. - 2. * a[o], a[o]}, {a[o], a[o], 1. - 2. * a[o]}};
        for (int p = 0; p < 3; p++, s++) {
            float* out = nodes32 + 4 * s;
            for (int k = 0; k < 3; k++) {
                const double q = b[p][0] * g[KG_V + k] + b[p][1] * g[KG_V + 3 + k] + b[p][2] * g[KG_V + 6 + k];
                out[k] = (float)(q - g[KG_CX + k]);
            }
            out[3] = (float)w[o];
        }
    }
}

void prep_target(int type, const double* d, int bc, double flux, double* t)
{
    if (type == KEM_CONICSECTION || type == KEM_RING) {
        double n[3];
        axisym_normal(type, d, n);
        for (int k = 0; k < 3; k++) {
            // KConicSection::Centroid (P0 + P1)/2, KRing::Centroid P (KConicSection.hh, KRing.hh)
            t[KT_CX + k] = type == KEM_RING ? d[k] : (d[k] + d[3 + k]) * .5;
            t[KT_NX + k] = n[k];
        }
        t[KT_BC] = bc;
        t[KT_FLUX] = flux;
        return;
    }
    for (int k = 0; k < 3; k++) {
        if (type == KEM_TRIANGLE)
            t[KT_CX + k] = d[2 + k] + (d[0] * d[5 + k] + d[1] * d[8 + k]) / 3.;  // KTriangle.hh:31-34
        else if (type == KEM_RECTANGLE)
            t[KT_CX + k] = d[2 + k] + d[0] * d[5 + k] * .5 + d[1] * d[8 + k] * .5;  // KRectangle.hh:31-34
        else
            t[KT_CX + k] = (d[k] + d[3 + k]) * .5;  // KLineSegment.hh:29-32
    }
    if (type == KEM_LINESEGMENT) {
        t[KT_NX] = t[KT_NY] = t[KT_NZ] = 0.;
    } else {
        // KTriangle::fN3 = unit(N1 x N2) (KThreeVector::Unit divides by the magnitude)
        double n[3] = {d[6] * d[10] - d[7] * d[9], d[7] * d[8] - d[5] * d[10], d[5] * d[9] - d[6] * d[8]};
        const double m = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        for (int k = 0; k < 3; k++) t[KT_NX + k] = n[k] / m;
    }
    t[KT_BC] = bc;
    t[KT_FLUX] = flux;
}

namespace {
uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffULL;
    v = (v | v << 16) & 0x1f0000ff0000ffULL;
    v = (v | v << 8) & 0x100f00f00f00f00fULL;
    v = (v | v << 4) & 0x10c30c30c30c30c3ULL;
    v = (v | v << 2) & 0x1249249249249249ULL;
    return v;
}
}  // namespace

void build_clusters(int64_t n, const double* g, int32_t* order, double* clusters)
{
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int64_t j = 0; j < n; j++)
        for (int k = 0; k < 3; k++) {
            lo[k] = std::min(lo[k], g[j * KG_STRIDE + KG_CX + k]);
            hi[k] = std::max(hi[k], g[j * KG_STRIDE + KG_CX + k]);
        }
    double ext = 0.;
    for (int k = 0; k < 3; k++) ext = std::max(ext, hi[k] - lo[k]);
    const double scale = ext > 0. ? 2097151. / ext : 0.;
    std::vector<std::pair<uint64_t, int32_t>> key((size_t)n);
    for (int64_t j = 0; j < n; j++) {
        uint64_t m = 0;
        for (int k = 0; k < 3; k++) {
            double q = (g[j * KG_STRIDE + KG_CX + k] - lo[k]) * scale;
            q = std::min(std::max(q, 0.), 2097151.);
            m |= spread21((uint64_t)q) << k;
        }
        key[(size_t)j] = {m, (int32_t)j};
    }
    std::sort(key.begin(), key.end());
    for (int64_t j = 0; j < n; j++) order[j] = key[(size_t)j].second;
    const int64_t ncl = (n + KEM_CLUSTER - 1) / KEM_CLUSTER;
    for (int64_t c = 0; c < ncl; c++) {
        const int64_t e0 = c * KEM_CLUSTER, e1 = std::min(n, e0 + KEM_CLUSTER);
        double blo[3] = {INFINITY, INFINITY, INFINITY}, bhi[3] = {-INFINITY, -INFINITY, -INFINITY};
        for (int64_t e = e0; e < e1; e++)
            for (int k = 0; k < 3; k++) {
                blo[k] = std::min(blo[k], g[order[e] * (int64_t)KG_STRIDE + KG_CX + k]);
                bhi[k] = std::max(bhi[k], g[order[e] * (int64_t)KG_STRIDE + KG_CX + k]);
            }
        double ctr[3];
        for (int k = 0; k < 3; k++) ctr[k] = 0.5 * (blo[k] + bhi[k]);
        double r = 0.;
        for (int64_t e = e0; e < e1; e++) {
            const double* ge = g + order[e] * (int64_t)KG_STRIDE;
            const double d = std::sqrt(sq(ge[KG_CX] - ctr[0]) + sq(ge[KG_CY] - ctr[1]) + sq(ge[KG_CZ] - ctr[2]));
            r = std::max(r, d + std::sqrt(ge[KG_NEAR2]));
        }
        clusters[4 * c + 0] = ctr[0];
        clusters[4 * c + 1] = ctr[1];
        clusters[4 * c + 2] = ctr[2];
        clusters[4 * c + 3] = r * (1. + 1e-12);
    }
}

void build_far6(int64_t n, const double* g, const float* nodes32, const int32_t* order, const double* clusters,
                float* far6, float* far3, double* far_cl)
{
    const int64_t ncl = (n + KEM_CLUSTER - 1) / KEM_CLUSTER;
    for (int64_t c = 0; c < ncl; c++) {
        const int64_t e0 = c * KEM_CLUSTER, e1 = std::min(n, e0 + KEM_CLUSTER);
        double reach = 0., reach3 = 0.;
        for (int64_t e = e0; e < e1; e++) {
            const int64_t j = order[e];
            const double* ge = g + j * (int64_t)KG_STRIDE;
            float* r = far6 + e * KEM_FAR6_STRIDE;
            for (int k = 0; k < KEM_FAR6_STRIDE; k++) r[k] = 0.f;
            // the 6 nodes as offsets from the cluster centre: (element centre - cluster centre)
            // in FP64 plus the element's FP32 node offset, rounded once
            for (int q = 0; q < 6; q++) {
                const float* nq = nodes32 + j * KEM_NODE32_STRIDE + 4 * (KEM_N6 + q);
                for (int k = 0; k < 3; k++)
                    r[4 * q + k] = (float)((ge[KG_CX + k] - clusters[4 * c + k]) + (double)nq[k]);
                r[4 * q + 3] = nq[3];
            }
            // the element centre (for the per-element 6-point test) and its 6-point reach
            double ce[4] = {ge[KG_CX], ge[KG_CY], ge[KG_CZ], INFINITY};
            if ((int)ge[KG_TYPE] == KEM_TRIANGLE) ce[3] = std::sqrt(ge[KG_MIX6_2]) * (1. + 1e-9);
            std::memcpy(r + KEM_FAR6_CENTRE, ce, sizeof ce);
            const double d = std::sqrt(sq(ce[0] - clusters[4 * c]) + sq(ce[1] - clusters[4 * c + 1]) +
                                       sq(ce[2] - clusters[4 * c + 2]));
            reach = std::max(reach, d + ce[3]);
            // the 3-point rule: nodes (2 V_q + (V_q+1 + V_q+2) / 2) / 3 relative to the cluster centre
            float* r3 = far3 + e * KEM_FAR3_STRIDE;
            for (int k = 0; k < KEM_FAR3_STRIDE; k++) r3[k] = 0.f;
            if ((int)ge[KG_TYPE] == KEM_TRIANGLE) {
                const double* v = ge + KG_V;
                for (int q = 0; q < 3; q++)
                    for (int k = 0; k < 3; k++) {
                        const double x = (2. / 3.) * v[3 * q + k] + (1. / 6.) * v[3 * ((q + 1) % 3) + k] +
                                         (1. / 6.) * v[3 * ((q + 2) % 3) + k];
                        r3[4 * q + k] = (float)(x - clusters[4 * c + k]);
                    }
                reach3 = std::max(reach3, d + KEM_FAR3_RATIO / ge[KG_INVSIZE] * (1. + 1e-9));
            } else {
                reach3 = INFINITY;
            }
        }
        double* cl = far_cl + KEM_FAR_CL_STRIDE * c;
        for (int k = 0; k < KEM_FAR_CL_STRIDE; k++) cl[k] = 0.;
        for (int k = 0; k < 3; k++) cl[k] = clusters[4 * c + k];
        cl[3] = reach * (1. + 1e-12);
        cl[4] = reach3 * (1. + 1e-12);
    }
}

}  // namespace kem
