/*
 * kem_oracle.c -- CPU FP64 restatement of KEMField's electrostatic BEM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product): see kem_oracle.h.
 *
 * Written for this project from the algorithms of the cited KEMField files.  The
 * reference's branch structure, thresholds and quadrature tables are kept because the
 * last bits of the results depend on them; loops and data flow are this file's own.
 * Paths are relative to KEMField/Source/ unless they start with Kommon/.
 */
#include "kem_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define SQ(x) ((x) * (x))

/* ---------------------------------------------------------------- constants */
/* Kommon/Base/Utility/KConst_2021.h:37 (pi), :139-146 (eps0, 4 pi eps0);
   Core/src/KEMConstants.cc:9-12 */
static const double KO_PI = 3.141592653589793238462643383279502884;
static const double KO_EPS0 = 8.8541878128E-12;
#define KO_K (1. / (4. * KO_PI * KO_EPS0))

double ko_pi(void) { return KO_PI; }
double ko_eps0(void) { return KO_EPS0; }
double ko_one_over_4pi_eps0(void) { return KO_K; }

static void cross3(const double a[3], const double b[3], double c[3])
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double a[3], const double b[3]) { return (a[0] * b[0] + a[1] * b[1]) + a[2] * b[2]; }
static double norm3(const double a[3]) { return sqrt(SQ(a[0]) + SQ(a[1]) + SQ(a[2])); }

/* ---------------------------------------------------------------- geometry */
/* KTriangle::SetValues(p0,p1,p2): Surfaces/src/KTriangle.cc:18-28 */
void ko_set_from_points(const double p0[3], const double p1[3], const double p2[3], double rec[KO_REC])
{
    double n1[3], n2[3];
    for (int k = 0; k < 3; k++) {
        n1[k] = p1[k] - p0[k];
        n2[k] = p2[k] - p0[k];
    }
    double a = norm3(n1), b = norm3(n2);
    rec[0] = a;
    rec[1] = b;
    for (int k = 0; k < 3; k++) {
        rec[2 + k] = p0[k];
        /* KThreeVector::Unit() divides by the magnitude */
        rec[5 + k] = n1[k] / a;
        rec[8 + k] = n2[k] / b;
    }
}

/* unit normal N3 = unit(N1 x N2): KTriangle.cc:16,28 */
double ko_axisym_area(int shape, const double* d);
void ko_axisym_normal(int shape, const double* d, double n[3]);

void ko_normal(int shape, const double* d, double n[3])
{
    if (shape == KO_CONICSECTION || shape == KO_RING) {
        ko_axisym_normal(shape, d, n);
        return;
    }
    double c[3];
    cross3(d + 5, d + 8, c);
    double m = norm3(c);
    for (int k = 0; k < 3; k++) n[k] = c[k] / m;
}

/* KTriangle.hh:27-30, KRectangle.hh:27-30, KLineSegment.hh:25-28 */
double ko_area(int shape, const double* d)
{
    if (shape == KO_CONICSECTION || shape == KO_RING) return ko_axisym_area(shape, d);
    if (shape == KO_TRIANGLE) {
        double c[3];
        cross3(d + 5, d + 8, c);
        return .5 * d[0] * d[1] * norm3(c);
    }
    if (shape == KO_RECTANGLE) return d[0] * d[1];
    double l[3] = {d[0] - d[3], d[1] - d[4], d[2] - d[5]};
    return KO_PI * d[6] * norm3(l);
}

/* KTriangle.hh:31-34, KRectangle.hh:31-34, KLineSegment.hh:29-32 */
void ko_centroid(int shape, const double* d, double c[3])
{
    if (shape == KO_RING) { /* KRing.hh Centroid: P */
        for (int k = 0; k < 3; k++) c[k] = d[k];
        return;
    }
    for (int k = 0; k < 3; k++) { /* conic section: (P0 + P1)/2, KConicSection.hh */
        if (shape == KO_TRIANGLE)
            c[k] = d[2 + k] + (d[0] * d[5 + k] + d[1] * d[8 + k]) / 3.;
        else if (shape == KO_RECTANGLE)
            c[k] = d[2 + k] + d[0] * d[5 + k] * .5 + d[1] * d[8 + k] * .5;
        else
            c[k] = (d[k] + d[3 + k]) * .5;
    }
}

/* ------------------------------------------------- triangle cubature tables */
/* Numeric/include/KElectrostaticCubatureTriangleIntegrator.hh:56-258.
   Published rules: 7-point Radon/Stroud (degree 5), 12-point Gatermann 1988 (degree 7),
   33-point Papanicolopulos 2011 (degree 12).  Barycentric orbits and normalised weights
   (sum 1). */
static double T7a[3], T7b[3], T7w[3];
static const double T12a[4] = {0.6238226509439084e-1, 0.5522545665692000e-1, 0.3432430294509488e-1,
                               0.5158423343536001};
static const double T12b[4] = {0.6751786707392436e-1, 0.3215024938520156, 0.6609491961867980,
                               0.2777161669764050};
static const double T12c[4] = {0.8700998678316848, 0.6232720494910644, 0.3047265008681072,
                               0.2064414986699949};
static const double T12w[4] = {0.2651702815743450e-1, 0.4388140871444811e-1, 0.2877504278497528e-1,
                               0.6749318700980879e-1};
static const double T33a[8] = {4.570749859701478e-01, 1.197767026828138e-01, 2.359249810891690e-02,
                               7.814843446812914e-01, 9.507072731273288e-01, 1.162960196779266e-01,
                               2.303415635526714e-02, 2.138249025617059e-02};
static const double T33b[8] = {2.714625070149261e-01, 4.401116486585931e-01, 4.882037509455416e-01,
                               1.092578276593543e-01, 2.464636343633559e-02, 2.554542286385173e-01,
                               2.916556797383410e-01, 1.272797172335894e-01};
static const double T33w[8] = {6.254121319590276e-02, 4.991833492806094e-02, 2.426683808145203e-02,
                               2.848605206887754e-02, 7.931642509973638e-03, 4.322736365941421e-02,
                               2.178358503860756e-02, 1.508367757651144e-02};
static int tables_ready = 0;

static void init_tables(void)
{
    if (tables_ready) return;
    const double t = 1. / 3., q = sqrt(15), r = (1. + q) / 7., s = (1. - q) / 7.;
    T7a[0] = t;
    T7a[1] = t + (2. * t * s);
    T7a[2] = t + (2. * t * r);
    T7b[0] = t;
    T7b[1] = t - (t * s);
    T7b[2] = t - (t * r);
    T7w[0] = 9 / 40.;
    T7w[1] = (155. + q) / 1200.;
    T7w[2] = (155. - q) / 1200.;
    tables_ready = 1;
}

/* one node: Q = a*A + b*B + c*C, component-wise in the reference order */
static void bary(double a, double b, double c, const double A[3], const double B[3], const double C[3], double* Q)
{
    for (int k = 0; k < 3; k++) Q[k] = a * A[k] + b * B[k] + c * C[k];
}

static void tri_vertices(const double* d, double A[3], double B[3], double C[3])
{
    for (int k = 0; k < 3; k++) {
        A[k] = d[2 + k];
        B[k] = d[2 + k] + (d[0] * d[5 + k]);
        C[k] = d[2 + k] + (d[1] * d[8 + k]);
    }
}

/* node generation: KElectrostaticCubatureTriangleIntegrator.cc:34-68 (7), :70-115 (12), :212-269 (33).
   The permutation pattern of each orbit follows the reference so that the node order (and
   hence the summation order) is the reference's. */
int ko_tri_nodes(int n, const double* d, double* Q, double* w)
{
    init_tables();
    double A[3], B[3], C[3];
    tri_vertices(d, A, B, C);
    int m = 0;
    if (n == 7) {
        bary(T7a[0], T7b[0], T7b[0], A, B, C, Q + 3 * m);
        w[m++] = T7w[0];
        for (int o = 1; o < 3; o++) {
            double a = T7a[o], b = T7b[o], c = T7b[o];
            bary(a, b, c, A, B, C, Q + 3 * m);
            w[m++] = T7w[o];
            bary(b, a, c, A, B, C, Q + 3 * m);
            w[m++] = T7w[o];
            bary(c, b, a, A, B, C, Q + 3 * m);
            w[m++] = T7w[o];
        }
        return m;
    }
    if (n == 12) {
        for (int o = 0; o < 4; o++) {
            double a = T12a[o], b = T12b[o], c = T12c[o];
            bary(a, b, c, A, B, C, Q + 3 * m);
            w[m++] = T12w[o] * 2.;
            bary(b, c, a, A, B, C, Q + 3 * m);
            w[m++] = T12w[o] * 2.;
            bary(c, a, b, A, B, C, Q + 3 * m);
            w[m++] = T12w[o] * 2.;
        }
        return m;
    }
    if (n == 33) {
        for (int o = 0; o < 8; o++) {
            double a = T33a[o], b = T33b[o], c = 1. - T33a[o] - T33b[o];
            bary(a, b, c, A, B, C, Q + 3 * m);
            w[m++] = T33w[o];
            bary(b, a, c, A, B, C, Q + 3 * m);
            w[m++] = T33w[o];
            bary(c, b, a, A, B, C, Q + 3 * m);
            w[m++] = T33w[o];
            if (o >= 5) {
                bary(a, c, b, A, B, C, Q + 3 * m);
                w[m++] = T33w[o];
                bary(c, a, b, A, B, C, Q + 3 * m);
                w[m++] = T33w[o];
                bary(b, c, a, A, B, C, Q + 3 * m);
                w[m++] = T33w[o];
            }
        }
        return m;
    }
    return 0;
}

/* weighted node sum for (E, phi): ElectricFieldAndPotential_TriNP (:387-451) and _RectNP.
   prefactor = area / (4 pi eps0). */
static void node_sum(int n, const double* Q, const double* w, double prefac, const double P[3], double out[4])
{
    double s[4] = {0., 0., 0., 0.};
    for (int g = 0; g < n; g++) {
        double dv[3], r2 = 0.;
        for (int k = 0; k < 3; k++) {
            dv[k] = P[k] - Q[3 * g + k];
            r2 += SQ(dv[k]);
        }
        double ir = 1. / sqrt(r2);
        double f = ir * ir * ir;
        f = f * w[g];
        s[0] += dv[0] * f;
        s[1] += dv[1] * f;
        s[2] += dv[2] * f;
        s[3] += w[g] * ir;
    }
    for (int k = 0; k < 4; k++) out[k] = prefac * s[k];
}

static double tri_area_from_rec(const double* d)
{
    /* area exactly as in ..._TriNP: 0.5*A*B*|N1 x N2| with the reference's component order */
    return 0.5 * d[0] * d[1] *
           sqrt(SQ((d[6] * d[10]) - (d[7] * d[9])) + SQ((d[7] * d[8]) - (d[5] * d[10])) +
                SQ((d[5] * d[9]) - (d[6] * d[8])));
}

void ko_tri_cubature(int n, const double* d, const double P[3], double out[4])
{
    double Q[99], w[33];
    int m = ko_tri_nodes(n, d, Q, w);
    node_sum(m, Q, w, tri_area_from_rec(d) * KO_K, P, out);
}

/* ---------------------------------------------------------------- solid angle */
/* KSolidAngle::SolidAngleTriangleAsArray: Surfaces/include/KSolidAngle.hh:58-144 */
double ko_solid_angle_tri(const double* d, const double P[3])
{
    double V[3][3];
    tri_vertices(d, V[0], V[1], V[2]);
    double n[3];
    cross3(d + 5, d + 8, n);
    double im = 1. / norm3(n);
    for (int k = 0; k < 3; k++) n[k] *= im;
    double c[3];
    ko_centroid(KO_TRIANGLE, d, c);
    double pc[3] = {P[0] - c[0], P[1] - c[1], P[2] - c[2]};
    double h = (n[0] * pc[0]) + (n[1] * pc[1]) + (n[2] * pc[2]);
    double res;
    if (norm3(pc) <= 1.e-12) {
        res = 2. * KO_PI;
    } else {
        double u[3][3];
        for (int v = 0; v < 3; v++) {
            double dv[3] = {V[v][0] - P[0], V[v][1] - P[1], V[v][2] - P[2]};
            double mi = 1. / norm3(dv);
            for (int k = 0; k < 3; k++) u[v][k] = mi * dv[k];
        }
        double x = 1. + dot3(u[0], u[1]) + dot3(u[0], u[2]) + dot3(u[1], u[2]);
        double a12[3];
        cross3(u[1], u[2], a12);
        double y = fabs(dot3(u[0], a12));
        res = fabs(2. * atan2(y, x));
    }
    if (h < 0.) res *= -1.;
    return res;
}

static void rect_vertices(const double* d, double V[4][3])
{
    for (int k = 0; k < 3; k++) {
        V[0][k] = d[2 + k];
        V[1][k] = d[2 + k] + (d[0] * d[5 + k]);
        V[2][k] = d[2 + k] + (d[0] * d[5 + k]) + (d[1] * d[8 + k]);
        V[3][k] = d[2 + k] + (d[1] * d[8 + k]);
    }
}

/* KSolidAngle::SolidAngleRectangleAsArray: KSolidAngle.hh:166-305 (two triangles P0P1P2, P2P3P0;
   in-plane points get 2 pi inside / 0 outside) */
double ko_solid_angle_rect(const double* d, const double P[3])
{
    double V[4][3];
    rect_vertices(d, V);
    double c[3], n[3];
    for (int k = 0; k < 3; k++) c[k] = d[2 + k] + d[0] * d[5 + k] * .5 + d[1] * d[8 + k] * .5;
    cross3(d + 5, d + 8, n); /* not normalised in the reference (N1, N2 orthonormal) */
    double h = (n[0] * (P[0] - c[0])) + (n[1] * (P[1] - c[1])) + (n[2] * (P[2] - c[2]));
    double res;
    if (fabs(h) < 1.e-12) {
        double s1[3], s2[3], dp[3];
        for (int k = 0; k < 3; k++) {
            s1[k] = d[0] * d[5 + k];
            s2[k] = d[1] * d[8 + k];
            dp[k] = d[2 + k] - P[k];
        }
        double l1 = (-1) * dot3(dp, s1) / SQ(d[0]);
        double l2 = (-1) * dot3(dp, s2) / SQ(d[1]);
        res = (l1 > 0. && l1 < 1. && l2 > 0. && l2 < 1.) ? 2. * KO_PI : 0.;
    } else {
        double u[4][3];
        for (int v = 0; v < 4; v++) {
            double dv[3] = {V[v][0] - P[0], V[v][1] - P[1], V[v][2] - P[2]};
            double mi = 1. / norm3(dv);
            for (int k = 0; k < 3; k++) u[v][k] = mi * dv[k];
        }
        double x1 = 1. + dot3(u[0], u[1]) + dot3(u[0], u[2]) + dot3(u[1], u[2]);
        double a12[3];
        cross3(u[1], u[2], a12);
        double y1 = fabs(dot3(u[0], a12));
        double x2 = 1. + dot3(u[2], u[3]) + dot3(u[2], u[0]) + dot3(u[3], u[0]);
        double a30[3];
        cross3(u[3], u[0], a30);
        double y2 = fabs(dot3(u[2], a30));
        res = fabs(2. * atan2(y1, x1)) + fabs(2. * atan2(y2, x2));
    }
    if (h < 0.) res *= -1.;
    return res;
}

/* ------------------------------------------------------ closed-form near field */
/* Common structure of KElectrostaticRWGTriangleIntegrator.cc and ...RectangleIntegrator.cc:
   phi = (-h*Omega - sum_e t_e*L_e)/(4 pi eps0), E = (n*Omega + sum_e m_e*L_e)/(4 pi eps0),
   L_e = ln((r+ + s+)/(r- + s-)) along edge e. */
typedef struct {
    int ne;           /* edges: 3 (triangle) or 4 (rectangle) */
    double V[4][3];   /* vertices in edge order */
    double u[4][3];   /* unit edge directions */
    double mid[4][3]; /* edge midpoints */
    double m[4][3];   /* in-plane outward normals u x n */
    double n[3];      /* unit plane normal */
    double c[3];      /* centroid used for h */
    double sv[4][3];  /* full edge vectors (edge detection) */
    double invlen[4]; /* 1/edge length (edge detection) */
} ko_poly;

static void poly_tri(const double* d, ko_poly* p)
{
    p->ne = 3;
    tri_vertices(d, p->V[0], p->V[1], p->V[2]);
    cross3(d + 5, d + 8, p->n);
    double im = 1. / norm3(p->n);
    for (int k = 0; k < 3; k++) p->n[k] = p->n[k] * im;
    ko_centroid(KO_TRIANGLE, d, p->c);
    /* IqL*: RWG/src/KElectrostaticRWGTriangleIntegrator.cc:20-100 */
    double e12[3] = {p->V[2][0] - p->V[1][0], p->V[2][1] - p->V[1][1], p->V[2][2] - p->V[1][2]};
    double len12 = sqrt(SQ(e12[0]) + SQ(e12[1]) + SQ(e12[2]));
    double i12 = 1. / len12;
    double hl[3] = {0.5 * d[0], 0.5 * len12, 0.5 * d[1]};
    for (int k = 0; k < 3; k++) {
        p->u[0][k] = d[5 + k];
        p->u[1][k] = i12 * e12[k];
        p->u[2][k] = -d[8 + k];
        /* edge detection vectors: ...TriangleIntegrator.cc:645-657 */
        p->sv[0][k] = d[0] * d[5 + k];
        p->sv[1][k] = e12[k];
        p->sv[2][k] = (-1) * d[1] * d[8 + k];
    }
    p->invlen[0] = 1. / d[0];
    p->invlen[1] = 1. / len12;
    p->invlen[2] = 1. / d[1];
    for (int e = 0; e < 3; e++) {
        for (int k = 0; k < 3; k++) p->mid[e][k] = hl[e] * p->u[e][k] + p->V[e][k];
        cross3(p->u[e], p->n, p->m[e]);
    }
}

static void poly_rect(const double* d, ko_poly* p)
{
    p->ne = 4;
    rect_vertices(d, p->V);
    cross3(d + 5, d + 8, p->n);
    double im = 1. / sqrt(SQ(p->n[0]) + SQ(p->n[1]) + SQ(p->n[2]));
    for (int k = 0; k < 3; k++) p->n[k] = p->n[k] * im;
    /* RWG rectangle centroid: ...RectangleIntegrator.cc:1212-1214 */
    for (int k = 0; k < 3; k++) p->c[k] = d[2 + k] + 0.5 * (d[0] * d[5 + k]) + 0.5 * (d[1] * d[8 + k]);
    for (int k = 0; k < 3; k++) {
        p->u[0][k] = d[5 + k];
        p->u[1][k] = d[8 + k];
        p->u[2][k] = -d[5 + k];
        p->u[3][k] = -d[8 + k];
        /* midpoints: ...RectangleIntegrator.cc:60-75 */
        p->mid[0][k] = (0.5 * d[0] * d[5 + k]) + p->V[0][k];
        p->mid[1][k] = (0.5 * d[1] * d[8 + k]) + p->V[1][k];
        p->mid[2][k] = (-0.5 * d[0] * d[5 + k]) + p->V[2][k];
        p->mid[3][k] = (-0.5 * d[1] * d[8 + k]) + p->V[3][k];
        p->sv[0][k] = d[0] * d[5 + k];
        p->sv[1][k] = d[1] * d[8 + k];
        p->sv[2][k] = -d[0] * d[5 + k];
        p->sv[3][k] = -d[1] * d[8 + k];
    }
    p->invlen[0] = 1. / d[0];
    p->invlen[1] = 1. / d[1];
    p->invlen[2] = 1. / d[0];
    p->invlen[3] = 1. / d[1];
    for (int e = 0; e < 4; e++) cross3(p->u[e], p->n, p->m[e]);
}

/* LogArgTaylor: ...TriangleIntegrator.cc:8-18 */
static double log_arg_taylor(double sM, double dist)
{
    double q = fabs(dist / sM);
    if (q < 1.e-14) q = 1.e-14;
    return 0.5 * fabs(sM) * SQ(q);
}

/* line-integral term IqLFieldAndPotential (...TriangleIntegrator.cc:393-604):
   returns iL[0..2] (field part) and iL[3] (potential part) */
static void iql(const ko_poly* p, const double P[3], int count, int line, double dist, double iL[4])
{
    double D[4][3], r[4], t[4];
    for (int v = 0; v < p->ne; v++) {
        for (int k = 0; k < 3; k++) D[v][k] = p->V[v][k] - P[k];
        r[v] = sqrt(SQ(D[v][0]) + SQ(D[v][1]) + SQ(D[v][2]));
    }
    for (int e = 0; e < p->ne; e++)
        t[e] = (p->m[e][0] * (P[0] - p->mid[e][0])) + (p->m[e][1] * (P[1] - p->mid[e][1])) +
               (p->m[e][2] * (P[2] - p->mid[e][2]));
    iL[0] = iL[1] = iL[2] = iL[3] = 0.;
    for (int e = 0; e < p->ne; e++) {
        int e1 = (e + 1) % p->ne;
        double rM = r[e], rP = r[e1];
        double sM = dot3(D[e], p->u[e]), sP = dot3(D[e1], p->u[e]);
        double val;
        int have = 0;
        if (count == 1 && line == e) {
            if (fabs(dist / sM) < 1.E-6) {
                val = log(rP + sP) - log(log_arg_taylor(sM, dist));
                iL[0] += p->m[e][0] * val;
                iL[1] += p->m[e][1] * val;
                iL[2] += p->m[e][2] * val;
                iL[3] += t[e] * val;
            }
        }
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
            have = 1;
        }
        if (have) {
            iL[0] += p->m[e][0] * val;
            iL[1] += p->m[e][1] * val;
            iL[2] += p->m[e][2] * val;
            iL[3] += t[e] * val;
        }
    }
}

/* point-on-edge detection: ...TriangleIntegrator.cc:667-737 (fMinDistanceToSideLine 1e-14,
   fToleranceLambda 1e-15) */
static int edge_check(const ko_poly* p, const double P[3], int* line, double* distmin)
{
    int count = 0;
    *line = 99;
    *distmin = 0.;
    for (int e = 0; e < p->ne; e++) {
        double D[3] = {p->V[e][0] - P[0], p->V[e][1] - P[1], p->V[e][2] - P[2]};
        double cv[3];
        cross3(p->sv[e], D, cv);
        double dl = sqrt(SQ(cv[0]) + SQ(cv[1]) + SQ(cv[2])) * p->invlen[e];
        double lam = ((-D[0] * p->sv[e][0]) + (-D[1] * p->sv[e][1]) + (-D[2] * p->sv[e][2])) * SQ(p->invlen[e]);
        if (dl < 1.E-14 && lam >= -1.E-15 && lam <= (1. + 1.E-15)) {
            *distmin = dl;
            count++;
            *line = e;
        }
    }
    return count;
}

static void rwg_eval(const ko_poly* p, int shape, const double* d, const double P[3], double out[4])
{
    int line;
    double dist;
    int count = edge_check(p, P, &line, &dist);
    if (count == 2) {
        /* field point on a vertex: average of points moved by +-1e-7 along the normal (:730-771) */
        double acc[4] = {0., 0., 0., 0.};
        for (int s = 0; s < 2; s++) {
            double sg = s == 0 ? 1. : -1.;
            double Q[3] = {P[0] + sg * 1.E-7 * p->n[0], P[1] + sg * 1.E-7 * p->n[1], P[2] + sg * 1.E-7 * p->n[2]};
            double h = (p->n[0] * (Q[0] - p->c[0])) + (p->n[1] * (Q[1] - p->c[1])) + (p->n[2] * (Q[2] - p->c[2]));
            double om = shape == KO_TRIANGLE ? ko_solid_angle_tri(d, Q) : ko_solid_angle_rect(d, Q);
            double iL[4];
            iql(p, Q, 9, 9, 9., iL);
            for (int k = 0; k < 3; k++) acc[k] += (p->n[k] * om + iL[k]);
            acc[3] += (-h * om - iL[3]);
        }
        for (int k = 0; k < 4; k++) out[k] = KO_K * 0.5 * acc[k];
        return;
    }
    double h = (p->n[0] * (P[0] - p->c[0])) + (p->n[1] * (P[1] - p->c[1])) + (p->n[2] * (P[2] - p->c[2]));
    double om = shape == KO_TRIANGLE ? ko_solid_angle_tri(d, P) : ko_solid_angle_rect(d, P);
    double iL[4];
    iql(p, P, count, line, dist, iL);
    for (int k = 0; k < 3; k++) out[k] = KO_K * (p->n[k] * om + iL[k]);
    out[3] = KO_K * ((-h * om) - iL[3]);
}

void ko_tri_rwg(const double* d, const double P[3], double out[4])
{
    ko_poly p;
    poly_tri(d, &p);
    rwg_eval(&p, KO_TRIANGLE, d, P, out);
}

void ko_rect_rwg(const double* d, const double P[3], double out[4])
{
    ko_poly p;
    poly_rect(d, &p);
    rwg_eval(&p, KO_RECTANGLE, d, P, out);
}

/* ---------------------------------------------------- triangle numeric dispatch */
/* KElectrostaticCubatureTriangleIntegrator::ElectricFieldAndPotential (:595-665):
   size = mean side, dr = |centroid - P| / size; cut-offs 131.4 / 30.4 / 3.42 (.hh:51-53) */
int ko_tri_numeric(const double* d, const double P[3], double out[4])
{
    double A[3], B[3], C[3];
    tri_vertices(d, A, B, C);
    double l12 = sqrt(SQ(C[0] - B[0]) + SQ(C[1] - B[1]) + SQ(C[2] - B[2]));
    double size = (d[0] + d[1] + l12) / 3.;
    double c[3];
    ko_centroid(KO_TRIANGLE, d, c);
    double dv[3] = {c[0] - P[0], c[1] - P[1], c[2] - P[2]};
    double dr = sqrt(SQ(dv[0]) + SQ(dv[1]) + SQ(dv[2])) * (1. / size);
    int n = dr > 131.4 ? 7 : dr > 30.4 ? 12 : dr > 3.42 ? 33 : 0;
    if (n)
        ko_tri_cubature(n, d, P, out);
    else
        ko_tri_rwg(d, P, out);
    return n;
}

/* ------------------------------------------------------ rectangle cubature */
/* Numeric/include/KElectrostaticCubatureRectangleIntegrator.hh:64-160 and
   Numeric/src/...RectangleIntegrator.cc:43-290: nodes Q = centre + x*(A/2)*N1 + y*(B/2)*N2
   on [-1,1]^2 (Stroud 5-2 7-point, Tyler 12-point, 33-point degree-15 rule); weights sum 1. */
int ko_rect_nodes(int n, const double* d, double* Q, double* w)
{
    double x[33], y[33];
    int m = 0;
    if (n == 7) {
        const double t2 = sqrt((7. - sqrt(24.)) / 15.), s2 = sqrt((7. + sqrt(24.)) / 15.), r2 = sqrt(7. / 15.);
        const double xs[7] = {0., r2, -r2, s2, -s2, t2, -t2};
        const double ys[7] = {0., r2, -r2, -t2, t2, -s2, s2};
        const double ws[7] = {2. / 7., 25. / 168., 25. / 168., 5. / 48., 5. / 48., 5. / 48., 5. / 48.};
        for (m = 0; m < 7; m++) {
            x[m] = xs[m];
            y[m] = ys[m];
            w[m] = ws[m];
        }
    } else if (n == 12) {
        const double r = sqrt(6. / 7.), s = sqrt((114. - 3. * sqrt(583.)) / 287.),
                     t = sqrt((114. + 3. * sqrt(583.)) / 287.);
        const double xs[12] = {r, -r, 0., 0., s, s, -s, -s, t, t, -t, -t};
        const double ys[12] = {0., 0., r, -r, s, -s, s, -s, t, -t, t, -t};
        const double B1 = 49 / 810., B2 = (178981. + 2769. * sqrt(583)) / 1888920.,
                     B3 = (178981. - 2769. * sqrt(583)) / 1888920.;
        for (m = 0; m < 12; m++) {
            x[m] = xs[m];
            y[m] = ys[m];
            w[m] = m < 4 ? B1 : m < 8 ? B2 : B3;
        }
    } else if (n == 33) {
        static const double X[9] = {0.00000000000000000000, 0.77880971155441942252, 0.95729769978630736566,
                                    0.13818345986246535375, 0.94132722587292523695, 0.47580862521827590507,
                                    0.75580535657208143627, 0.69625007849174941396, 0.34271655604040678941};
        static const double Y[9] = {0.00000000000000000000, 0.98348668243987226379, 0.85955600564163892859,
                                    0.95892517028753485754, 0.39073621612946100068, 0.85007667369974857597,
                                    0.64782163718701073204, 0.70741508996444936217e-1, 0.40930456169403884330};
        static const double W[9] = {0.30038211543122536139 / 4.,  0.29991838864499131666e-1 / 4.,
                                    0.38174421317083669640e-1 / 4., 0.60424923817749980681e-1 / 4.,
                                    0.77492738533105339358e-1 / 4., 0.11884466730059560108 / 4.,
                                    0.12976355037000271129 / 4.,  0.21334158145718938943 / 4.,
                                    0.25687074948196783651 / 4.};
        x[0] = X[0];
        y[0] = Y[0];
        w[0] = W[0];
        m = 1;
        /* each orbit: (X,Y), (-X,-Y), (-Y,X), (Y,-X) (.cc:250-257) */
        for (int o = 1; o < 9; o++) {
            double xo[4] = {X[o], -X[o], -Y[o], Y[o]}, yo[4] = {Y[o], -Y[o], X[o], -X[o]};
            for (int q = 0; q < 4; q++) {
                x[m] = xo[q];
                y[m] = yo[q];
                w[m] = W[o];
                m++;
            }
        }
    } else {
        return 0;
    }
    const double a2 = 0.5 * d[0], b2 = 0.5 * d[1];
    double cen[3];
    for (int k = 0; k < 3; k++) cen[k] = d[2 + k] + a2 * d[5 + k] + b2 * d[8 + k];
    for (int i = 0; i < m; i++)
        for (int k = 0; k < 3; k++) Q[3 * i + k] = cen[k] + (x[i] * a2) * d[5 + k] + (y[i] * b2) * d[8 + k];
    return m;
}

void ko_rect_cubature(int n, const double* d, const double P[3], double out[4])
{
    double Q[99], w[33];
    int m = ko_rect_nodes(n, d, Q, w);
    node_sum(m, Q, w, d[0] * d[1] * KO_K, P, out);
}

/* KElectrostaticCubatureRectangleIntegrator::ElectricFieldAndPotential: size=(A+B)/2,
   centre P0 + 0.5*(A N1 + B N2), cut-offs 196.2 / 44.6 / 4.6 (.hh:52-54) */
int ko_rect_numeric(const double* d, const double P[3], double out[4])
{
    double size = 0.5 * (d[0] + d[1]);
    double c[3];
    for (int k = 0; k < 3; k++) c[k] = d[2 + k] + 0.5 * (d[0] * d[5 + k] + d[1] * d[8 + k]);
    double dv[3] = {c[0] - P[0], c[1] - P[1], c[2] - P[2]};
    double dr = sqrt(SQ(dv[0]) + SQ(dv[1]) + SQ(dv[2])) * (1. / size);
    int n = dr > 196.2 ? 7 : dr > 44.6 ? 12 : dr > 4.6 ? 33 : 0;
    if (n)
        ko_rect_cubature(n, d, P, out);
    else
        ko_rect_rwg(d, P, out);
    return n;
}

/* ------------------------------------------------------------ line segment */
/* Gauss-Legendre half-tables: Numeric/include/KElectrostaticQuadratureLineSegmentIntegrator.hh:59-83 */
static const double GLx4[2] = {0.339981043584856, 0.861136311594053};
static const double GLw4[2] = {0.652145154862546, 0.347854845137454};
static const double GLx16[8] = {0.09501250983763744, 0.28160355077925891, 0.45801677765722739,
                                0.61787624440264375, 0.75540440835500303, 0.86563120238783174,
                                0.94457502307323258, 0.98940093499164993};
static const double GLw16[8] = {0.189450610455068496, 0.182603415044923589, 0.169156519395002532,
                                0.149595988816576731, 0.124628971255533872, 0.095158511682492785,
                                0.062253523938647892, 0.027152459411754095};

/* ElectricFieldAndPotential_nNodes (...LineSegmentIntegrator.cc:126-191): symmetric node pairs
   about the midpoint, prefactor d/(4 eps0) */
void ko_line_quadrature(int nn, const double* d, const double P[3], double out[4])
{
    const double* xs = nn == 4 ? GLx4 : GLx16;
    const double* ws = nn == 4 ? GLw4 : GLw16;
    int half = nn / 2;
    double c[3] = {0.5 * (d[0] + d[3]), 0.5 * (d[1] + d[4]), 0.5 * (d[2] + d[5])};
    double L = sqrt(SQ(d[3] - d[0]) + SQ(d[4] - d[1]) + SQ(d[5] - d[2]));
    double hl = 0.5 * L, iu = 1. / L;
    double u[3] = {iu * (d[3] - d[0]), iu * (d[4] - d[1]), iu * (d[5] - d[2])};
    double sf[3] = {0., 0., 0.}, sp = 0.;
    for (int i = 0; i < half; i++) {
        double wf = ws[i] * hl, x = xs[i] * hl;
        double pd[3], nd[3], pm = 0., nm = 0.;
        for (int k = 0; k < 3; k++) {
            pd[k] = P[k] - (c[k] + (u[k] * x));
            pm += SQ(pd[k]);
            nd[k] = P[k] - (c[k] - (u[k] * x));
            nm += SQ(nd[k]);
        }
        pm = 1. / sqrt(pm);
        nm = 1. / sqrt(nm);
        sp += wf * (pm + nm);
        pm = pm * pm * pm;
        nm = nm * nm * nm;
        for (int k = 0; k < 3; k++) sf[k] += wf * (pm * pd[k] + nm * nd[k]);
    }
    double pre = 0.25 * d[6] / KO_EPS0;
    for (int k = 0; k < 3; k++) out[k] = pre * sf[k];
    out[3] = pre * sp;
}

/* KElectrostaticAnalyticLineSegmentIntegrator::Potential / ElectricField (.cc:15-83):
   ln((Da+Db+L)/(Da+Db-L)) with the inside-the-wire substitution */
void ko_line_analytic(const double* d, const double P[3], double out[4])
{
    const double* p0 = d;
    const double* p1 = d + 3;
    double dia = d[6];
    double l[3] = {p1[0] - p0[0], p1[1] - p0[1], p1[2] - p0[2]};
    double L = norm3(l);
    double a[3] = {p0[0] - P[0], p0[1] - P[1], p0[2] - P[2]};
    double b[3] = {p1[0] - P[0], p1[1] - P[1], p1[2] - P[2]};
    double Da = norm3(a), Db = norm3(b);
    /* potential */
    double DaP = Da, DbP = Db, DaE = Da, DbE = Db;
    if (!((Da + Db) > (L + dia))) {
        double u[3] = {l[0] / L, l[1] / L, l[2] / L};
        double pp[3] = {P[0] - p0[0], P[1] - p0[1], P[2] - p0[2]};
        double p = dot3(pp, u);
        if (!(p < (-dia * .5) || p > (L + dia * .5))) {
            double pf[3] = {p0[0] + p * u[0], p0[1] + p * u[1], p0[2] + p * u[2]};
            double dd[3] = {P[0] - pf[0], P[1] - pf[1], P[2] - pf[2]};
            if (!(norm3(dd) >= dia * .5)) {
                double qa[3] = {p0[0] - pf[0], p0[1] - pf[1], p0[2] - pf[2]};
                double qb[3] = {p1[0] - pf[0], p1[1] - pf[1], p1[2] - pf[2]};
                double qa2 = SQ(qa[0]) + SQ(qa[1]) + SQ(qa[2]);
                double qb2 = SQ(qb[0]) + SQ(qb[1]) + SQ(qb[2]);
                DaP = sqrt(qa2 + dia * dia * .25);
                DbP = sqrt(qb2 + dia * dia * .25);
                DaE = norm3(qa);
                DbE = norm3(qb);
            }
        }
    }
    double ln = log((DaP + DbP + L) / (DaP + DbP - L));
    out[3] = dia / (4. * KO_EPS0) * ln;
    double den = (DaE * (DaE + DbE + L) * (DaE + DbE - L) * DbE) * 4. * KO_EPS0;
    den = -1. / den;
    for (int k = 0; k < 3; k++)
        out[k] = (2. * L * (p1[k] * DaE - P[k] * (DaE + DbE) + p0[k] * DbE)) * dia * den;
}

/* KElectrostaticQuadratureLineSegmentIntegrator::ElectricFieldAndPotential (.cc:266-310):
   dr = |midpoint - P|/L; > 27.7 -> 4 nodes, > 2.0 -> 16 nodes, else analytic */
int ko_line_numeric(const double* d, const double P[3], double out[4])
{
    double L = sqrt(((d[3] - d[0]) * (d[3] - d[0])) + ((d[4] - d[1]) * (d[4] - d[1])) +
                    ((d[5] - d[2]) * (d[5] - d[2])));
    double c[3] = {(d[0] + d[3]) * 0.5, (d[1] + d[4]) * 0.5, (d[2] + d[5]) * 0.5};
    double dv[3] = {c[0] - P[0], c[1] - P[1], c[2] - P[2]};
    double dr = sqrt(SQ(dv[0]) + SQ(dv[1]) + SQ(dv[2])) * (1. / L);
    if (dr > 27.7) {
        ko_line_quadrature(4, d, P, out);
        return 4;
    }
    if (dr > 2.) {
        ko_line_quadrature(16, d, P, out);
        return 16;
    }
    ko_line_analytic(d, P, out);
    return 0;
}

/* ------------------------------------------------------------ analytic policy */
/* KElectrostaticAnalyticTriangleIntegrator.cc (Analytic/src): the triangle is put in a local
   frame with the field point on the z axis; the potential and the field are sums of the
   one-dimensional primitives below.  Branch structure, thresholds and expression order follow
   the cited lines. */
#define SGN(v) (((v) > 0.) - ((v) < 0.))
static const double KO_PI_2 = 3.141592653589793238462643383279502884 / 2.;

/* F1 :209-212 */
static double at_F1(double a, double b, double u) { return u * asinh((a + b * u) / sqrt(u * u + 1.)); }

/* shared g1/g2 of I3 and I3p with the negative-argument patch :216-227, :243-253 */
static void at_g12(double a, double b, double u1, double u2, double* g1, double* g2)
{
    *g1 = (sqrt(b * b + 1.) * sqrt(a * a + 2 * a * b * u1 + (b * b + 1.) * u1 * u1 + 1.) + b * (a + b * u1) + u1);
    *g2 = (sqrt(b * b + 1.) * sqrt(a * a + 2 * a * b * u2 + (b * b + 1.) * u2 * u2 + 1.) + b * (a + b * u2) + u2);
    if (*g1 <= 0) *g1 = -(1 + a * a + b * b) / (2 * (b * b + 1) * u1);
    if (*g2 <= 0) *g2 = -(1 + a * a + b * b) / (2 * (b * b + 1) * u2);
}

/* I3 :214-241 */
static double at_I3(double a, double b, double u1, double u2)
{
    double g1, g2;
    at_g12(a, b, u1, u2, &g1, &g2);
    if (fabs(g1) < 1.e-12 && fabs(a) < 1.e-14) g1 = 1.e-12;
    if (fabs(g2) < 1.e-12 && fabs(a) < 1.e-14) g2 = 1.e-12;
    return a / sqrt(b * b + 1.) * log(fabs(g2 / g1));
}

/* I3p :243-256 */
static double at_I3p(double a, double b, double u1, double u2)
{
    double g1, g2;
    at_g12(a, b, u1, u2, &g1, &g2);
    return 1. / sqrt(b * b + 1.) * log(fabs(g2 / g1));
}

/* I4 in the substituted variable t :258-275 */
static double at_I4t(double alpha, double gamma, double q2, double prefac, double t1, double t2)
{
    double q = sqrt(q2);
    double g1 = sqrt(gamma * t1 * t1 + alpha);
    double g2 = sqrt(gamma * t2 * t2 + alpha);
    if (t1 > 1.e15 || t2 > 1.e15) {
        if (t2 < 1.e15) return (prefac * 1. / q * (atan(g2 / q) - KO_PI_2));
        if (t1 < 1.e15) return (prefac * 1. / q * (KO_PI_2 - atan(g1 / q)));
        return 0.;
    }
    return prefac * 1. / q * atan(q * (g2 - g1) / (q2 + g1 * g2));
}

/* I4 :277-333.  The reference nudges u1, u2 or a by 1e-14 and re-enters; the loop below takes
   the same steps (bounded: each nudge leaves the 1e-14 window within two steps). */
static double at_I4(double a, double b, double u1, double u2)
{
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
    double alpha = ((double)1.) + (a * a) / (b * b);
    double gamma = (a * a + b * b) * (a * a + b * b + ((double)1.)) / (b * b);
    double q2 = (a * a + b * b) * (a * a + b * b) / (b * b);
    double prefac = (a * a / b + b);
    double t1 = (a * u1 != b) ? (b * u1 + a) / (a * u1 - b) : 1.e15;
    double t2 = (a * u2 != b) ? (b * u2 + a) / (a * u2 - b) : 1.e15;
    double sign = 1.;
    if (a < 0.) sign = -sign;
    if (b < 0.) sign = -sign;
    if (u1 > b / a) sign = -sign;
    if (SGN(u1 - b / a) != SGN(u2 - b / a))
        return sign * (at_I4t(alpha, gamma, q2, prefac, t1, 1.e16) + at_I4t(alpha, gamma, q2, prefac, t2, 1.e16));
    return sign * at_I4t(alpha, gamma, q2, prefac, t1, t2);
}

/* I4_2 in t :335-346 */
static double at_I42t(double alpha, double gamma, double prefac, double t1, double t2)
{
    double g1 = sqrt(alpha * t1 * t1 + gamma);
    double g2 = sqrt(alpha * t2 * t2 + gamma);
    if (gamma - alpha <= 0) return prefac * (g2 - g1) / ((alpha - gamma) + g1 * g2);
    double q = sqrt(gamma - alpha);
    return prefac * 1. / q * atanh(q * (g2 - g1) / ((alpha - gamma) + g1 * g2));
}

/* I4_2 :348-377 */
static double at_I42(double a, double b, double u1, double u2)
{
    double alpha = ((double)1.) + (a * a) / (b * b);
    double gamma = (a * a + b * b) * (a * a + b * b + 1) / (b * b);
    double lambda = -a / b;
    double prefac = (a * a / b + b);
    double t1 = (b * u1 != -a) ? (b - a * u1) / (a + b * u1) : 1.e15;
    double t2 = (b * u2 != -a) ? (b - a * u2) / (a + b * u2) : 1.e15;
    if (SGN(u1 - lambda) != SGN(u2 - lambda)) {
        if (u1 > lambda)
            return (at_I42t(alpha, gamma, prefac, fabs(t1), 1.e15) + at_I42t(alpha, gamma, prefac, fabs(t2), 1.e15));
        return (at_I42t(alpha, gamma, prefac, 1.e15, fabs(t1)) + at_I42t(alpha, gamma, prefac, 1.e15, fabs(t2)));
    }
    if (u1 > lambda) return at_I42t(alpha, gamma, prefac, t1, t2);
    return at_I42t(alpha, gamma, prefac, t2, t1);
}

/* I1 :379-382 */
static double at_I1(double a, double b, double u1, double u2)
{
    return at_F1(a, b, u2) - at_F1(a, b, u1) + at_I3(a, b, u1, u2) - at_I4(a, b, u1, u2);
}

/* I6 :384-389 */
static double at_I6(double x, double u1, double u2)
{
    if (fabs(x) < 1.e-15) return 0;
    return x * log((sqrt(u2 * u2 + x * x + 1.) + u2) / (sqrt(u1 * u1 + x * x + 1.) + u1));
}

/* I7 :391-408 */
static double at_I7(double x, double u1, double u2)
{
    double t1 = (fabs(u1) > 1.e-16) ? 1. / u1 : 1.e16;
    double t2 = (fabs(u2) > 1.e-16) ? 1. / u2 : 1.e16;
    double g1 = sqrt(1. + t1 * t1 * (1. + x * x));
    double g2 = sqrt(1. + t2 * t2 * (1. + x * x));
    return atan(x * (g2 - g1) / (x * x + g2 * g1));
}

/* I2 :410-427 */
static double at_I2(double x, double u1, double u2)
{
    double f = (at_F1(x, 0., u2) - at_F1(x, 0., u1));
    if (SGN(u1) != SGN(u2)) {
        if (u1 <= 0.) return f + at_I6(x, u1, u2) + at_I7(x, 0., fabs(u1)) + at_I7(x, 0., fabs(u2));
        return f + at_I6(x, u1, u2) + at_I7(x, fabs(u1), 0.) + at_I7(x, fabs(u2), 0.);
    }
    if (u1 <= 0.) return f + at_I6(x, u1, u2) + at_I7(x, u2, u1);
    return f + at_I6(x, u1, u2) + at_I7(x, u1, u2);
}

/* J2 :429-438 */
static double at_J2(double a, double u1, double u2)
{
    if (a == 0.) return 0.;
    double g1 = sqrt(u1 * u1 + a * a + 1.);
    double g2 = sqrt(u2 * u2 + a * a + 1.);
    return a / (2. * fabs(a)) * log(((g2 - fabs(a)) * (g1 + fabs(a))) / ((g2 + fabs(a)) * (g1 - fabs(a))));
}

/* Local_Ez's per-edge term :473-487 */
static double at_ez_edge(double a, double b, double u0, double u1)
{
    if (fabs(b) > 1.e-14) return at_I4(a, b, u0, u1);
    if (SGN(u0) != SGN(u1)) {
        if (u0 <= 0.) return -(at_I7(a, 0., fabs(u0)) + at_I7(a, 0., fabs(u1)));
        return -(at_I7(a, fabs(u0), 0.) + at_I7(a, fabs(u1), 0.));
    }
    if (u0 <= 0.) return at_I7(a, u0, u1);
    return at_I7(a, u1, u0);
}

/* Local_Ey's per-edge term :452-462 */
static double at_ey_edge(double a, double b, double u0, double u1)
{
    if (fabs(b) > 1.e-14) return at_I42(a, b, u0, u1) + b * at_I3p(a, b, u0, u1);
    return at_J2(a, u0, u1);
}

/* Potential_noZ's one-edge term :178-208 (ylim: 1e-14 for the upper edge, 5e-14 for the lower) */
static double at_noz_edge(double a, double b, double y, double ylim)
{
    double logArg = (1. + b * b) * y + a * b + sqrt(1. + b * b) * sqrt((1. + b * b) * y * y + 2 * a * b * y + a * a);
    if (logArg > 1.e-14) {
        if (fabs(y) > ylim) return y * asinh((a + b * y) / fabs(y)) + a / sqrt(1. + b * b) * log(logArg);
        return a / sqrt(1. + b * b) * log(logArg);
    }
    if (fabs(y) > ylim) return y * asinh(y * b / fabs(y));
    return 0.;
}

static double at_noz(double a2, double b2, double a1, double b1, double y)
{
    return at_noz_edge(a2, b2, y, 1.e-14) - at_noz_edge(a1, b1, y, 5.e-14);
}

/* local frame of Potential :13-70 / ElectricField :84-136 */
typedef struct {
    double N2p[3], a[2], b[2], u[2], z, zsign;
} at_frame;

static void at_local(const double* d, const double P[3], at_frame* f)
{
    const double* N1 = d + 5;
    const double* N2 = d + 8;
    double N3[3];
    ko_normal(KO_TRIANGLE, d, N3);
    double p[3] = {d[2] - P[0], d[3] - P[1], d[4] - P[2]};
    double n1dotn2 = dot3(N1, N2);
    double t[3] = {N2[0] - n1dotn2 * N1[0], N2[1] - n1dotn2 * N1[1], N2[2] - n1dotn2 * N1[2]};
    double tm = norm3(t);
    for (int k = 0; k < 3; k++) f->N2p[k] = tm > 0. ? t[k] / tm : t[k];
    double n2dotn2p = dot3(N2, f->N2p);
    double x0 = dot3(p, N1), y0 = dot3(p, f->N2p), z0 = -dot3(p, N3);
    double x1 = x0 + d[0];
    double x2 = x0 + d[1] * n1dotn2;
    double y1;
    if (z0 < 0) {
        y0 = -y0;
        y1 = y0 - d[1] * n2dotn2p;
        f->zsign = -1;
    } else {
        y1 = y0 + d[1] * n2dotn2p;
        f->zsign = 1;
    }
    double z = fabs(z0);
    f->z = z;
    if (z > 1.e-14) {
        f->u[0] = y0 / z;
        if (f->u[0] == 0.) f->u[0] = 1.e-18;
        f->u[1] = y1 / z;
        if (f->u[1] == 0.) f->u[1] = 1.e-18;
        f->a[0] = (x0 * y1 - x2 * y0) / (z * (y1 - y0));
        f->a[1] = (x1 * y1 - x2 * y0) / (z * (y1 - y0));
    } else {
        f->u[0] = y0;
        if (f->u[0] == 0.) f->u[0] = 1.e-18;
        f->u[1] = y1;
        if (f->u[1] == 0.) f->u[1] = 1.e-18;
        f->a[0] = (x0 * y1 - x2 * y0) / (y1 - y0);
        f->a[1] = (x1 * y1 - x2 * y0) / (y1 - y0);
    }
    f->b[0] = (x2 - x0) / (y1 - y0);
    f->b[1] = (x2 - x1) / (y1 - y0);
}

/* Potential :13-83 and ElectricField :85-174 */
void ko_tri_analytic(const double* d, const double P[3], double out[4])
{
    at_frame f;
    at_local(d, P, &f);
    double I;
    if (f.z > 1.e-14) {
        if (fabs(f.b[0]) < 1.e-13)
            I = f.z * (at_I1(f.a[1], f.b[1], f.u[0], f.u[1]) - at_I2(f.a[0], f.u[0], f.u[1]));
        else if (fabs(f.b[1]) < 1.e-13)
            I = f.z * (at_I2(f.a[1], f.u[0], f.u[1]) - at_I1(f.a[0], f.b[0], f.u[0], f.u[1]));
        else
            I = f.z * (at_I1(f.a[1], f.b[1], f.u[0], f.u[1]) - at_I1(f.a[0], f.b[0], f.u[0], f.u[1]));
    } else {
        I = (at_noz(f.a[1], f.b[1], f.a[0], f.b[0], f.u[1]) - at_noz(f.a[1], f.b[1], f.a[0], f.b[0], f.u[0]));
    }
    out[3] = fabs(I) / (4. * KO_PI * KO_EPS0);

    double c[3];
    ko_centroid(KO_TRIANGLE, d, c);
    double dc[3] = {c[0] - P[0], c[1] - P[1], c[2] - P[2]};
    double dist = norm3(dc);
    double prefac = 1. / (4. * KO_PI * KO_EPS0);
    double loc[3];
    loc[0] = f.zsign * prefac * (at_I3p(f.a[1], f.b[1], f.u[0], f.u[1]) - at_I3p(f.a[0], f.b[0], f.u[0], f.u[1]));
    loc[1] = prefac * (at_ey_edge(f.a[0], f.b[0], f.u[0], f.u[1]) - at_ey_edge(f.a[1], f.b[1], f.u[0], f.u[1]));
    if (f.z > 1.e-14)
        loc[2] = prefac * (at_ez_edge(f.a[1], f.b[1], f.u[0], f.u[1]) - at_ez_edge(f.a[0], f.b[0], f.u[0], f.u[1]));
    else
        loc[2] = (dist < 1.e-12) ? 1. / (2. * KO_EPS0) : 0.;
    double N3[3];
    ko_normal(KO_TRIANGLE, d, N3);
    for (int j = 0; j < 3; j++) out[j] = (d[5 + j] * loc[0] + f.N2p[j] * loc[1] + N3[j] * loc[2]);
}

/* KElectrostaticAnalyticRectangleIntegrator.cc: Integral_ln :128-150 */
static double ar_ln(double x, double y, double w)
{
    double r = sqrt(fabs(x * x + y * y + w * w));
    double r0 = sqrt(fabs(y * y + w * w));
    double xa = fabs(x);
    double c1 = (xa < 1.e-10) ? 0. : xa * log(fabs(y + r) + 1.e-12);
    double c2 = (fabs(y) < 1.e-12) ? 0. : y * log(fabs((xa + r) / r0) + 1.e-12);
    double c3 = (fabs(w) < 1.e-12) ? 0. : w * (atan(xa / w) + atan(y * w / (x * x + w * w + xa * r)) - atan(y / w));
    double ret = c1 + c2 - xa + c3;
    return (x < 0.) ? -ret : ret;
}

/* EFieldLocalXY :152-180 */
static double ar_exy(double x1, double x2, double y1, double y2, double z)
{
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

/* EFieldLocalZ :182-205 */
static double ar_ez(double x1, double x2, double y1, double y2, double z)
{
    double t1 = (fabs(y1) > 1.e-15 ? z / y1 : 1.e15);
    double t2 = (fabs(y2) > 1.e-15 ? z / y2 : 1.e15);
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

/* Potential :25-43 and ElectricField :45-121 */
void ko_rect_analytic(const double* d, const double P[3], double out[4])
{
    const double* N1 = d + 5;
    const double* N2 = d + 8;
    double N3[3];
    ko_normal(KO_RECTANGLE, d, N3);
    double p[3] = {P[0] - d[2], P[1] - d[3], P[2] - d[4]};
    double uP = dot3(p, N1), vP = dot3(p, N2), w = dot3(p, N3);
    double xmin = -uP, xmax = -uP + d[0], ymin = -vP, ymax = -vP + d[1];
    double I = (ar_ln(xmax, ymax, w) - ar_ln(xmin, ymax, w) - ar_ln(xmax, ymin, w) + ar_ln(xmin, ymin, w));
    out[3] = I / (4. * KO_PI * KO_EPS0);
    double prefac = 1. / (4. * KO_PI * KO_EPS0);
    double loc[3];
    loc[0] = prefac * ar_exy(xmin, xmax, ymin, ymax, w);
    loc[1] = prefac * ar_exy(ymin, ymax, xmin, xmax, w);
    double tmin = w / ymin, tmax = w / ymax;
    double sign_z = (w < 0) ? -1. : 1.;
    if (fabs(w) < 1.e-13)
        loc[2] = (xmin < 0. && xmax > 0. && ymin < 0. && ymax > 0.) ? 1.0 / (2. * KO_EPS0) : 0.;
    else if (SGN(tmin) != SGN(tmax))
        loc[2] = prefac * sign_z * fabs(ar_ez(xmin, xmax, 0, fabs(ymin), w) + ar_ez(xmin, xmax, 0, fabs(ymax), w));
    else
        loc[2] = prefac * sign_z * fabs(ar_ez(xmin, xmax, ymax, ymin, w));
    for (int i = 0; i < 3; i++) out[i] = (N1[i] * loc[0] + N2[i] * loc[1] + N3[i] * loc[2]);
}

/* ------------------------------------------------------------ reference policy */
/* Gauss-Legendre 32-point rule, positive nodes ascending (mirror nodes -x, same weights);
   KGaussLegendreQuadrature.cc (Math/src) applies them as sum_j w_j f(B + A x_j) followed by
   sum_j w_j f(B - A x_j), times A = (b - a)/2, B = (b + a)/2 */
static const double GL32x[16] = {
    0.0483076656877383162348, 0.144471961582796493485, 0.239287362252137074545, 0.33186860228212764978,
    0.421351276130635345364,  0.506899908932229390024, 0.587715757240762329041, 0.663044266930215200975,
    0.732182118740289680387,  0.794483795967942406963, 0.849367613732569970134, 0.896321155766052123965,
    0.934906075937739689171,  0.964762255587506430774, 0.9856115115452683354,   0.997263861849481563545};
static const double GL32w[16] = {
    0.0965400885147278005668, 0.0956387200792748594191, 0.0938443990808045656392, 0.0911738786957638847129,
    0.0876520930044038111428, 0.0833119242269467552222, 0.0781938957870703064717, 0.0723457941088485062254,
    0.0658222227763618468377, 0.0586840934785355471453, 0.0509980592623761761962, 0.0428358980222266806569,
    0.0342738629130214331027, 0.0253920653092620594558, 0.0162743947309056706052, 0.00701861000947009660041};

/* point charge kernel for the nested rules: (E, phi) of unit charge at Q, times 1/(4 pi eps0);
   triF (BiQuadratureTriangleIntegrator.cc:51-70) computes C = 1/R, C^3; rectF
   (BiQuadratureRectangleIntegrator.cc:33-56) 1/R and R^3 */
static void bq_point(int rect, const double Q[3], const double P[3], double v[4])
{
    double QP[3] = {P[0] - Q[0], P[1] - Q[1], P[2] - Q[2]};
    double R = norm3(QP);
    if (rect) {
        double R3 = R * R * R;
        for (int j = 0; j < 3; j++) v[j] = KO_K / R3 * QP[j];
        v[3] = KO_K / R;
    } else {
        double C = 1. / R, C3 = C * C * C;
        for (int j = 0; j < 3; j++) v[j] = KO_K * C3 * QP[j];
        v[3] = KO_K * C;
    }
}

/* KElectrostaticBiQuadratureTriangleIntegrator.cc:24-150: outer 32-node rule over the height
   y in [0, h] of vertex P1 above side P2-P0, inner rule along the chord parallel to P2->P0 */
void ko_tri_biquad(const double* d, const double P[3], double out[4])
{
    double A[3], B[3], C[3];
    for (int k = 0; k < 3; k++) {
        A[k] = d[2 + k];
        B[k] = d[2 + k] + d[5 + k] * d[0];
        C[k] = d[2 + k] + d[8 + k] * d[1];
    }
    double a[3], b[3], c[3];
    for (int k = 0; k < 3; k++) {
        a[k] = A[k] - C[k];
        b[k] = B[k] - C[k];
        c[k] = B[k] - A[k];
    }
    double la = norm3(a), lb = norm3(b);
    double ua[3];
    for (int k = 0; k < 3; k++) ua[k] = (1. / la) * a[k];
    double cosg = dot3(a, b) / (la * lb);
    double D[3], DB[3];
    for (int k = 0; k < 3; k++) {
        D[k] = C[k] + (lb * cosg) * ua[k];
        DB[k] = D[k] - B[k];
    }
    double h = norm3(DB);
    double so[4] = {0., 0., 0., 0.};
    double Ao = (h - 0.) / 2., Bo = (h + 0.) / 2.;
    for (int half = 0; half < 2; half++)
        for (int jo = 0; jo < 16; jo++) {
            double y = half == 0 ? Bo + Ao * GL32x[jo] : Bo - Ao * GL32x[jo];
            double E[3], F[3], EF[3];
            for (int k = 0; k < 3; k++) {
                E[k] = B[k] - ((y / h) * b[k]);
                F[k] = B[k] - ((y / h) * c[k]);
                EF[k] = E[k] - F[k];
            }
            double L = norm3(EF);
            double Ai = (L - 0.) / 2., Bi = (L + 0.) / 2.;
            double si[4] = {0., 0., 0., 0.};
            for (int hi = 0; hi < 2; hi++)
                for (int ji = 0; ji < 16; ji++) {
                    double x = hi == 0 ? Bi + Ai * GL32x[ji] : Bi - Ai * GL32x[ji];
                    double Q[3] = {E[0] + x * ua[0], E[1] + x * ua[1], E[2] + x * ua[2]};
                    double v[4];
                    bq_point(0, Q, P, v);
                    for (int k = 0; k < 4; k++) si[k] += GL32w[ji] * v[k];
                }
            for (int k = 0; k < 4; k++) so[k] += GL32w[jo] * (Ai * si[k]);
        }
    for (int k = 0; k < 4; k++) out[k] = Ao * so[k];
}

/* KElectrostaticBiQuadratureRectangleIntegrator.cc:18-104: outer 32-node rule over x in
   [0, A] along N1, inner over y in [0, B] along N2 */
void ko_rect_biquad(const double* d, const double P[3], double out[4])
{
    double so[4] = {0., 0., 0., 0.};
    double Ao = (d[0] - 0.) / 2., Bo = (d[0] + 0.) / 2.;
    double Ai = (d[1] - 0.) / 2., Bi = (d[1] + 0.) / 2.;
    for (int half = 0; half < 2; half++)
        for (int jo = 0; jo < 16; jo++) {
            double x = half == 0 ? Bo + Ao * GL32x[jo] : Bo - Ao * GL32x[jo];
            double si[4] = {0., 0., 0., 0.};
            for (int hi = 0; hi < 2; hi++)
                for (int ji = 0; ji < 16; ji++) {
                    double y = hi == 0 ? Bi + Ai * GL32x[ji] : Bi - Ai * GL32x[ji];
                    double Q[3];
                    for (int k = 0; k < 3; k++) Q[k] = d[2 + k] + x * d[5 + k] + y * d[8 + k];
                    double v[4];
                    bq_point(1, Q, P, v);
                    for (int k = 0; k < 4; k++) si[k] += GL32w[ji] * v[k];
                }
            for (int k = 0; k < 4; k++) so[k] += GL32w[jo] * (Ai * si[k]);
        }
    for (int k = 0; k < 4; k++) out[k] = Ao * so[k];
}

/* KElectrostatic256NodeQuadratureLineSegmentIntegrator.cc:192-270 (ElectricFieldAndPotential):
   the segment in 8 equal parts, the 32-node rule on each */
void ko_line_256(const double* d, const double P[3], double out[4])
{
    double L = sqrt(SQ(d[3] - d[0]) + SQ(d[4] - d[1]) + SQ(d[5] - d[2]));
    double pu = 1. / L;
    double u[3] = {pu * (d[3] - d[0]), pu * (d[4] - d[1]), pu * (d[5] - d[2])};
    double seg = L / 8, hs = 0.5 * seg;
    double prefac = 0.25 * d[6] / KO_EPS0;
    double sf[3] = {0., 0., 0.}, sp = 0.;
    for (int k = 0; k < 8; k++) {
        double ctr[3];
        for (int j = 0; j < 3; j++) {
            double s0 = d[j] + (k * seg * u[j]);
            double s1 = d[j] + ((k + 1) * seg * u[j]);
            ctr[j] = 0.5 * (s0 + s1);
        }
        for (int i = 0; i < 16; i++) {
            double wf = GL32w[i] * hs, nd = GL32x[i] * hs;
            double pd[3], ndir[3], pm = 0., nm = 0.;
            for (int j = 0; j < 3; j++) {
                pd[j] = P[j] - (ctr[j] + (u[j] * nd));
                pm += SQ(pd[j]);
                ndir[j] = P[j] - (ctr[j] - (u[j] * nd));
                nm += SQ(ndir[j]);
            }
            pm = 1. / sqrt(pm);
            nm = 1. / sqrt(nm);
            sp += wf * (pm + nm);
            pm = pm * pm * pm;
            nm = nm * nm * nm;
            for (int j = 0; j < 3; j++) sf[j] += wf * (pm * pd[j] + nm * ndir[j]);
        }
    }
    for (int j = 0; j < 3; j++) out[j] = prefac * sf[j];
    out[3] = prefac * sp;
}

/* ------------------------------------------------------ axisymmetric elements */
/* Complete elliptic integrals from Carlson's symmetric forms (B. C. Carlson, Numer. Algorithms
   10 (1995) 13-26, algorithms 1 and 4: duplication to a relative spread below TOL, then the
   fifth-order series):  K(k) = R_F(0, 1-k^2, 1),  E(k) = K - (k^2/3) R_D(0, 1-k^2, 1),
   (E - K)/k^2 = -R_D(0, 1-k^2, 1)/3.  This is an independent method: the reference evaluates
   Cody's Chebyshev fits (Math/src/KEllipticIntegrals.cc:14-99) and a power series
   (:101-125); the values agree to rounding except where the reference regularises the
   logarithm of K, log(eta + 1e-19) with eta = 1 - k^2 (:54), which ko_ell_K reproduces. */
static double carlson_rf(double x, double y, double z)
{
    const double a0 = (x + y + z) / 3.;
    double q = fmax(fabs(a0 - x), fmax(fabs(a0 - y), fabs(a0 - z))) * 400.; /* (3 * 1e-16)^(-1/6) ~ 389 */
    double a = a0, f = 1.;
    while (f * q >= fabs(a)) {
        const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        const double lam = sx * sy + sx * sz + sy * sz;
        x = .25 * (x + lam);
        y = .25 * (y + lam);
        z = .25 * (z + lam);
        a = .25 * (a + lam);
        f *= .25;
    }
    const double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y);
    (void)z;
    const double e2 = X * Y - Z * Z, e3 = X * Y * Z;
    return (1. - e2 / 10. + e3 / 14. + e2 * e2 / 24. - 3. * e2 * e3 / 44.) / sqrt(a);
}

static double carlson_rd(double x, double y, double z)
{
    const double a0 = (x + y + 3. * z) / 5.;
    double q = fmax(fabs(a0 - x), fmax(fabs(a0 - y), fabs(a0 - z))) * 600.; /* (1e-16 / 4)^(-1/6) ~ 541 */
    double a = a0, f = 1., s = 0.;
    while (f * q >= fabs(a)) {
        const double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        const double lam = sx * sy + sx * sz + sy * sz;
        s += f / (sz * (z + lam));
        x = .25 * (x + lam);
        y = .25 * (y + lam);
        z = .25 * (z + lam);
        a = .25 * (a + lam);
        f *= .25;
    }
    const double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y) / 3.;
    const double e2 = X * Y - 6. * Z * Z, e3 = (3. * X * Y - 8. * Z * Z) * Z, e4 = 3. * (X * Y - Z * Z) * Z * Z,
                 e5 = X * Y * Z * Z * Z;
    const double ser = 1. - 3. * e2 / 14. + e3 / 6. + 9. * e2 * e2 / 88. - 3. * e4 / 22. - 9. * e2 * e3 / 52. +
                       3. * e5 / 26.;
    return f * ser / (a * sqrt(a)) + 3. * s;
}

/* the reference's regularisation of K: its fit is ln4 + A(eta) - (1/2 + B(eta)) log(eta + 1e-19)
   where the exact K has log(eta); B(eta) = O(eta) makes the difference -log1p(1e-19/eta)/2 */
static double ell_k_reg(double eta) { return eta > 0. ? -.5 * log1p(1e-19 / eta) : 0.; }

double ko_ell_K(double k)
{
    const double eta = 1. - k * k; /* KEllipticIntegrals.cc:24 */
    if (!(eta > 0.)) return 1.386294361119890618834464242916 - .5 * log(fabs(eta) + 1e-19);
    return carlson_rf(0., eta, 1.) + ell_k_reg(eta);
}

double ko_ell_E(double k)
{
    const double eta = 1. - k * k;
    if (!(eta > 0.)) return 1.;
    return carlson_rf(0., eta, 1.) - (k * k / 3.) * carlson_rd(0., eta, 1.);
}

/* (E(k) - K(k))/k^2: the reference takes the difference above k^2 = 0.8 (and so inherits the
   regularised K there) and a power series below (KEllipticIntegrals.cc:101-125) */
double ko_ell_EK(double k)
{
    const double k2 = k * k, eta = 1. - k2;
    if (!(eta > 0.)) return (1. - ko_ell_K(k)) / k2;
    double ek = -carlson_rd(0., eta, 1.) / 3.;
    if (k2 > 0.8) ek -= ell_k_reg(eta) / k2;
    return ek;
}

/* KConicSection: P0, P1 (Surfaces/src/KConicSection.cc:7-11), radius = x, axial position = z
   of each point (Surfaces/include/KConicSection.hh GetR0/GetZ0); KRing: P (KRing.cc:8-11). */
double ko_axisym_area(int shape, const double* d)
{
    if (shape == KO_RING) return 2. * KO_PI * d[0]; /* KRing.cc:20-23 */
    return KO_PI * (d[0] + d[3]) * sqrt((d[0] - d[3]) * (d[0] - d[3]) + (d[2] - d[5]) * (d[2] - d[5]));
}

/* KConicSection::Normal (KConicSection.cc:59-104); KRing::Normal (KRing.cc:40-43) */
void ko_axisym_normal(int shape, const double* d, double n[3])
{
    n[1] = 0.;
    if (shape == KO_RING) {
        n[0] = 1.;
        n[2] = 0.;
        return;
    }
    const double r0 = d[0], z0 = d[2], r1 = d[3], z1 = d[5];
    if (fabs(z0 - z1) < 1.e-10) {
        n[0] = 0.;
        n[2] = r0 < r1 ? -1. : 1.;
    } else if (fabs(r0 - r1) < 1.e-10) {
        n[2] = 0.;
        n[0] = z0 < z1 ? 1. : -1.;
    } else {
        const double m = (z0 - z1) / (r1 - r0);
        n[0] = m / sqrt(1. + m * m);
        n[2] = 1. / sqrt(1. + m * m);
    }
}

/* Unscaled integrands of a charged ring at axial position Z, radius R seen from (z, r)
   (Analytic/src/KElectrostaticAnalyticRingIntegrator.cc:93-110 potential, :138-160 radial and
   :185-205 axial field); the ring runs along the generating line of a conic section,
   R and Z at distance p from its first point (par: z, r, zA, rA, zB, rB, L). */
static void ring_integrands(double p, const double par[7], double f[3])
{
    const double Z = par[2] + p / par[6] * (par[4] - par[2]);
    const double R = par[3] + p / par[6] * (par[5] - par[3]);
    const double dz = par[0] - Z, dr = par[1] - R, sumr = R + par[1];
    const double eta = (dr * dr + dz * dz) / (sumr * sumr + dz * dz);
    const double S = sqrt(sumr * sumr + dz * dz);
    const double kp = eta > 1. ? 0. : sqrt(1. - eta);
    f[0] = R * ko_ell_K(kp) / S;
    const double k = sqrt(1. - eta);
    const double E = ko_ell_E(k);
    f[1] = R / (S * S * S) * (-2. * R * ko_ell_EK(k) + dr / eta * E);
    f[2] = dz / (S * S * S) * E / eta * R;
}

/* Gluck's composite rule on n+1 equidistant nodes of [a, b] with end corrections
   (Math/src/KGaussianQuadrature.cc:18-88): the end weights are the reference's table.  They
   are not fixed by polynomial exactness (other end weights are also exact to degree 5 and 9
   for every n), and the conic-section results depend on the rule's error term, so the same
   weights are used. */
static const double GL5[6] = {51. / 160., 991. / 720., 59. / 90., 97. / 80., 1333. / 1440., 91. / 90.};
static const double GL9[10] = {0.2803440531305107e0, 0.1648702325837748e1, -0.2027449845679092e0,
                               0.2797927414021179e1, -0.9761199294532843e0, 0.2556499393738999e1,
                               0.1451083002645404e0, 0.1311227127425048e1, 0.9324249063051143e0,
                               0.1006631393298060e1};

static void gluck_quad(double a, double b, const double par[7], int n, double ans[3])
{
    if (n < 12) n = 12;
    const double h = (b - a) / n;
    const int m = n < 20 ? 6 : 10;
    const double* w = n < 20 ? GL5 : GL9;
    ans[0] = ans[1] = ans[2] = 0.;
    for (int i = 0; i <= n; i++) {
        const double x = a + h * i;
        const double wi = i < m ? w[i] : (i > n - m ? w[n - i] : 1.);
        double f[3];
        ring_integrands(x, par, f);
        for (int j = 0; j < 3; j++) ans[j] += wi * f[j];
    }
    for (int j = 0; j < 3; j++) ans[j] *= h;
}

/* field and potential of a uniformly charged ring (unit charge: the reference's prefactor
   1/(2 pi^2 eps0 R) times R K/S, KElectrostaticAnalyticRingIntegrator.cc:26-67) */
void ko_ring(const double* d, const double P[3], double out[4])
{
    const double R = d[0], Z = d[2];
    const double par[7] = {P[2], sqrt(P[0] * P[0] + P[1] * P[1]), Z, R, Z, R, 1.};
    const double c = 1. / (par[3] * 2. * KO_PI * KO_PI * KO_EPS0);
    double f[3];
    ring_integrands(0., par, f);
    out[3] = c * f[0];
    out[2] = c * f[2];
    const double er = c * f[1];
    if (par[1] < 1.e-14) {
        out[0] = out[1] = 0.;
    } else {
        out[0] = P[0] / par[1] * er;
        out[1] = P[1] / par[1] * er;
    }
}

/* field and potential of a conic section of unit surface charge density: the integral of
   ring integrands over the generating line, 20 / 100 / 500 nodes by the distance measure
   |Da + Db - L|/L, and near the line a geometric split towards the projected point with the
   logarithmic remainder added in closed form (Analytic/src/
   KElectrostaticAnalyticConicSectionIntegrator.cc:28-104 potential, :107-221 field) */
void ko_conic(const double* d, const double P[3], double out[4])
{
    static const double ln4 = 1.386294361119890;
    const double z = P[2], r = sqrt(P[0] * P[0] + P[1] * P[1]);
    const double za = d[2], ra = d[0], zb = d[5], rb = d[3];
    const double L = sqrt((d[0] - d[3]) * (d[0] - d[3]) + (d[1] - d[4]) * (d[1] - d[4]) + (d[2] - d[5]) * (d[2] - d[5]));
    const double par[7] = {z, r, za, ra, zb, rb, L};
    const double Da = sqrt((za - z) * (za - z) + (ra - r) * (ra - r));
    const double Db = sqrt((zb - z) * (zb - z) + (rb - r) * (rb - r));
    const double D = fabs(Da + Db - L) / L;
    double q[3];
    if (D >= 5.e-2)
        gluck_quad(0., L, par, 20, q);
    else if (D >= 5.e-3)
        gluck_quad(0., L, par, 100, q);
    else if (D >= 5.e-4)
        gluck_quad(0., L, par, 500, q);
    else {
        const double u0 = (zb - za) / L, u1 = (rb - ra) / L;
        const double pp = (z - za) * u0 + (r - ra) * u1;
        q[0] = q[1] = q[2] = 0.;
        if (pp < L) {
            const double pmax = L, pmin = pp < 0. ? 0. : pp;
            double b = pmax, a = pmin + (b - pmin) * 0.3, t[3];
            for (int it = 1; it <= 35; it++) {
                gluck_quad(a, b, par, 50, t);
                for (int j = 0; j < 3; j++) q[j] += t[j];
                if (fabs(a - pmin) / L < 1.e-8) break;
                b = a;
                a = pmin + (b - pmin) * 0.3;
            }
            const double h = fabs(a - pmin);
            const double tail = h * (ln4 + log(2. * par[1] + 1.e-12) + 1. - log(h)) / 2.;
            for (int j = 0; j < 3; j++) q[j] += tail;
        }
        if (pp > 0.) {
            const double pmax = pp > L ? L : pp, pmin = 0.;
            double b = pmin, a = pmax - (pmax - b) * 0.3, t[3];
            for (int it = 1; it <= 35; it++) {
                gluck_quad(b, a, par, 50, t);
                for (int j = 0; j < 3; j++) q[j] += t[j];
                if (fabs(pmax - a) / L < 1.e-8) break;
                b = a;
                a = pmax - (pmax - b) * 0.3;
            }
            const double h = fabs(pmax - a);
            const double tail = h * (ln4 + log(2. * par[1] + 1.e-12) + 1. - log(h)) / 2.;
            for (int j = 0; j < 3; j++) q[j] += tail;
        }
    }
    const double c = 1. / (KO_PI * KO_EPS0);
    out[3] = c * q[0];
    const double er = c * q[1];
    out[2] = c * q[2];
    if (r < 1.e-14) {
        out[0] = out[1] = 0.;
    } else {
        out[0] = P[0] / r * er;
        out[1] = P[1] / r * er;
    }
}

/* ------------------------------------------------------------ policies */
/* KElectrostaticBoundaryIntegratorFactory.cc:29-88: numeric = cubature/RWG dispatch (default),
   rwg = closed form everywhere (line: analytic), analytic = the Analytic integrators,
   reference = nested 32-node Gauss-Legendre (line: 256 nodes). */
int ko_element_eval(int policy, int shape, const double* d, const double P[3], double out[4])
{
    /* every policy takes the analytic conic-section and ring integrators (...Factory.cc:39-73) */
    if (shape == KO_CONICSECTION) {
        ko_conic(d, P, out);
        return 0;
    }
    if (shape == KO_RING) {
        ko_ring(d, P, out);
        return 0;
    }
    if (policy == KO_ANALYTIC) {
        if (shape == KO_TRIANGLE) ko_tri_analytic(d, P, out);
        else if (shape == KO_RECTANGLE) ko_rect_analytic(d, P, out);
        else ko_line_analytic(d, P, out);
        return 0;
    }
    if (policy == KO_REFERENCE) {
        if (shape == KO_TRIANGLE) ko_tri_biquad(d, P, out);
        else if (shape == KO_RECTANGLE) ko_rect_biquad(d, P, out);
        else ko_line_256(d, P, out);
        return 0;
    }
    if (shape == KO_TRIANGLE) {
        if (policy == KO_RWG) {
            ko_tri_rwg(d, P, out);
            return 0;
        }
        return ko_tri_numeric(d, P, out);
    }
    if (shape == KO_RECTANGLE) {
        if (policy == KO_RWG) {
            ko_rect_rwg(d, P, out);
            return 0;
        }
        return ko_rect_numeric(d, P, out);
    }
    if (policy == KO_RWG) {
        ko_line_analytic(d, P, out);
        return 0;
    }
    return ko_line_numeric(d, P, out);
}

/* ------------------------------------------------------------ collocation */
/* KElectrostaticBoundaryIntegrator::ComputeBoundaryIntegral (Core/include/...hh:260-280) with
   the boundary visitor prefactor (1+flux)/(1-flux) for Neumann targets (Core/src/...cc:49-62) */
double ko_matrix_element(int policy, int64_t i, int64_t j, const int32_t* shape, const double* recs,
                         const int32_t* bc, const double* flux)
{
    const double* di = recs + KO_REC * i;
    const double* dj = recs + KO_REC * j;
    double ci[3], out[4];
    ko_centroid(shape[i], di, ci);
    if (bc == NULL || bc[i] == KO_DIRICHLET) {
        ko_element_eval(policy, shape[j], dj, ci, out);
        return out[3];
    }
    double cj[3];
    ko_centroid(shape[j], dj, cj);
    double dd[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
    if (norm3(dd) >= 1.e-12) {
        double n[3];
        ko_normal(shape[i], di, n);
        ko_element_eval(policy, shape[j], dj, ci, out);
        return out[0] * n[0] + out[1] * n[1] + out[2] * n[2];
    }
    double f = flux ? flux[i] : 0.;
    return ((1. + f) / (1. - f)) / (2. * KO_EPS0);
}

void ko_assemble(int policy, int64_t n, const int32_t* shape, const double* recs, const int32_t* bc,
                 const double* flux, double* A)
{
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) A[i * n + j] = ko_matrix_element(policy, i, j, shape, recs, bc, flux);
}

void ko_rhs(int64_t n, const int32_t* bc, const double* value, double* b)
{
    for (int64_t i = 0; i < n; i++) b[i] = (bc == NULL || bc[i] == KO_DIRICHLET) ? value[i] : 0.;
}

/* ------------------------------------------------------------ field */
/* KIntegratingFieldSolver<...,ElectrostaticSingleThread>::ElectricFieldAndPotential
   (KElectrostaticIntegratingFieldSolver.hh:303-339): Kahan sums in container order */
static void field_point(int policy, int64_t n, const int32_t* shape, const double* recs, const double* sigma,
                        const double* P, double* phi, double* E)
{
    double s[4] = {0., 0., 0., 0.}, c[4] = {0., 0., 0., 0.};
    for (int64_t j = 0; j < n; j++) {
        double out[4];
        ko_element_eval(policy, shape[j], recs + KO_REC * j, P, out);
        for (int k = 0; k < 4; k++) {
            double y = out[k] * sigma[j] - c[k];
            double t = s[k] + y;
            c[k] = (t - s[k]) - y;
            s[k] = t;
        }
    }
    if (phi) *phi = s[3];
    if (E)
        for (int k = 0; k < 3; k++) E[k] = s[k];
}

void ko_field(int policy, int64_t n, const int32_t* shape, const double* recs, const double* sigma, int64_t m,
              const double* P, double* phi, double* E)
{
    for (int64_t p = 0; p < m; p++) field_point(policy, n, shape, recs, sigma, P + 3 * p, phi ? phi + p : 0, E ? E + 3 * p : 0);
}

void ko_field_mt(int nthreads, int policy, int64_t n, const int32_t* shape, const double* recs, const double* sigma,
                 int64_t m, const double* P, double* phi, double* E)
{
    if (m <= 0) return;
    /* the first point serially: initialises the lazily built rule tables before the threads */
    field_point(policy, n, shape, recs, sigma, P, phi ? phi : 0, E ? E : 0);
#pragma omp parallel for num_threads(nthreads) schedule(dynamic, 8)
    for (int64_t p = 1; p < m; p++)
        field_point(policy, n, shape, recs, sigma, P + 3 * p, phi ? phi + p : 0, E ? E + 3 * p : 0);
}

/* ------------------------------------------------------ symmetry groups */
/* A group's integral: its members in order, plainly summed (Core/include/
   KElectrostaticElementIntegrator.hh:52-79); members [gs[g], gs[g+1]). */
static void group_eval(int policy, const int64_t* gs, int64_t g, const int32_t* shape, const double* recs,
                       const double P[3], double out[4])
{
    out[0] = out[1] = out[2] = out[3] = 0.;
    for (int64_t m = gs[g]; m < gs[g + 1]; m++) {
        double v[4];
        ko_element_eval(policy, shape[m], recs + KO_REC * m, P, v);
        for (int k = 0; k < 4; k++) out[k] += v[k];
    }
}

/* collocation with groups as sources and targets: the target is the group's first member
   (KSymmetryGroup::Centroid / Normal, Surfaces/include/KSymmetryGroup.hh:48-59), the Neumann
   self term compares the groups' (first members') centroids as for single elements
   (Core/include/KElectrostaticBoundaryIntegrator.hh:260-280) */
void ko_assemble_groups(int policy, int64_t n, const int64_t* gs, const int32_t* shape, const double* recs,
                        const int32_t* bc, const double* flux, double* A)
{
    for (int64_t i = 0; i < n; i++) {
        const int64_t fi = gs[i];
        double ci[3], ni[3];
        ko_centroid(shape[fi], recs + KO_REC * fi, ci);
        ko_normal(shape[fi], recs + KO_REC * fi, ni);
        for (int64_t j = 0; j < n; j++) {
            double out[4];
            if (bc == NULL || bc[i] == KO_DIRICHLET) {
                group_eval(policy, gs, j, shape, recs, ci, out);
                A[i * n + j] = out[3];
                continue;
            }
            double cj[3];
            ko_centroid(shape[gs[j]], recs + KO_REC * gs[j], cj);
            double dd[3] = {cj[0] - ci[0], cj[1] - ci[1], cj[2] - ci[2]};
            if (norm3(dd) >= 1.e-12) {
                group_eval(policy, gs, j, shape, recs, ci, out);
                A[i * n + j] = out[0] * ni[0] + out[1] * ni[1] + out[2] * ni[2];
            } else {
                const double f = flux ? flux[i] : 0.;
                A[i * n + j] = ((1. + f) / (1. - f)) / (2. * KO_EPS0);
            }
        }
    }
}

/* Kahan-summed field over groups, sigma per group (KElectrostaticIntegratingFieldSolver.hh:262-339
   with the group integral above) */
void ko_field_groups(int policy, int64_t n, const int64_t* gs, const int32_t* shape, const double* recs,
                     const double* sigma, int64_t m, const double* P, double* phi, double* E)
{
    for (int64_t p = 0; p < m; p++) {
        double s[4] = {0., 0., 0., 0.}, c[4] = {0., 0., 0., 0.};
        for (int64_t j = 0; j < n; j++) {
            double out[4];
            group_eval(policy, gs, j, shape, recs, P + 3 * p, out);
            for (int k = 0; k < 4; k++) {
                double y = out[k] * sigma[j] - c[k];
                double t = s[k] + y;
                c[k] = (t - s[k]) - y;
                s[k] = t;
            }
        }
        if (phi) phi[p] = s[3];
        if (E)
            for (int k = 0; k < 3; k++) E[3 * p + k] = s[k];
    }
}

/* ------------------------------------------------------------ solvers */
static double dotn(int64_t n, const double* a, const double* b)
{
    double r = 0.;
    for (int64_t i = 0; i < n; i++) r += a[i] * b[i];
    return r;
}

static void matvec(int64_t n, const double* A, const double* x, double* y)
{
    for (int64_t i = 0; i < n; i++) {
        double s = 0.;
        for (int64_t j = 0; j < n; j++) s += A[i * n + j] * x[j];
        y[i] = s;
    }
}

/* GMRES: KGeneralizedMinimalResidual.hh:178-337 driven by KSimpleIterativeKrylovSolver::Solve
   (:140-194).  Convergence: |p_{j+1}| / ||b|| <= tol.  The small least-squares system H y = p is
   upper triangular after the Givens rotations; it is solved by back substitution (the
   reference uses an SVD, which gives the same y to rounding for a non-singular H).
   Restart: x <- x0 + V y (the reference's UpdateSolution writes V y alone). */
int ko_gmres(int64_t n, const double* A, const double* b, double* x, double tol, int maxit, int restart,
             double* resid, double* hist)
{
    return ko_gmres_ortho(n, A, b, x, tol, maxit, restart, 0, resid, hist);
}

/* ortho = 0: modified Gram-Schmidt in the reference's order (KGeneralizedMinimalResidual.hh:225-236);
   ortho = 1: classical Gram-Schmidt with one reorthogonalisation pass (CGS2), the scheme of the
   HIP path. */
int ko_gmres_ortho(int64_t n, const double* A, const double* b, double* x, double tol, int maxit, int restart,
                   int ortho, double* resid, double* hist)
{
    int kmax = (restart > 0 && restart < maxit) ? restart : maxit;
    double* V = (double*)malloc(sizeof(double) * (size_t)n * (size_t)(kmax + 1));
    double* H = (double*)calloc((size_t)(kmax + 1) * (size_t)(kmax + 1), sizeof(double));
    double* cs = (double*)malloc(sizeof(double) * (size_t)(kmax + 1));
    double* sn = (double*)malloc(sizeof(double) * (size_t)(kmax + 1));
    double* pv = (double*)malloc(sizeof(double) * (size_t)(kmax + 2));
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* y = (double*)malloc(sizeof(double) * (size_t)(kmax + 1));
    double bnorm = sqrt(dotn(n, b, b));
    int it = 0;
    double rn = 0.;
    int done = 0;
    while (!done) {
        /* ResetAndInitialize (:194-212) */
        matvec(n, A, x, w);
        for (int64_t i = 0; i < n; i++) w[i] = b[i] - w[i];
        double nr = sqrt(dotn(n, w, w));
        double inr = 1.0 / nr;
        for (int64_t i = 0; i < n; i++) V[i] = w[i] * inr;
        pv[0] = nr;
        int j = 0;
        for (;;) {
            /* AugmentKrylovSubspace (:214-271): MGS then Givens */
            double* vj1 = V + (size_t)(j + 1) * n;
            matvec(n, A, V + (size_t)j * n, w);
            double* h = H + (size_t)j * (kmax + 1);
            if (ortho == 0) {
                for (int r = 0; r <= j; r++) {
                    double hh = dotn(n, w, V + (size_t)r * n);
                    h[r] = hh;
                    const double* vr = V + (size_t)r * n;
                    for (int64_t i = 0; i < n; i++) w[i] -= hh * vr[i];
                }
            } else {
                for (int r = 0; r <= j; r++) h[r] = 0.;
                for (int pass = 0; pass < 2; pass++) {
                    double hp[kmax + 1];
                    for (int r = 0; r <= j; r++) hp[r] = dotn(n, w, V + (size_t)r * n);
                    for (int64_t i = 0; i < n; i++) {
                        double s = 0.;
                        for (int r = 0; r <= j; r++) s += V[(size_t)r * n + i] * hp[r];
                        w[i] -= s;
                    }
                    for (int r = 0; r <= j; r++) h[r] += hp[r];
                }
            }
            double beta = sqrt(dotn(n, w, w));
            double ib = 1.0 / beta;
            for (int64_t i = 0; i < n; i++) vj1[i] = w[i] * ib;
            for (int r = 0; r < j; r++) {
                double a = h[r], bb = h[r + 1];
                h[r] = cs[r] * a + sn[r] * bb;
                h[r + 1] = -1.0 * sn[r] * a + cs[r] * bb;
            }
            double g = sqrt(h[j] * h[j] + beta * beta);
            cs[j] = h[j] / g;
            sn[j] = beta / g;
            h[j] = g;
            pv[j + 1] = -1.0 * sn[j] * pv[j];
            pv[j] = cs[j] * pv[j];
            rn = fabs(pv[j + 1]);
            j++;
            it++;
            if (hist) hist[it - 1] = rn / bnorm;
            int conv = (rn / bnorm) <= tol || it >= maxit;
            if (conv || j >= kmax) {
                /* UpdateSolution (:273-337): solve the triangular system */
                for (int r = j - 1; r >= 0; r--) {
                    double s = pv[r];
                    for (int q = r + 1; q < j; q++) s -= H[(size_t)q * (kmax + 1) + r] * y[q];
                    y[r] = s / H[(size_t)r * (kmax + 1) + r];
                }
                for (int r = 0; r < j; r++) {
                    const double* vr = V + (size_t)r * n;
                    for (int64_t i = 0; i < n; i++) x[i] += y[r] * vr[i];
                }
                if (conv) done = 1;
                break;
            }
        }
    }
    if (resid) *resid = rn / bnorm;
    free(V);
    free(H);
    free(cs);
    free(sn);
    free(pv);
    free(w);
    free(y);
    return it;
}

/* Robin Hood: KRobinHood.hh:62-104 with KRobinHood_SingleThread.hh:53-120.
   Every check_interval sub-iterations ||r||_inf/||b||_inf is compared with tol. */
int64_t ko_robin_hood(int64_t n, const double* A, const double* b, double* x, double tol, int64_t check_interval,
                      int64_t max_sub, double* resid)
{
    if (check_interval <= 0) check_interval = n;
    double* bi = (double*)calloc((size_t)n, sizeof(double));
    double* r = (double*)calloc((size_t)n, sizeof(double));
    double xinf = 0., binf = 0.;
    for (int64_t i = 0; i < n; i++) {
        if (fabs(x[i]) > xinf) xinf = fabs(x[i]);
        if (fabs(b[i]) > binf) binf = fabs(b[i]);
    }
    if (xinf > 1.e-16) matvec(n, A, x, bi);
    double rnorm = 1.; /* KIterativeSolver: fResidualNorm(1.) */
    int64_t sub = 0, total = 0;
    while (rnorm > tol && total < max_sub) {
        sub++;
        total++;
        for (int64_t i = 0; i < n; i++) r[i] = b[i] - bi[i];
        if (sub % check_interval == 0) {
            sub = 0;
            double ri = 0.;
            for (int64_t i = 0; i < n; i++)
                if (fabs(r[i]) > ri) ri = fabs(r[i]);
            rnorm = ri / binf;
        }
        int64_t im = 0;
        double mx = fabs(r[0]);
        for (int64_t i = 0; i < n; i++)
            if (fabs(r[i]) > mx) {
                mx = fabs(r[i]);
                im = i;
            }
        double corr = (b[im] - bi[im]) / A[im * n + im];
        x[im] += corr;
        for (int64_t i = 0; i < n; i++) bi[i] += A[i * n + im] * corr;
    }
    if (resid) *resid = rnorm;
    free(bi);
    free(r);
    return total;
}

/* KGaussianElimination::Solve (LinearAlgebra/Solvers/include/KGaussianElimination.hh:24-58):
   Gauss-Jordan elimination on the augmented matrix without pivoting, x_i = a_in / a_ii */
int ko_gauss(int64_t n, const double* A, const double* b, double* x)
{
    int64_t w = n + 1;
    double* M = (double*)malloc(sizeof(double) * (size_t)n * (size_t)w);
    for (int64_t i = 0; i < n; i++) {
        for (int64_t j = 0; j < n; j++) M[i * w + j] = A[i * n + j];
        M[i * w + n] = b[i];
    }
    for (int64_t col = 0; col < n; col++) {
        for (int64_t row = 0; row < n; row++) {
            if (row == col) continue;
            double f = M[row * w + col] / M[col * w + col];
            for (int64_t k = 0; k < w; k++) M[row * w + k] = M[row * w + k] - f * M[col * w + k];
        }
    }
    for (int64_t i = 0; i < n; i++) x[i] = M[i * w + n] / M[i * w + i];
    free(M);
    return 0;
}
