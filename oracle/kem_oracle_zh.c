/*
 * kem_oracle_zh.c -- CPU FP64 restatement of KEMField's electrostatic zonal-harmonic field
 * solver for axisymmetric elements (rings and conic sections sharing the z axis).
 *
 * TEST INFRASTRUCTURE ONLY: the checker of the HIP product path; only tests/ load it.
 *
 * Restated from (paths relative to KEMField/Source/FieldSolvers/ZonalHarmonic):
 *   Generator/src/KZHLegendreCoefficients.cc:20-60       recursion coefficients, rows 0-3
 *   Generator/src/KZHCoefficientGeneratorRing.cc:18-88     ring coefficients and rho
 *   Generator/src/KZHCoefficientGeneratorConicSection.cc:37-180 conic-section coefficients
 *                                                           (31-node modified Newton-Cotes)
 *   Generator/include/KZonalHarmonicCoefficientGenerator.hh:147-512 bifurcation, source points
 *   Generator/include/KZonalHarmonicContainer.hh:246-345    container tree, which levels
 *                                                           carry source points
 *   Solver/include/KZonalHarmonicComputer.hh:84-230          choice of source point (first
 *                                                           call: the whole list is searched)
 *   Solver/src/KElectrostaticZonalHarmonicFieldSolver.cc:98-124, 356-495, 719-857
 *                                                           ElectricFieldAndPotential: central,
 *                                                           remote, subsets, direct fallback
 * The direct fallback is the integrating field solver over the node's elements (Kahan sums in
 * container order, as ko_field).  Non-axisymmetric elements go to a node of their own without
 * source points (GroupCoaxialElements), which always integrates directly.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "kem_oracle.h"

#define ZH_PI 3.141592653589793238462643383279502884197169
#define ZH_EPS0 8.8541878128e-12

/* parameters (KZonalHarmonicParameters.hh:11-27 defaults) */
typedef struct {
    int nbif;
    double conv_ratio, prox, conv_param;
    int ncen;
    int frac_spacing;
    double frac_dist, cen_dz, cen_z1, cen_z2;
    int nrem, nrem_sp;
    double rem_z1, rem_z2;
} zh_par;

typedef struct {
    double z0, rho;
    int n;
    double* c;
} zh_sp;

typedef struct zh_node {
    int64_t* el;
    int64_t nel;
    int axisym; /* 0: a node of non-axisymmetric elements (never has source points) */
    zh_sp* cen;
    int ncen;
    zh_sp* rem;
    int nrem;
    struct zh_node* sub;
    int nsub;
} zh_node;

typedef struct {
    int64_t n;
    const int32_t* shape;
    const double* recs;
    const double* sigma;
    zh_par par;
    double* leg[4];
    int nleg;
    zh_node head;
    int64_t counts[3]; /* central, remote, direct evaluations */
} zh_tree;

static int is_axi(int s) { return s == KO_CONICSECTION || s == KO_RING; }

/* rows 0-3 of the recursion coefficients (KZHLegendreCoefficients.cc:38-44) */
static void legendre_rows(zh_tree* t, int nmax)
{
    t->nleg = nmax < 2 ? 2 : nmax;
    for (int r = 0; r < 4; r++) t->leg[r] = calloc((size_t)t->nleg, sizeof(double));
    for (int n = 2; n < t->nleg; n++) {
        t->leg[0][n] = (2. * n - 1.) / (1. * n);
        t->leg[1][n] = (n - 1.) / (1. * n);
        t->leg[2][n] = (2. * n - 1.) / (1. * (n - 1.));
        t->leg[3][n] = (1. * n) / (1. * (n - 1.));
    }
}

/* ---- per-element generators: z extent, rho and coefficients */
static void el_extrema(const zh_tree* t, int64_t e, double* z1, double* z2)
{
    const double* d = t->recs + KO_REC * e;
    if (t->shape[e] == KO_RING) {
        *z1 = *z2 = d[2];
    } else {
        *z1 = d[2] < d[5] ? d[2] : d[5];
        *z2 = d[2] > d[5] ? d[2] : d[5];
    }
}

static double el_rho(const zh_tree* t, int64_t e, double z0, int cen)
{
    const double* d = t->recs + KO_REC * e;
    if (t->shape[e] == KO_RING) {
        /* KZHCoefficientGeneratorRing.cc:74-88 (null distance 0) */
        const double rr = sqrt((d[2] - z0) * (d[2] - z0) + d[0] * d[0]);
        return rr < 0. ? 0. : rr;
    }
    /* KZHCoefficientGeneratorConicSection.cc:133-173 */
    const double Z = (d[2] + d[5]) * .5 - z0, R = (d[0] + d[3]) * .5;
    const double L = sqrt((d[2] - d[5]) * (d[2] - d[5]) + (d[0] - d[3]) * (d[0] - d[3]));
    const double c = (d[5] - d[2]) / L, s = (d[3] - d[0]) / L;
    const double x = -Z * c - R * s, xm = L / 2.;
    const double ra = sqrt((Z - xm * c) * (Z - xm * c) + (R - xm * s) * (R - xm * s));
    const double rb = sqrt((Z + xm * c) * (Z + xm * c) + (R + xm * s) * (R + xm * s));
    double rmin, rmax;
    if (x >= xm) {
        rmin = rb;
        rmax = ra;
    } else if (x <= -xm) {
        rmin = ra;
        rmax = rb;
    } else {
        rmin = sqrt((Z + x * c) * (Z + x * c) + (R + x * s) * (R + x * s));
        rmax = ra > rb ? ra : rb;
    }
    return cen ? rmin : rmax;
}

static void el_coeffs(const zh_tree* t, int64_t e, double z0, double rho, int nc, double* coeff, int cen)
{
    const double* d = t->recs + KO_REC * e;
    const double sig = t->sigma[e];
    double* P1 = calloc((size_t)nc, sizeof(double));
    if (t->shape[e] == KO_RING) {
        /* KZHCoefficientGeneratorRing.cc:18-72 */
        const double rr = sqrt((z0 - d[2]) * (z0 - d[2]) + d[0] * d[0]);
        const double pref = sig / (4. * ZH_PI * ZH_EPS0 * rr);
        const double ratio = cen ? rho / rr : rr / rho;
        P1[0] = 1.;
        P1[1] = (d[2] - z0) / rr;
        double ph = cen ? pref : pref * ratio;
        coeff[0] += ph;
        ph *= ratio;
        coeff[1] += ph * P1[1];
        for (int i = 2; i < nc; i++) {
            ph *= ratio;
            P1[i] = t->leg[0][i] * P1[1] * P1[i - 1] - t->leg[1][i] * P1[i - 2];
            coeff[i] += ph * P1[i];
        }
        free(P1);
        return;
    }
    /* KZHCoefficientGeneratorConicSection.cc:37-130: 31-node rule along the generating line
       with the modified Newton-Cotes end weights */
    static const double w9[10] = {0.2803440531305107e0, 0.1648702325837748e1, -0.2027449845679092e0,
                                  0.2797927414021179e1, -0.9761199294532843e0, 0.2556499393738999e1,
                                  0.1451083002645404e0, 0.1311227127425048e1, 0.9324249063051143e0,
                                  0.1006631393298060e1};
    const int M = 30;
    double w[31];
    for (int m = 0; m <= 9; m++) w[m] = w9[m];
    for (int m = 10; m <= M - 10; m++) w[m] = 1.;
    for (int m = M - 9; m <= M; m++) w[m] = w9[M - m];
    const double z0c = d[2], z1c = d[5], r0c = d[0], r1c = d[3];
    const double L = sqrt((z0c - z1c) * (z0c - z1c) + (r0c - r1c) * (r0c - r1c));
    const double del = L / M;
    for (int m = 0; m <= M; m++) {
        const double x = -L / 2. + del * m;
        const double Z = (z0c + z1c) / 2. + x * (z1c - z0c) / L;
        const double R = (r0c + r1c) / 2. + x * (r1c - r0c) / L;
        double ro0 = sqrt(R * R + (Z - z0) * (Z - z0));
        if (ro0 == 0) ro0 = 1.e-17;
        const double Rper = R / ro0, u0 = (Z - z0) / ro0;
        P1[0] = 1.;
        P1[1] = u0;
        double rc, rcn;
        if (cen)
            rc = rcn = rho / ro0;
        else
            rc = rcn = ro0 / rho;
        const double c = sig / (2. * ZH_EPS0) * w[m] * del * Rper;
        if (cen) {
            coeff[0] += c * P1[0];
            coeff[1] += c * rcn * P1[1];
        } else {
            coeff[0] += c * rcn * P1[0];
            rcn *= rc;
            coeff[1] += c * rcn * P1[1];
        }
        for (int n = 2; n < nc; n++) {
            rcn *= rc;
            P1[n] = t->leg[0][n] * u0 * P1[n - 1] - t->leg[1][n] * P1[n - 2];
            coeff[n] += c * rcn * P1[n];
            if (rcn < 1.e-17) break;
        }
    }
    free(P1);
}

/* ---- container-level generation (KZonalHarmonicCoefficientGenerator.hh) */
static void node_extrema(const zh_tree* t, const zh_node* nd, double* z1, double* z2)
{
    el_extrema(t, nd->el[0], z1, z2);
    for (int64_t i = 1; i < nd->nel; i++) {
        double a, b;
        el_extrema(t, nd->el[i], &a, &b);
        if (a < *z1) *z1 = a;
        if (b > *z2) *z2 = b;
    }
}

static double node_rho(const zh_tree* t, const zh_node* nd, double z, int cen)
{
    double rho = cen ? 1.e10 : 0.;
    for (int64_t i = 0; i < nd->nel; i++) {
        const double v = el_rho(t, nd->el[i], z, cen);
        if (cen ? v < rho : v > rho) rho = v;
    }
    return rho;
}

static zh_sp make_sp(const zh_tree* t, const zh_node* nd, double z, int nc, int cen)
{
    zh_sp sp;
    sp.z0 = z;
    sp.rho = node_rho(t, nd, z, cen);
    sp.n = nc;
    sp.c = calloc((size_t)nc, sizeof(double));
    for (int64_t i = 0; i < nd->nel; i++) el_coeffs(t, nd->el[i], z, sp.rho, nc, sp.c, cen);
    return sp;
}

static void push_sp(zh_sp** v, int* n, int* cap, zh_sp sp)
{
    if (*n == *cap) {
        *cap = *cap ? 2 * *cap : 8;
        *v = realloc(*v, sizeof(zh_sp) * (size_t)*cap);
    }
    (*v)[(*n)++] = sp;
}

static void node_source_points(const zh_tree* t, zh_node* nd, int head)
{
    const zh_par* p = &t->par;
    double z1 = head ? p->cen_z1 : 0., z2 = head ? p->cen_z2 : 0.;
    int cap = 0;
    if (z2 < z1) {
        const double tmp = z1;
        z1 = z2;
        z2 = tmp;
    }
    if (fabs(z1 - z2) < 1.e-10) node_extrema(t, nd, &z1, &z2);
    if (p->frac_spacing) {
        /* GenerateCentralSourcePointsByFractionalDistance (.hh:263-313) */
        push_sp(&nd->cen, &nd->ncen, &cap, make_sp(t, nd, z1, p->ncen, 1));
        double z = z1;
        for (;;) {
            double dZ = p->frac_dist * node_rho(t, nd, z, 1);
            if (dZ < p->cen_dz) dZ = p->cen_dz;
            z += dZ;
            if (z >= z2) break;
            push_sp(&nd->cen, &nd->ncen, &cap, make_sp(t, nd, z, p->ncen, 1));
        }
        push_sp(&nd->cen, &nd->ncen, &cap, make_sp(t, nd, z2, p->ncen, 1));
    } else {
        /* GenerateCentralSourcePointsByFixedDistance (.hh:221-261) */
        unsigned nsp = (unsigned)((z2 - z1) / p->cen_dz);
        if (nsp == 0) nsp = 1;
        for (unsigned i = 0; i < nsp; i++)
            push_sp(&nd->cen, &nd->ncen, &cap, make_sp(t, nd, z1 + p->cen_dz * i, p->ncen, 1));
    }
    /* GenerateRemoteSourcePoints (.hh:315-360) */
    z1 = head ? p->rem_z1 : 0.;
    z2 = head ? p->rem_z2 : 0.;
    if (z2 < z1) {
        const double tmp = z1;
        z1 = z2;
        z2 = tmp;
    }
    if (fabs(z1 - z2) < 1.e-10) node_extrema(t, nd, &z1, &z2);
    unsigned nsp = (unsigned)p->nrem_sp;
    double dz = (z2 - z1) / (nsp - 1);
    if (fabs(z1 - z2) < 1.e-10) {
        nsp = 1;
        dz = 0.;
    }
    cap = 0;
    for (unsigned i = 0; i < nsp; i++) push_sp(&nd->rem, &nd->nrem, &cap, make_sp(t, nd, z1 + dz * i, p->nrem, 0));
}

static void node_init(zh_node* nd, int64_t cap)
{
    memset(nd, 0, sizeof(*nd));
    nd->el = malloc(sizeof(int64_t) * (size_t)(cap > 0 ? cap : 1));
    nd->axisym = 1;
}

/* ConstructSubContainers (KZonalHarmonicContainer.hh:246-282): level -1 groups coaxial
   elements (all axisymmetric elements here share the z axis; the others form a group of their
   own), then every level below NBifurcations splits a node at the middle of its source-point
   extrema by element centre */
static void construct(zh_tree* t, zh_node* nd, int level)
{
    if (nd->nel == 0) return;
    if (level == -1) {
        int64_t nax = 0;
        for (int64_t i = 0; i < nd->nel; i++) nax += is_axi(t->shape[nd->el[i]]);
        if (nax > 0 && nax < nd->nel) {
            nd->nsub = 2;
            nd->sub = calloc(2, sizeof(zh_node));
            node_init(&nd->sub[0], nax);
            node_init(&nd->sub[1], nd->nel - nax);
            for (int64_t i = 0; i < nd->nel; i++) {
                zh_node* s = is_axi(t->shape[nd->el[i]]) ? &nd->sub[0] : &nd->sub[1];
                s->el[s->nel++] = nd->el[i];
            }
            nd->sub[1].axisym = 0;
            construct(t, &nd->sub[0], level + 1);
            return;
        }
        if (nax == 0) {
            nd->axisym = 0;
            return;
        }
    }
    if (level < t->par.nbif && nd->nel >= 2) {
        double z1, z2;
        node_extrema(t, nd, &z1, &z2);
        const double zmid = (z1 + z2) * .5;
        zh_node lo, hi;
        node_init(&lo, nd->nel);
        node_init(&hi, nd->nel);
        for (int64_t i = 0; i < nd->nel; i++) {
            double a, b;
            el_extrema(t, nd->el[i], &a, &b);
            zh_node* s = (a + b) * .5 < zmid ? &lo : &hi;
            s->el[s->nel++] = nd->el[i];
        }
        const int n = (lo.nel > 0) + (hi.nel > 0);
        if (n > 1) {
            nd->nsub = 2;
            nd->sub = calloc(2, sizeof(zh_node));
            nd->sub[0] = lo;
            nd->sub[1] = hi;
            construct(t, &nd->sub[0], level + 1);
            construct(t, &nd->sub[1], level + 1);
            return;
        }
        free(lo.el);
        free(hi.el);
    }
}

/* ComputeCoefficients (KZonalHarmonicContainer.hh:284-345): the head carries source points
   only when it has no subcontainers; every lower level does */
static void coefficients(zh_tree* t, zh_node* nd, int level)
{
    if (nd->nel == 0) return;
    if (nd->axisym && ((level == -1 && nd->nsub == 0) || level != -1)) node_source_points(t, nd, level == -1);
    for (int i = 0; i < nd->nsub; i++) coefficients(t, &nd->sub[i], level + 1);
}

/* ---- evaluation */
static int use_central(const zh_tree* t, const zh_node* nd, const double P[3], int* idx)
{
    if (nd->ncen == 0) return 0;
    const float p0 = (float)P[0], p1 = (float)P[1], z = (float)P[2];
    const float r2 = p0 * p0 + p1 * p1;
    const float crs = (float)t->par.conv_ratio * (float)t->par.conv_ratio;
    float best = 1.e20f;
    for (int i = 0; i < nd->ncen; i++) {
        const float rhof = (float)nd->cen[i].rho;
        const double inv = 1. / (double)(rhof * rhof);
        const float delz = z - (float)nd->cen[i].z0;
        const float rc2 = (float)((r2 + delz * delz) * inv);
        if (rc2 < best) {
            best = rc2;
            *idx = i;
        }
    }
    return !(best > crs);
}

static int use_remote(const zh_tree* t, const zh_node* nd, const double P[3], int* idx)
{
    if (nd->nrem == 0) return 0;
    const float p0 = (float)P[0], p1 = (float)P[1], z = (float)P[2];
    const float r2 = p0 * p0 + p1 * p1;
    const float crs = (float)t->par.conv_ratio * (float)t->par.conv_ratio;
    float best = 1.e20f;
    for (int i = 0; i < nd->nrem; i++) {
        const float rhof = (float)nd->rem[i].rho;
        const double rho2 = (double)(rhof * rhof);
        const float delz = z - (float)nd->rem[i].z0;
        const float rr2 = (float)(rho2 / (r2 + delz * delz));
        if (rr2 < best) {
            best = rr2;
            *idx = i;
        }
    }
    return !(best > crs);
}

/* CentralExpansionFieldAndPotential (.cc:356-495) */
static int central(const zh_tree* t, const zh_sp* sp, const double P[3], double out[4])
{
    const double r = sqrt(P[0] * P[0] + P[1] * P[1]), z = P[2];
    const double* c = sp->c;
    const double prox = t->par.prox, conv = t->par.conv_param;
    if (r < prox && fabs(z - sp->z0) < prox) {
        out[2] = -c[1] / sp->rho;
        out[0] = out[1] = 0.;
        out[3] = c[0];
        return 1;
    }
    const double delz = z - sp->z0, rho = sqrt(r * r + delz * delz);
    const double u = delz / rho, s = r / rho, rc = rho / sp->rho;
    double rcn = rc;
    double Ez = c[1] + c[2] * 2. * rc * u, Er = c[2] * rc, Phi = c[0] + c[1] * rc * u;
    int cp = 0, cz = 0, cr = 0;
    double lp = 1.e30, lz = 1.e30, lr = 1.e30;
    double p1m1 = u, p1m2 = 1., pp1 = 1., pp2 = 0.;
    for (int n = 2; n < sp->n - 1; n++) {
        const double p1 = t->leg[0][n] * u * p1m1 - t->leg[1][n] * p1m2;
        const double p1p = t->leg[2][n] * u * pp1 - t->leg[3][n] * pp2;
        rcn *= rc;
        const double ez = c[n + 1] * (n + 1.) * rcn * p1, er = c[n + 1] * rcn * p1p, ph = c[n] * rcn * p1;
        Phi += ph;
        Ez += ez;
        Er += er;
        const double rp = conv * fabs(Phi), rz = conv * fabs(Ez), rrr = conv * fabs(Er);
        if (fabs(ph) < rp && fabs(lp) < rp) cp = 1;
        if (fabs(ez) < rz && fabs(lz) < rz) cz = 1;
        if ((fabs(er) < rrr && fabs(lr) < rrr) || r < prox) cr = 1;
        if (cp && cz && cr) break;
        lp = ph;
        lz = ez;
        lr = er;
        p1m2 = p1m1;
        p1m1 = p1;
        pp2 = pp1;
        pp1 = p1p;
    }
    if (!(cp && cz && cr)) return 0;
    Ez *= -1. / sp->rho;
    Er *= s / sp->rho;
    out[3] = Phi;
    out[2] = Ez;
    if (r < prox) {
        out[0] = out[1] = 0.;
    } else {
        out[0] = P[0] / r * Er;
        out[1] = P[1] / r * Er;
    }
    return 1;
}

/* RemoteExpansionFieldAndPotential (.cc:719-857) */
static int remote(const zh_tree* t, const zh_sp* sp, const double P[3], double out[4])
{
    const double r = sqrt(P[0] * P[0] + P[1] * P[1]), z = P[2];
    const double* c = sp->c;
    const double prox = t->par.prox, conv = t->par.conv_param;
    const double delz = z - sp->z0;
    double rho = sqrt(r * r + delz * delz);
    if (rho < 1.e-9) rho = 1.e-9;
    const double u = delz / rho, s = r / rho, rr = sp->rho / rho;
    double pn1 = c[0] * rr, zn1 = 0., rn1 = 0.;
    double rrn = rr * rr;
    double pn = c[1] * rrn * u, zn = c[0] * rrn * u, rn = c[0] * rrn;
    double Phi = pn1 + pn, Ez = zn1 + zn, Er = rn1 + rn;
    int cp = 0, cz = 0, cr = 0;
    double p1m1 = u, p1m2 = 1., pp1 = 1., pp2 = 0.;
    for (int n = 2; n < sp->n; n++) {
        const double p1 = t->leg[0][n] * u * p1m1 - t->leg[1][n] * p1m2;
        const double p1p = t->leg[2][n] * u * pp1 - t->leg[3][n] * pp2;
        rrn *= rr;
        pn = c[n] * rrn * p1;
        zn = c[n - 1] * n * rrn * p1;
        rn = c[n - 1] * rrn * p1p;
        Phi += pn;
        Ez += zn;
        Er += rn;
        const double rp = conv * fabs(Phi), rz = conv * fabs(Ez), rrr = conv * fabs(Er);
        if (fabs(pn) < rp && fabs(pn1) < rp) cp = 1;
        if (fabs(zn) < rz && fabs(zn1) < rz) cz = 1;
        if ((fabs(rn) < rrr && fabs(rn1) < rrr) || r < prox) cr = 1;
        if (cp && cz && cr) break;
        pn1 = pn;
        zn1 = zn;
        rn1 = rn;
        p1m2 = p1m1;
        p1m1 = p1;
        pp2 = pp1;
        pp1 = p1p;
    }
    if (!(cp && cz && cr)) return 0;
    Ez *= 1. / sp->rho;
    Er *= s / sp->rho;
    out[3] = Phi;
    out[2] = Ez;
    if (r < prox) {
        out[0] = out[1] = 0.;
    } else {
        out[0] = P[0] / r * Er;
        out[1] = P[1] / r * Er;
    }
    return 1;
}

/* the node's integrating field solver (Kahan sums over its elements in container order) */
static void direct(const zh_tree* t, const zh_node* nd, const double P[3], double out[4])
{
    double s[4] = {0., 0., 0., 0.}, cc[4] = {0., 0., 0., 0.};
    for (int64_t i = 0; i < nd->nel; i++) {
        const int64_t j = nd->el[i];
        double v[4];
        ko_element_eval(KO_NUMERIC, t->shape[j], t->recs + KO_REC * j, P, v);
        for (int k = 0; k < 4; k++) {
            const double y = v[k] * t->sigma[j] - cc[k];
            const double tt = s[k] + y;
            cc[k] = (tt - s[k]) - y;
            s[k] = tt;
        }
    }
    for (int k = 0; k < 4; k++) out[k] = s[k];
}

/* ElectricFieldAndPotential (.cc:98-124): central, remote, the sum over the subsets, direct */
static void eval_node(zh_tree* t, const zh_node* nd, const double P[3], double out[4])
{
    int i = -1;
    if (use_central(t, nd, P, &i) && central(t, &nd->cen[i], P, out)) {
        t->counts[0]++;
        return;
    }
    i = -1;
    if (use_remote(t, nd, P, &i) && remote(t, &nd->rem[i], P, out)) {
        t->counts[1]++;
        return;
    }
    if (nd->nsub > 0) {
        double acc[4] = {0., 0., 0., 0.};
        for (int k = 0; k < nd->nsub; k++) {
            double o[4];
            eval_node(t, &nd->sub[k], P, o);
            for (int q = 0; q < 4; q++) acc[q] = acc[q] + o[q];
        }
        for (int q = 0; q < 4; q++) out[q] = acc[q];
        return;
    }
    t->counts[2]++;
    direct(t, nd, P, out);
}

/* params: nbif, conv_ratio, prox, conv_param, ncen, frac_spacing, frac_dist, cen_dz, cen_z1,
   cen_z2, nrem, nrem_sp, rem_z1, rem_z2 (NULL: the reference's defaults).  The arrays must
   outlive the handle. */
void* ko_zh_build(int64_t n, const int32_t* shape, const double* recs, const double* sigma, const double* params)
{
    zh_tree* t = calloc(1, sizeof(zh_tree));
    t->n = n;
    t->shape = shape;
    t->recs = recs;
    t->sigma = sigma;
    const double def[14] = {5, 0.99, 1.e-14, 1.e-15, 500, 1, .2, .01, 0., 0., 500, 3, 0., 0.};
    const double* q = params ? params : def;
    t->par.nbif = (int)q[0];
    t->par.conv_ratio = q[1];
    t->par.prox = q[2];
    t->par.conv_param = q[3];
    t->par.ncen = (int)q[4];
    t->par.frac_spacing = q[5] != 0.;
    t->par.frac_dist = q[6];
    t->par.cen_dz = q[7];
    t->par.cen_z1 = q[8];
    t->par.cen_z2 = q[9];
    t->par.nrem = (int)q[10];
    t->par.nrem_sp = (int)q[11];
    t->par.rem_z1 = q[12];
    t->par.rem_z2 = q[13];
    legendre_rows(t, t->par.ncen > t->par.nrem ? t->par.ncen : t->par.nrem);
    node_init(&t->head, n);
    for (int64_t i = 0; i < n; i++) t->head.el[t->head.nel++] = i;
    construct(t, &t->head, -1);
    coefficients(t, &t->head, -1);
    return t;
}

void ko_zh_eval(void* h, int64_t m, const double* P, double* out, int64_t* counts)
{
    zh_tree* t = h;
    t->counts[0] = t->counts[1] = t->counts[2] = 0;
    for (int64_t p = 0; p < m; p++) eval_node(t, &t->head, P + 3 * p, out + 4 * p);
    if (counts)
        for (int k = 0; k < 3; k++) counts[k] = t->counts[k];
}

static void info_node(const zh_node* nd, int64_t* o)
{
    o[0]++;
    o[1] += nd->ncen;
    o[2] += nd->nrem;
    for (int i = 0; i < nd->nsub; i++) info_node(&nd->sub[i], o);
}

/* out[3]: nodes, central and remote source points over the tree */
void ko_zh_info(void* h, int64_t* out)
{
    out[0] = out[1] = out[2] = 0;
    info_node(&((zh_tree*)h)->head, out);
}

/* the first node (depth first) with central source points: its k-th source point's z0, rho
   and coefficients (out[2 + n]); returns the number of coefficients, 0 if none */
static const zh_node* first_with_sp(const zh_node* nd)
{
    if (nd->ncen) return nd;
    for (int i = 0; i < nd->nsub; i++) {
        const zh_node* r = first_with_sp(&nd->sub[i]);
        if (r) return r;
    }
    return 0;
}

int ko_zh_central_sp(void* h, int k, double* out)
{
    const zh_node* nd = first_with_sp(&((zh_tree*)h)->head);
    if (!nd || k < 0 || k >= nd->ncen) return 0;
    out[0] = nd->cen[k].z0;
    out[1] = nd->cen[k].rho;
    memcpy(out + 2, nd->cen[k].c, sizeof(double) * (size_t)nd->cen[k].n);
    return nd->cen[k].n;
}

static void free_node(zh_node* nd)
{
    for (int i = 0; i < nd->ncen; i++) free(nd->cen[i].c);
    for (int i = 0; i < nd->nrem; i++) free(nd->rem[i].c);
    free(nd->cen);
    free(nd->rem);
    for (int i = 0; i < nd->nsub; i++) free_node(&nd->sub[i]);
    free(nd->sub);
    free(nd->el);
}

void ko_zh_free(void* h)
{
    zh_tree* t = h;
    if (!t) return;
    free_node(&t->head);
    for (int r = 0; r < 4; r++) free(t->leg[r]);
    free(t);
}
