// kem_zh_gen.cpp -- zonal-harmonic expansion of the field of axisymmetric elements (host side):
// the container tree, its source points along the axis and their Legendre coefficients,
// flattened into the arrays the device evaluator (kem_zh.hip) walks.
//
// KEMField's KZonalHarmonicContainer / KZonalHarmonicCoefficientGenerator
// (FieldSolvers/ZonalHarmonic/Generator): rings and conic sections about the z axis; the
// container is split at the middle of its axial extent NBifurcations times, every node below
// the head gets central source points (fractional or fixed spacing) and remote source points,
// each holding the Legendre coefficients of the node's potential on the axis about that point.
#include "kem_zh.h"

#include <algorithm>
#include <cmath>

namespace kem {
namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;
constexpr double kEps0 = 8.8541878128e-12;

struct Node {
    std::vector<int64_t> el;
    bool axisym = true;
    std::vector<int32_t> sub;  // child node indices
    std::vector<int32_t> cen, rem;  // source point indices
};

struct Builder {
    const ZHElement* e;
    ZHParams p;
    std::vector<Node> nodes;
    ZHTables out;

    double leg(int row, int n) const { return out.leg[(size_t)row * out.nleg + n]; }

    void extrema(int64_t k, double& z1, double& z2) const
    {
        if (e[k].type == ZHElement::kRing) {
            z1 = z2 = e[k].z0;
        } else {
            z1 = std::min(e[k].z0, e[k].z1);
            z2 = std::max(e[k].z0, e[k].z1);
        }
    }

    // distance from the axis point z0 to the element: nearest (central) or farthest (remote)
    // (KZHCoefficientGeneratorRing.cc:74-88, KZHCoefficientGeneratorConicSection.cc:133-173)
    double rho(int64_t k, double z0, bool cen) const
    {
        const ZHElement& a = e[k];
        if (a.type == ZHElement::kRing) return std::sqrt((a.z0 - z0) * (a.z0 - z0) + a.r0 * a.r0);
        const double Z = (a.z0 + a.z1) * .5 - z0, R = (a.r0 + a.r1) * .5;
        const double L = std::sqrt((a.z0 - a.z1) * (a.z0 - a.z1) + (a.r0 - a.r1) * (a.r0 - a.r1));
        const double c = (a.z1 - a.z0) / L, s = (a.r1 - a.r0) / L;
        const double x = -Z * c - R * s, xm = L / 2.;
        const double ra = std::sqrt((Z - xm * c) * (Z - xm * c) + (R - xm * s) * (R - xm * s));
        const double rb = std::sqrt((Z + xm * c) * (Z + xm * c) + (R + xm * s) * (R + xm * s));
        if (x >= xm) return cen ? rb : ra;
        if (x <= -xm) return cen ? ra : rb;
        return cen ? std::sqrt((Z + x * c) * (Z + x * c) + (R + x * s) * (R + x * s)) : std::max(ra, rb);
    }

    // adds element k's coefficients about z0 (radius rho) to c[0..nc)
    void coeffs(int64_t k, double z0, double rh, int nc, double* c, bool cen, std::vector<double>& P) const
    {
        const ZHElement& a = e[k];
        P.assign((size_t)nc, 0.);
        if (a.type == ZHElement::kRing) {
            // the ring's on-axis potential sigma / (4 pi eps0 |Q - z0|) expanded in Legendre
            // polynomials of its polar angle (KZHCoefficientGeneratorRing.cc:18-72)
            const double rr = std::sqrt((z0 - a.z0) * (z0 - a.z0) + a.r0 * a.r0);
            const double pre = a.sigma / (4. * kPi * kEps0 * rr);
            const double ratio = cen ? rh / rr : rr / rh;
            P[0] = 1.;
            P[1] = (a.z0 - z0) / rr;
            double ph = cen ? pre : pre * ratio;
            c[0] += ph;
            ph *= ratio;
            c[1] += ph * P[1];
            for (int i = 2; i < nc; i++) {
                ph *= ratio;
                P[i] = leg(0, i) * P[1] * P[i - 1] - leg(1, i) * P[i - 2];
                c[i] += ph * P[i];
            }
            return;
        }
        // the conic section as rings along its generating line: a 31-node composite rule with
        // Glück's modified Newton-Cotes end weights (KZHCoefficientGeneratorConicSection.cc:37-130)
        static const double w9[10] = {0.2803440531305107e0, 0.1648702325837748e1, -0.2027449845679092e0,
                                      0.2797927414021179e1, -0.9761199294532843e0, 0.2556499393738999e1,
                                      0.1451083002645404e0, 0.1311227127425048e1, 0.9324249063051143e0,
                                      0.1006631393298060e1};
        constexpr int M = 30;
        double w[M + 1];
        for (int m = 0; m <= M; m++) w[m] = m <= 9 ? w9[m] : (m >= M - 9 ? w9[M - m] : 1.);
        const double L = std::sqrt((a.z0 - a.z1) * (a.z0 - a.z1) + (a.r0 - a.r1) * (a.r0 - a.r1));
        const double del = L / M;
        for (int m = 0; m <= M; m++) {
            const double x = -L / 2. + del * m;
            const double Z = (a.z0 + a.z1) / 2. + x * (a.z1 - a.z0) / L;
            const double R = (a.r0 + a.r1) / 2. + x * (a.r1 - a.r0) / L;
            double ro0 = std::sqrt(R * R + (Z - z0) * (Z - z0));
            if (ro0 == 0) ro0 = 1.e-17;
            const double rper = R / ro0, u0 = (Z - z0) / ro0;
            P[0] = 1.;
            P[1] = u0;
            const double rc = cen ? rh / ro0 : ro0 / rh;
            double rcn = rc;
            const double cm = a.sigma / (2. * kEps0) * w[m] * del * rper;
            if (cen) {
                c[0] += cm * P[0];
                c[1] += cm * rcn * P[1];
            } else {
                c[0] += cm * rcn * P[0];
                rcn *= rc;
                c[1] += cm * rcn * P[1];
            }
            for (int n = 2; n < nc; n++) {
                rcn *= rc;
                P[n] = leg(0, n) * u0 * P[n - 1] - leg(1, n) * P[n - 2];
                c[n] += cm * rcn * P[n];
                if (rcn < 1.e-17) break;
            }
        }
    }

    void node_extrema(const Node& nd, double& z1, double& z2) const
    {
        extrema(nd.el[0], z1, z2);
        for (size_t i = 1; i < nd.el.size(); i++) {
            double a, b;
            extrema(nd.el[i], a, b);
            z1 = std::min(z1, a);
            z2 = std::max(z2, b);
        }
    }

    double node_rho(const Node& nd, double z, bool cen) const
    {
        double r = cen ? 1.e10 : 0.;
        for (int64_t k : nd.el) {
            const double v = rho(k, z, cen);
            if (cen ? v < r : v > r) r = v;
        }
        return r;
    }

    int32_t add_sp(const Node& nd, double z, int nc, bool cen)
    {
        ZHSourcePoint sp;
        sp.z0 = z;
        sp.rho = node_rho(nd, z, cen);
        sp.ncoef = nc;
        sp.offset = (int64_t)out.coeff.size();
        out.coeff.resize(out.coeff.size() + (size_t)nc, 0.);
        std::vector<double> P;
        for (int64_t k : nd.el) coeffs(k, z, sp.rho, nc, out.coeff.data() + sp.offset, cen, P);
        out.sp.push_back(sp);
        return (int32_t)out.sp.size() - 1;
    }

    // source points of one node (KZonalHarmonicCoefficientGenerator.hh:221-360); the head's
    // may be limited to a user range
    void source_points(int32_t id, bool head)
    {
        double z1 = head ? p.cen_z1 : 0., z2 = head ? p.cen_z2 : 0.;
        if (z2 < z1) std::swap(z1, z2);
        if (std::fabs(z1 - z2) < 1.e-10) node_extrema(nodes[id], z1, z2);
        std::vector<int32_t> cen;
        if (p.frac_spacing) {
            cen.push_back(add_sp(nodes[id], z1, p.ncen, true));
            double z = z1;
            for (;;) {
                double dz = p.frac_dist * node_rho(nodes[id], z, true);
                if (dz < p.cen_dz) dz = p.cen_dz;
                z += dz;
                if (z >= z2) break;
                cen.push_back(add_sp(nodes[id], z, p.ncen, true));
            }
            cen.push_back(add_sp(nodes[id], z2, p.ncen, true));
        } else {
            unsigned nsp = (unsigned)((z2 - z1) / p.cen_dz);
            if (nsp == 0) nsp = 1;
            for (unsigned i = 0; i < nsp; i++) cen.push_back(add_sp(nodes[id], z1 + p.cen_dz * i, p.ncen, true));
        }
        nodes[id].cen = cen;
        z1 = head ? p.rem_z1 : 0.;
        z2 = head ? p.rem_z2 : 0.;
        if (z2 < z1) std::swap(z1, z2);
        if (std::fabs(z1 - z2) < 1.e-10) node_extrema(nodes[id], z1, z2);
        unsigned nsp = (unsigned)p.nrem_sp;
        double dz = (z2 - z1) / (nsp - 1);
        if (std::fabs(z1 - z2) < 1.e-10) {
            nsp = 1;
            dz = 0.;
        }
        std::vector<int32_t> rem;
        for (unsigned i = 0; i < nsp; i++) rem.push_back(add_sp(nodes[id], z1 + dz * i, p.nrem, false));
        nodes[id].rem = rem;
    }

    int32_t new_node(std::vector<int64_t> el, bool axisym)
    {
        Node nd;
        nd.el = std::move(el);
        nd.axisym = axisym;
        nodes.push_back(std::move(nd));
        return (int32_t)nodes.size() - 1;
    }

    // the container tree (KZonalHarmonicContainer.hh:246-282): at the head, axisymmetric and
    // other elements form separate groups; every level below NBifurcations splits a node of at
    // least two elements at the middle of its axial extent by element centre
    void construct(int32_t id, int level)
    {
        if (nodes[id].el.empty()) return;
        if (level == -1) {
            std::vector<int64_t> ax, other;
            for (int64_t k : nodes[id].el) (e[k].type == ZHElement::kOther ? other : ax).push_back(k);
            if (!ax.empty() && !other.empty()) {
                const int32_t a = new_node(ax, true), b = new_node(other, false);
                nodes[id].sub = {a, b};
                construct(a, level + 1);
                return;
            }
            if (ax.empty()) {
                nodes[id].axisym = false;
                return;
            }
        }
        if (level < p.nbif && nodes[id].el.size() >= 2) {
            double z1, z2;
            node_extrema(nodes[id], z1, z2);
            const double zmid = (z1 + z2) * .5;
            std::vector<int64_t> lo, hi;
            for (int64_t k : nodes[id].el) {
                double a, b;
                extrema(k, a, b);
                ((a + b) * .5 < zmid ? lo : hi).push_back(k);
            }
            if (!lo.empty() && !hi.empty()) {
                const int32_t a = new_node(lo, true), b = new_node(hi, true);
                nodes[id].sub = {a, b};
                construct(a, level + 1);
                construct(b, level + 1);
            }
        }
    }

    // the head carries source points only without subcontainers, every other node always
    // (KZonalHarmonicContainer.hh:284-345)
    void coefficients(int32_t id, int level)
    {
        if (nodes[id].el.empty()) return;
        if (nodes[id].axisym && (level != -1 || nodes[id].sub.empty())) source_points(id, level == -1);
        const std::vector<int32_t> sub = nodes[id].sub;
        for (int32_t s : sub) coefficients(s, level + 1);
    }

    int depth(int32_t id) const
    {
        int d = 0;
        for (int32_t s : nodes[id].sub) d = std::max(d, depth(s));
        return d + 1;
    }
};

}  // namespace

ZHTables zh_build(const std::vector<ZHElement>& elements, const ZHParams& params)
{
    Builder b;
    b.e = elements.data();
    b.p = params;
    // recursion coefficients of P_n and P_n' (KZHLegendreCoefficients.cc:38-44), rows 0-3
    b.out.nleg = std::max(2, std::max(params.ncen, params.nrem));
    b.out.leg.assign((size_t)4 * b.out.nleg, 0.);
    for (int n = 2; n < b.out.nleg; n++) {
        b.out.leg[(size_t)0 * b.out.nleg + n] = (2. * n - 1.) / (1. * n);
        b.out.leg[(size_t)1 * b.out.nleg + n] = (n - 1.) / (1. * n);
        b.out.leg[(size_t)2 * b.out.nleg + n] = (2. * n - 1.) / (1. * (n - 1.));
        b.out.leg[(size_t)3 * b.out.nleg + n] = (1. * n) / (1. * (n - 1.));
    }
    std::vector<int64_t> all(elements.size());
    for (size_t i = 0; i < all.size(); i++) all[i] = (int64_t)i;
    b.new_node(all, true);
    if (!elements.empty()) {
        b.construct(0, -1);
        b.coefficients(0, -1);
    }
    // flatten: per node [central begin, count, remote begin, count, children begin, count,
    // elements begin, count]
    ZHTables& t = b.out;
    for (const Node& nd : b.nodes) {
        const int32_t rec[8] = {nd.cen.empty() ? 0 : nd.cen.front(), (int32_t)nd.cen.size(),
                                nd.rem.empty() ? 0 : nd.rem.front(), (int32_t)nd.rem.size(),
                                (int32_t)t.child.size(), (int32_t)nd.sub.size(),
                                (int32_t)t.elem.size(), (int32_t)nd.el.size()};
        t.node.insert(t.node.end(), rec, rec + 8);
        t.child.insert(t.child.end(), nd.sub.begin(), nd.sub.end());
        t.elem.insert(t.elem.end(), nd.el.begin(), nd.el.end());
    }
    t.depth = elements.empty() ? 1 : b.depth(0);
    return t;
}

}  // namespace kem
