// Drives the library through the C++ facade (include/kemhip.hpp) the way KEMField host code
// drives KEMField: build a container, assemble, solve (GMRES, Gauss, Robin Hood), evaluate the
// field.  Unit cube, 4 x 4 squares per face, two triangles per square, 1 V
// (CubeCapacitanceTest.cc-style).  Prints "facade OK" and exits 0 when every check passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "kemhip.hpp"

using namespace kemhip;

static int fails = 0;
static void check(bool ok, const char* what, double v)
{
    std::printf("%-48s %s (%.6e)\n", what, ok ? "ok" : "FAILED", v);
    if (!ok) fails++;
}

// raw dump of the step-level root field for the oracle check (tests/test_gpu_facade.py)
static void dump_mesh(std::FILE* f, const SurfaceData& s, const std::vector<double>& sigma)
{
    const int64_t n = (int64_t)s.size();
    std::fwrite(&n, sizeof n, 1, f);
    std::fwrite(s.shape.data(), sizeof(int32_t), (size_t)n, f);
    std::fwrite(s.data.data(), sizeof(double), (size_t)(11 * n), f);
    std::fwrite(sigma.data(), sizeof(double), (size_t)n, f);
}

int main(int argc, char** argv)
{
    Context ctx(0);
    SurfaceData s;
    const int n = 4;
    for (int axis = 0; axis < 3; axis++)
        for (int sign = -1; sign <= 1; sign += 2)
            for (int a = 0; a < n; a++)
                for (int b = 0; b < n; b++) {
                    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
                    auto P = [&](int ia, int ib) {
                        Position p{};
                        p[axis] = 0.5 * sign;
                        p[u] = -0.5 + (double)ia / n;
                        p[v] = -0.5 + (double)ib / n;
                        return p;
                    };
                    // outward orientation: (p1 - p0) x (p2 - p0) along sign * e_axis
                    if (sign > 0) {
                        s.AddTriangle(P(a, b), P(a + 1, b), P(a + 1, b + 1), 1.);
                        s.AddTriangle(P(a + 1, b + 1), P(a, b + 1), P(a, b), 1.);
                    } else {
                        s.AddTriangle(P(a, b), P(a, b + 1), P(a + 1, b + 1), 1.);
                        s.AddTriangle(P(a + 1, b + 1), P(a + 1, b), P(a, b), 1.);
                    }
                }
    KSurfaceContainer cont(ctx, s);
    KElectrostaticBoundaryIntegrator integ = KElectrostaticBoundaryIntegratorFactory::MakeDefault(ctx);
    KBoundaryIntegralMatrix A(cont, integ);
    const std::vector<double> b = cont.BoundaryValues();
    check(A.Dimension() == cont.size(), "KSquareMatrix::Dimension", A.Dimension());

    KKrylovChargeDensitySolver gmres;
    gmres.SetTolerance(1e-10);
    std::vector<double> x;
    kem_solve_stats st = gmres.Solve(A, b, x);
    check(st.residual <= 1e-10, "GMRES converged to 1e-10", st.residual);
    KGaussianEliminationChargeDensitySolver gauss;
    std::vector<double> xg;
    gauss.Solve(A, b, xg);
    double d = 0., nx = 0.;
    for (size_t i = 0; i < x.size(); i++) {
        d += (x[i] - xg[i]) * (x[i] - xg[i]);
        nx += x[i] * x[i];
    }
    check(std::sqrt(d / nx) < 1e-8, "Gauss-Jordan == GMRES", std::sqrt(d / nx));
    KRobinHoodChargeDensitySolver rh;
    rh.SetTolerance(1e-8);
    std::vector<double> xr;
    rh.Solve(A, b, xr);
    d = 0.;
    for (size_t i = 0; i < x.size(); i++) d += (x[i] - xr[i]) * (x[i] - xr[i]);
    check(std::sqrt(d / nx) < 1e-5, "Robin Hood == GMRES", std::sqrt(d / nx));

    // operator()(i, j) against Multiply with a unit vector
    std::vector<double> e(A.Dimension(), 0.), col;
    e[7] = 1.;
    A.Multiply(e, col);
    check(std::fabs(col[3] - A(3, 7)) <= 1e-15 * std::fabs(col[3]), "operator() == Multiply(e_j)", col[3]);

    // capacitance Q / (4 pi eps0 V); area of each triangle = (1/n)^2 / 2
    double Q = 0.;
    for (double sig : x) Q += sig * 0.5 / (n * n);
    const double C = Q / (4. * M_PI * 8.8541878128E-12);
    check(std::fabs(C - 0.6533) < 2e-3, "capacitance / (4 pi eps0)", C);

    KElectrostaticBoundaryField field(cont, integ, gmres);
    field.Initialize();
    const double phic = field.Potential({0.1, -0.05, 0.2});
    check(std::fabs(phic - 1.) < 1e-2, "potential inside the conductor ~ 1 V", phic);
    const Position far{0., 0., 20.};
    const auto ep = field.ElectricFieldAndPotential(far);
    check(std::fabs(ep.second - C / 20.) < 2e-3 * C / 20., "far potential ~ C/r (monopole)", ep.second);
    check(std::fabs(ep.first[2] * 400. / C - 1.) < 1e-2, "far field ~ C/r^2", ep.first[2] * 400. / C);

    KIntegratingElectrostaticFieldSolver mixed(cont, x, true);
    const auto em = mixed.ElectricFieldAndPotential(far);
    check(std::fabs(em.second - ep.second) < 1e-6 * std::fabs(ep.second), "mixed field == FP64 field (1e-6)",
          std::fabs(em.second - ep.second) / std::fabs(ep.second));

    // Kassiopeia's field component: single-point calls and one batched call over many sample
    // points give the field solver's answers
    {
        KSElectricKEMField ks(field);
        ks.InitializeComponent();
        std::vector<Position> pts;
        for (int k = 0; k < 5000; k++) {
            const double t = 0.37 * k, r = 2. + 3. * (k % 17) / 17.;
            pts.push_back({r * std::cos(t) * std::sin(0.1 + 0.0006 * k), r * std::sin(t) * std::sin(0.1 + 0.0006 * k),
                           r * std::cos(0.1 + 0.0006 * k)});
        }
        std::vector<FieldVector> Eb;
        std::vector<double> pb;
        ks.CalculateFieldAndPotential(pts, 0., Eb, pb);
        double worst = 0.;
        for (int k = 0; k < 5000; k += 499) {
            const auto ref = field.ElectricFieldAndPotential(pts[(size_t)k]);
            double phi1;
            FieldVector E1;
            ks.CalculatePotential(pts[(size_t)k], 0., phi1);
            ks.CalculateField(pts[(size_t)k], 0., E1);
            const double en = std::sqrt(ref.first[0] * ref.first[0] + ref.first[1] * ref.first[1] +
                                        ref.first[2] * ref.first[2]);
            for (int c = 0; c < 3; c++) {
                worst = std::max(worst, std::fabs(Eb[(size_t)k][(size_t)c] - ref.first[(size_t)c]) / en);
                worst = std::max(worst, std::fabs(E1[(size_t)c] - ref.first[(size_t)c]) / en);
            }
            worst = std::max(worst, std::fabs(pb[(size_t)k] - ref.second) / std::fabs(ref.second));
            worst = std::max(worst, std::fabs(phi1 - ref.second) / std::fabs(ref.second));
        }
        check(worst < 1e-12, "KSElectricKEMField batch and single calls == field solver (1e-12)", worst);
        ks.DeinitializeComponent();
        bool threw_state = false;
        try {
            ks.CalculateFieldAndPotential(pts, 0., Eb, pb);
        } catch (const Error&) {
            threw_state = true;
        }
        check(threw_state, "KSElectricKEMField before InitializeComponent throws", 0.);
    }

    // the same cube from rectangles, 6 x 6 per face (CubeCapacitanceTest.cc:106-116: the
    // capacitance 0.6606785 within 1e-2)
    SurfaceData r;
    const int nr = 6;
    for (int axis = 0; axis < 3; axis++)
        for (int sign = -1; sign <= 1; sign += 2)
            for (int a = 0; a < nr; a++)
                for (int b = 0; b < nr; b++) {
                    const int u = (axis + 1) % 3, v = (axis + 2) % 3;
                    auto P = [&](int ia, int ib) {
                        Position p{};
                        p[axis] = 0.5 * sign;
                        p[u] = -0.5 + (double)ia / nr;
                        p[v] = -0.5 + (double)ib / nr;
                        return p;
                    };
                    if (sign > 0)
                        r.AddRectangle(P(a, b), P(a + 1, b), P(a, b + 1), 1.);
                    else
                        r.AddRectangle(P(a, b), P(a, b + 1), P(a + 1, b), 1.);
                }
    KSurfaceContainer rcont(ctx, r);
    KBoundaryIntegralMatrix Ar(rcont, integ);
    std::vector<double> xr2;
    gmres.Solve(Ar, rcont.BoundaryValues(), xr2);
    double Qr = 0.;
    for (double sig : xr2) Qr += sig / (nr * nr);
    const double Cr = Qr / (4. * M_PI * 8.8541878128E-12);
    check(std::fabs(Cr - 0.6606785) < 1e-2, "rectangle cube capacitance (0.6606785, 1e-2)", Cr);

    // a conducting unit disk of 64 conic sections (DiskCapacitanceTest.cc:63-113: C = 8 eps0 R
    // within 1e-2), graded towards rim and centre as the reference's test grades it
    {
        SurfaceData dsk;
        const int nseg = 64;
        std::vector<double> seg(2 * nseg);
        for (int i = 0; i < nseg; i++) {
            const double a = std::pow((double)i / nseg, 1.5), b = std::pow((double)(i + 1) / nseg, 1.5);
            seg[i] = seg[2 * nseg - 1 - i] = b - a;
        }
        double r0 = 1.;
        for (int i = 0; i < nseg; i++) {
            const double r1 = i == nseg - 1 ? 0. : r0 - seg[i];
            dsk.AddConicSection(r0, 0., r1, 0., 1.);
            r0 = r1;
        }
        KSurfaceContainer dcont(ctx, dsk);
        KBoundaryIntegralMatrix Ad(dcont, integ);
        std::vector<double> xd;
        gmres.Solve(Ad, dcont.BoundaryValues(), xd);
        double Qd = 0.;
        for (int i = 0; i < nseg; i++) {
            const double a = dsk.data[11 * i], b = dsk.data[11 * i + 3];
            Qd += M_PI * (a + b) * std::fabs(a - b) * xd[(size_t)i];  // KConicSection::Area
        }
        const double Cd = Qd / 8.8541878128E-12;
        check(std::fabs(Cd - 8.) < 1e-2, "conic-section disk capacitance C/eps0 (8, 1e-2)", Cd);
        SurfaceData rg;
        rg.AddRing(0.5, 0., 1.);
        KSurfaceContainer rcont1(ctx, rg);
        KIntegratingElectrostaticFieldSolver ring(rcont1, {1.}, false);
        // unit charge on the ring: phi on the axis = 1/(4 pi eps0 sqrt(R^2 + z^2))
        const double pr = ring.Potential({0., 0., 0.3});
        const double want = 1. / (4. * M_PI * 8.8541878128E-12 * std::sqrt(0.25 + 0.09));
        check(std::fabs(pr - want) < 1e-12 * want, "ring potential on the axis (closed form)", pr / want - 1.);
    }

    // Kassiopeia's root field over two components (the triangle cube and the rectangle cube):
    // one call per step for all sample points, the components summed per point in the order
    // they were added; equal to the two components' own batches added
    {
        KElectrostaticBoundaryField field_r(rcont, integ, gmres);
        KSElectricKEMField k1(field), k2(field_r);
        k1.InitializeComponent();
        k2.InitializeComponent();
        KSRootElectricField root;
        root.AddElectricField(&k1);
        root.AddElectricField(&k2);
        std::vector<Position> pts;
        for (int k = 0; k < 3000; k++) {
            const double t = 0.71 * k, r = 0.7 + 2.5 * (k % 23) / 23.;
            pts.push_back({r * std::cos(t) * std::sin(0.2 + 0.0009 * k), r * std::sin(t) * std::sin(0.2 + 0.0009 * k),
                           r * std::cos(0.2 + 0.0009 * k)});
        }
        std::vector<FieldVector> E, E1, E2;
        std::vector<double> phi, p1, p2;
        root.CalculateFieldAndPotential(pts, 0., E, phi);
        k1.CalculateFieldAndPotential(pts, 0., E1, p1);
        k2.CalculateFieldAndPotential(pts, 0., E2, p2);
        double worst = 0.;
        for (size_t k = 0; k < pts.size(); k++) {
            worst = std::max(worst, std::fabs(phi[k] - (p1[k] + p2[k])));
            for (int c = 0; c < 3; c++) worst = std::max(worst, std::fabs(E[k][(size_t)c] - (E1[k][(size_t)c] + E2[k][(size_t)c])));
        }
        check(worst == 0., "KSRootElectricField == sum of its components", worst);
        FieldVector Es;
        double ps;
        root.CalculateFieldAndPotential(pts[17], 0., Es, ps);
        check(ps == phi[17] && Es[2] == E[17][2], "KSRootElectricField single point == batch", ps - phi[17]);
        if (argc > 1) {
            std::FILE* f = std::fopen(argv[1], "wb");
            if (!f) {
                check(false, "dump file opens", 0.);
            } else {
                const int64_t m = (int64_t)pts.size();
                std::fwrite(&m, sizeof m, 1, f);
                for (const Position& P : pts) std::fwrite(P.data(), sizeof(double), 3, f);
                for (const FieldVector& v : E) std::fwrite(v.data(), sizeof(double), 3, f);
                std::fwrite(phi.data(), sizeof(double), (size_t)m, f);
                dump_mesh(f, s, field.ChargeDensities());
                dump_mesh(f, r, field_r.ChargeDensities());
                std::fclose(f);
            }
        }
    }

    bool threw = false;
    try {
        KElectrostaticBoundaryIntegratorFactory::Make(ctx, "no-such-integrator");
    } catch (const KEMSimpleException&) {
        threw = true;
    }
    check(threw, "unknown integrator throws KEMSimpleException", 0.);
    integ.Select();
    if (fails == 0) std::printf("facade OK\n");
    return fails == 0 ? 0 : 1;
}
