/*
 * kemhip.hpp -- C++ facade over the C-ABI (kemhip.h) with KEMField's class and method names,
 * header-only, for host code that drives the MI355X library the way it drives KEMField.
 *
 * Mirrors (paths under KEMField/Source):
 *   KSurfaceContainer                 Surfaces/include/KSurfaceContainer.hh
 *   KElectrostaticBoundaryIntegrator  BoundaryIntegrals/Electrostatic/Core/include/KElectrostaticBoundaryIntegrator.hh:57-61
 *   KElectrostaticBoundaryIntegratorFactory::Make  .../Core/src/KElectrostaticBoundaryIntegratorFactory.cc:78-88
 *   KBoundaryIntegralMatrix (KSquareMatrix<double>)  BoundaryIntegrals/Core/include/KBoundaryIntegralMatrix.hh:16-99,
 *                                     LinearAlgebra/Core/include/KMatrix.hh:15-111
 *   KKrylovChargeDensitySolver / KRobinHoodChargeDensitySolver / KGaussianEliminationChargeDensitySolver
 *                                     Interface/ChargeDensitySolvers/Electric/src/ (one .cc each)
 *   KIntegratingElectrostaticFieldSolver  Interface/FieldSolvers/Electric/src/KIntegratingElectrostaticFieldSolver.cc:44-89
 *   KElectrostaticBoundaryField       Interface/Fields/Electric/src/KElectrostaticBoundaryField.cc:16-199
 *   KSElectricKEMField                Kassiopeia/Fields/Source/KSElectricKEMField.cxx:44-72 (outside KEMField)
 *
 * Errors: a negative C-ABI code throws kemhip::Error (unknown integrator / solver names throw
 * kemhip::KEMSimpleException, as the reference's factory does); KEM_W_NOT_CONVERGED is
 * returned in the solver statistics, not thrown (the reference warns).
 */
#ifndef KEMHIP_HPP
#define KEMHIP_HPP

#include <array>
#include <cmath>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "kemhip.h"

namespace kemhip {

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
struct KEMSimpleException : Error {
    using Error::Error;
};

using Position = std::array<double, 3>;
using FieldVector = std::array<double, 3>;

class Context {
  public:
    explicit Context(int device = 0)
    {
        const int rc = kem_ctx_create(device, &h_);
        if (rc != KEM_OK) throw Error(rc, "kem_ctx_create failed (no HIP device?)");
    }
    ~Context() { kem_ctx_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    kem_ctx* get() const { return h_; }
    int check(int rc) const
    {
        if (rc >= 0) return rc;
        const std::string msg = kem_last_error(h_);
        if (rc == KEM_E_NAME) throw KEMSimpleException(rc, msg);
        throw Error(rc, msg);
    }

  private:
    kem_ctx* h_ = nullptr;
};

// A flattened container: shape type per element, the reference's 11-double shape records
// (triangles / rectangles [A, B, P0, N1, N2], segments [P0, P1, diameter]), boundary type,
// value (Dirichlet potential) and flux (Neumann) in container order.
struct SurfaceData {
    std::vector<int32_t> shape;
    std::vector<double> data;
    std::vector<int32_t> bc;
    std::vector<double> value, flux;
    size_t size() const { return shape.size(); }
    // KTriangle::SetValues(p0, p1, p2) (Surfaces/src/KTriangle.cc:18-28)
    void AddTriangle(const Position& p0, const Position& p1, const Position& p2, double potential)
    {
        double n1[3], n2[3];
        for (int k = 0; k < 3; k++) {
            n1[k] = p1[k] - p0[k];
            n2[k] = p2[k] - p0[k];
        }
        const double a = norm(n1), b = norm(n2);
        shape.push_back(KEM_SHAPE_TRIANGLE);
        data.insert(data.end(), {a, b, p0[0], p0[1], p0[2], n1[0] / a, n1[1] / a, n1[2] / a, n2[0] / b, n2[1] / b,
                                 n2[2] / b});
        bc.push_back(KEM_BC_DIRICHLET);
        value.push_back(potential);
        flux.push_back(0.);
    }
    // KRectangle::SetValues(p0, p1, p2, p3) (Surfaces/src/KRectangle.cc:18-28): sides p1 - p0
    // and p3 - p0 (p2, the corner opposite p0, is not used)
    void AddRectangle(const Position& p0, const Position& p1, const Position& p3, double potential)
    {
        double n1[3], n2[3];
        for (int k = 0; k < 3; k++) {
            n1[k] = p1[k] - p0[k];
            n2[k] = p3[k] - p0[k];
        }
        const double a = norm(n1), b = norm(n2);
        shape.push_back(KEM_SHAPE_RECTANGLE);
        data.insert(data.end(), {a, b, p0[0], p0[1], p0[2], n1[0] / a, n1[1] / a, n1[2] / a, n2[0] / b, n2[1] / b,
                                 n2[2] / b});
        bc.push_back(KEM_BC_DIRICHLET);
        value.push_back(potential);
        flux.push_back(0.);
    }
    // KLineSegment::SetValues(p0, p1, diameter) (Surfaces/include/KLineSegment.hh:23)
    void AddLineSegment(const Position& p0, const Position& p1, double diameter, double potential)
    {
        shape.push_back(KEM_SHAPE_LINESEGMENT);
        data.insert(data.end(), {p0[0], p0[1], p0[2], p1[0], p1[1], p1[2], diameter, 0., 0., 0., 0.});
        bc.push_back(KEM_BC_DIRICHLET);
        value.push_back(potential);
        flux.push_back(0.);
    }
    // KConicSection::SetValues(r0, z0, r1, z1) (Surfaces/src/KConicSection.cc:13-22): the surface
    // of revolution about z of the segment (r0, z0)-(r1, z1)
    void AddConicSection(double r0, double z0, double r1, double z1, double potential)
    {
        shape.push_back(KEM_SHAPE_CONICSECTION);
        data.insert(data.end(), {r0, 0., z0, r1, 0., z1, 0., 0., 0., 0., 0.});
        bc.push_back(KEM_BC_DIRICHLET);
        value.push_back(potential);
        flux.push_back(0.);
    }
    // KRing::SetValues(r, z) (Surfaces/src/KRing.cc:14-18): a charged circle of radius r at height z
    void AddRing(double r, double z, double potential)
    {
        shape.push_back(KEM_SHAPE_RING);
        data.insert(data.end(), {r, 0., z, 0., 0., 0., 0., 0., 0., 0., 0.});
        bc.push_back(KEM_BC_DIRICHLET);
        value.push_back(potential);
        flux.push_back(0.);
    }
    // turns element i into a Neumann (dielectric interface) element with the given flux
    // (KNeumannBoundary's normal boundary flux, the permittivity ratio of the interface)
    void SetNeumann(size_t i, double normal_flux)
    {
        bc.at(i) = KEM_BC_NEUMANN;
        value.at(i) = 0.;
        flux.at(i) = normal_flux;
    }

  private:
    static double norm(const double* v) { return std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }
};

class KSurfaceContainer {
  public:
    KSurfaceContainer(const Context& ctx, const SurfaceData& s) : ctx_(ctx), n_((int64_t)s.size())
    {
        ctx.check(kem_mesh_create(ctx.get(), n_, s.shape.data(), s.data.data(), s.bc.data(), s.value.data(),
                                  s.flux.data(), &h_));
    }
    ~KSurfaceContainer() { kem_mesh_destroy(h_); }
    KSurfaceContainer(const KSurfaceContainer&) = delete;
    KSurfaceContainer& operator=(const KSurfaceContainer&) = delete;
    size_t size() const { return (size_t)n_; }
    kem_mesh* get() const { return h_; }
    const Context& context() const { return ctx_; }
    // boundary values b_i (KBoundaryIntegralVector.hh:60-75)
    std::vector<double> BoundaryValues() const
    {
        std::vector<double> b(size());
        ctx_.check(kem_mesh_rhs(h_, b.data()));
        return b;
    }

  private:
    const Context& ctx_;
    int64_t n_;
    kem_mesh* h_ = nullptr;
};

class KElectrostaticBoundaryIntegrator {
  public:
    KElectrostaticBoundaryIntegrator(const Context& ctx, std::string name = "numeric") : ctx_(ctx), name_(std::move(name))
    {
        ctx_.check(kem_set_integrator(ctx_.get(), name_.c_str()));
    }
    const std::string& Name() const { return name_; }
    void Select() const { ctx_.check(kem_set_integrator(ctx_.get(), name_.c_str())); }
    // batched: out[k] = (E, phi) of element elem[k] at P[k] for unit charge density
    void ElectricFieldAndPotential(const KSurfaceContainer& c, const std::vector<int64_t>& elem,
                                   const std::vector<Position>& P, std::vector<std::pair<FieldVector, double>>& out) const
    {
        Select();
        std::vector<double> buf(4 * elem.size());
        ctx_.check(kem_element_eval(ctx_.get(), c.get(), (int64_t)elem.size(), elem.data(),
                                    reinterpret_cast<const double*>(P.data()), buf.data()));
        out.resize(elem.size());
        for (size_t k = 0; k < elem.size(); k++) out[k] = {{buf[4 * k], buf[4 * k + 1], buf[4 * k + 2]}, buf[4 * k + 3]};
    }
    std::pair<FieldVector, double> ElectricFieldAndPotential(const KSurfaceContainer& c, int64_t j, const Position& P) const
    {
        std::vector<std::pair<FieldVector, double>> out;
        ElectricFieldAndPotential(c, {j}, {P}, out);
        return out[0];
    }
    double Potential(const KSurfaceContainer& c, int64_t j, const Position& P) const
    {
        return ElectricFieldAndPotential(c, j, P).second;
    }
    FieldVector ElectricField(const KSurfaceContainer& c, int64_t j, const Position& P) const
    {
        return ElectricFieldAndPotential(c, j, P).first;
    }

  private:
    const Context& ctx_;
    std::string name_;
};

struct KElectrostaticBoundaryIntegratorFactory {
    static KElectrostaticBoundaryIntegrator Make(const Context& ctx, const std::string& name)
    {
        return KElectrostaticBoundaryIntegrator(ctx, name);
    }
    static KElectrostaticBoundaryIntegrator MakeDefault(const Context& ctx) { return Make(ctx, "default"); }
};

// KBoundaryIntegralMatrix as a KSquareMatrix<double>: device-resident, FP64 and/or FP32 storage
class KBoundaryIntegralMatrix {
  public:
    KBoundaryIntegralMatrix(const KSurfaceContainer& c, const KElectrostaticBoundaryIntegrator& integrator,
                            bool fp64 = true, bool fp32 = false)
        : ctx_(c.context())
    {
        integrator.Select();
        ctx_.check(kem_matrix_assemble(ctx_.get(), c.get(), (fp64 ? KEM_STORE_FP64 : 0) | (fp32 ? KEM_STORE_FP32 : 0),
                                       &h_));
    }
    ~KBoundaryIntegralMatrix() { kem_matrix_destroy(h_); }
    KBoundaryIntegralMatrix(const KBoundaryIntegralMatrix&) = delete;
    KBoundaryIntegralMatrix& operator=(const KBoundaryIntegralMatrix&) = delete;
    unsigned int Dimension() const { return (unsigned int)kem_matrix_dimension(h_); }
    double operator()(unsigned int i, unsigned int j) const
    {
        double v = 0.;
        ctx_.check(kem_matrix_get(h_, i, j, &v));
        return v;
    }
    void Multiply(const std::vector<double>& x, std::vector<double>& y) const
    {
        y.resize(Dimension());
        ctx_.check(kem_matrix_multiply(h_, 0, x.data(), y.data()));
    }
    void MultiplyTranspose(const std::vector<double>& x, std::vector<double>& y) const
    {
        y.resize(Dimension());
        ctx_.check(kem_matrix_multiply(h_, 1, x.data(), y.data()));
    }
    kem_matrix* get() const { return h_; }
    const Context& context() const { return ctx_; }

  private:
    const Context& ctx_;
    kem_matrix* h_ = nullptr;
};

// Charge-density solvers: Solve(A, b, x) with x holding the initial guess (KChargeDensitySolver
// writes sigma into the surfaces; here x is the caller's vector in container order)
class KChargeDensitySolver {
  public:
    virtual ~KChargeDensitySolver() = default;
    kem_solve_stats Solve(const KBoundaryIntegralMatrix& A, const std::vector<double>& b, std::vector<double>& x) const
    {
        x.resize(b.size(), 0.);
        kem_solve_stats st{};
        A.context().check(kem_solve(A.get(), &opts_, b.data(), x.data(), &st));
        return st;
    }

  protected:
    KChargeDensitySolver() { kem_solve_opts_default(&opts_); }
    kem_solve_opts opts_;
};

class KKrylovChargeDensitySolver : public KChargeDensitySolver {
  public:
    KKrylovChargeDensitySolver() { opts_.method = KEM_SOLVER_GMRES; }
    void SetTolerance(double t) { opts_.tolerance = t; }
    void SetMaxIterations(int n) { opts_.max_iterations = n; }
    void SetIterationsBetweenRestart(int n) { opts_.restart = n; }
    void SetTimeLimitSeconds(double s) { opts_.time_limit_s = s; }
    void SetMixedPrecision(bool on) { opts_.precision = on ? KEM_MIXED : KEM_FP64; }
};

class KRobinHoodChargeDensitySolver : public KChargeDensitySolver {
  public:
    KRobinHoodChargeDensitySolver() { opts_.method = KEM_SOLVER_ROBIN_HOOD; }
    void SetTolerance(double t) { opts_.tolerance = t; }
    void SetCheckSubInterval(int64_t n) { opts_.check_interval = n; }
};

class KGaussianEliminationChargeDensitySolver : public KChargeDensitySolver {
  public:
    KGaussianEliminationChargeDensitySolver() { opts_.method = KEM_SOLVER_GAUSS; }
};

// direct field of a solved container (fused, batched)
class KIntegratingElectrostaticFieldSolver {
  public:
    KIntegratingElectrostaticFieldSolver(const KSurfaceContainer& c, std::vector<double> sigma, bool mixed = false)
        : c_(c), sigma_(std::move(sigma)), precision_(mixed ? KEM_MIXED : KEM_FP64)
    {
    }
    void ElectricFieldAndPotential(const std::vector<Position>& P, std::vector<FieldVector>& E,
                                   std::vector<double>& phi) const
    {
        E.resize(P.size());
        phi.resize(P.size());
        c_.context().check(kem_field_eval(c_.context().get(), c_.get(), sigma_.data(), (int64_t)P.size(),
                                          reinterpret_cast<const double*>(P.data()), precision_, phi.data(),
                                          reinterpret_cast<double*>(E.data())));
    }
    std::pair<FieldVector, double> ElectricFieldAndPotential(const Position& P) const
    {
        std::vector<FieldVector> E;
        std::vector<double> phi;
        ElectricFieldAndPotential({P}, E, phi);
        return {E[0], phi[0]};
    }
    double Potential(const Position& P) const { return ElectricFieldAndPotential(P).second; }
    FieldVector ElectricField(const Position& P) const { return ElectricFieldAndPotential(P).first; }
    const std::vector<double>& ChargeDensities() const { return sigma_; }

  private:
    const KSurfaceContainer& c_;
    std::vector<double> sigma_;
    int precision_;
};

// KElectrostaticBoundaryField: Initialize() solves for the charge densities, then the field
// solver answers Potential / ElectricField / ElectricFieldAndPotential (the time argument is
// accepted and ignored, as for the static field of the reference)
class KElectrostaticBoundaryField {
  public:
    KElectrostaticBoundaryField(const KSurfaceContainer& c, KElectrostaticBoundaryIntegrator integrator,
                                const KChargeDensitySolver& solver, bool mixed_field = false)
        : c_(c), integrator_(std::move(integrator)), solver_(solver), mixed_(mixed_field)
    {
    }
    kem_solve_stats Initialize()
    {
        KBoundaryIntegralMatrix A(c_, integrator_);
        std::vector<double> x;
        const kem_solve_stats st = solver_.Solve(A, c_.BoundaryValues(), x);
        field_.reset(new KIntegratingElectrostaticFieldSolver(c_, std::move(x), mixed_));
        return st;
    }
    double Potential(const Position& P, double /*t*/ = 0.) const { return solver().Potential(P); }
    FieldVector ElectricField(const Position& P, double /*t*/ = 0.) const { return solver().ElectricField(P); }
    std::pair<FieldVector, double> ElectricFieldAndPotential(const Position& P, double /*t*/ = 0.) const
    {
        return solver().ElectricFieldAndPotential(P);
    }
    // many points per call (the batched entry a tracking step can use)
    void ElectricFieldAndPotential(const std::vector<Position>& P, std::vector<FieldVector>& E,
                                   std::vector<double>& phi) const
    {
        solver().ElectricFieldAndPotential(P, E, phi);
    }
    const std::vector<double>& ChargeDensities() const { return solver().ChargeDensities(); }
    const KSurfaceContainer& Container() const { return c_; }
    bool MixedPrecision() const { return mixed_; }

  private:
    const KIntegratingElectrostaticFieldSolver& solver() const
    {
        if (!field_) throw Error(KEM_E_STATE, "KElectrostaticBoundaryField: Initialize() first");
        return *field_;
    }
    const KSurfaceContainer& c_;
    KElectrostaticBoundaryIntegrator integrator_;
    const KChargeDensitySolver& solver_;
    bool mixed_;
    std::unique_ptr<KIntegratingElectrostaticFieldSolver> field_;
};

// KSElectricKEMField (Kassiopeia/Fields/Source/KSElectricKEMField.cxx:44-72): Kassiopeia's
// electric-field component over a KElectrostaticBoundaryField.  InitializeComponent() solves
// (KElectrostaticBoundaryField::Initialize) and keeps the charge densities on the device; the
// point and result buffers stay on the device between calls and grow to the largest batch, so
// a tracking step that collects the positions of all its particles evaluates them in one call
// (CalculateFieldAndPotential over a vector) instead of one launch and round trip per particle.
// The time argument is accepted and ignored (static field).
class KSElectricKEMField {
  public:
    explicit KSElectricKEMField(KElectrostaticBoundaryField& field) : field_(field) {}
    ~KSElectricKEMField() { release(); }
    KSElectricKEMField(const KSElectricKEMField&) = delete;
    KSElectricKEMField& operator=(const KSElectricKEMField&) = delete;

    void InitializeComponent()
    {
        field_.Initialize();
        release();
        const std::vector<double>& s = field_.ChargeDensities();
        sigma_ = alloc((int64_t)s.size());
        ctx().check(kem_dev_upload(ctx().get(), sigma_, s.data(), (int64_t)(sizeof(double) * s.size())));
    }
    void DeinitializeComponent() { release(); }

    void CalculatePotential(const Position& P, const double& /*t*/, double& phi)
    {
        FieldVector E;
        CalculateFieldAndPotential(P, 0., E, phi);
    }
    void CalculateField(const Position& P, const double& /*t*/, FieldVector& E)
    {
        double phi;
        CalculateFieldAndPotential(P, 0., E, phi);
    }
    void CalculateFieldAndPotential(const Position& P, const double& /*t*/, FieldVector& E, double& phi)
    {
        std::vector<FieldVector> Ev;
        std::vector<double> pv;
        CalculateFieldAndPotential(std::vector<Position>{P}, 0., Ev, pv);
        E = Ev[0];
        phi = pv[0];
    }
    // all sample points of one step in one evaluation
    void CalculateFieldAndPotential(const std::vector<Position>& P, const double& /*t*/, std::vector<FieldVector>& E,
                                    std::vector<double>& phi)
    {
        if (!sigma_) throw Error(KEM_E_STATE, "KSElectricKEMField: InitializeComponent() first");
        const int64_t n = (int64_t)P.size();
        E.resize(P.size());
        phi.resize(P.size());
        if (n == 0) return;
        reserve(n);
        host_.resize((size_t)(4 * n));
        for (int64_t k = 0; k < n; k++)
            for (int c = 0; c < 3; c++) host_[(size_t)(c * n + k)] = P[(size_t)k][(size_t)c];
        const int64_t bytes = (int64_t)sizeof(double) * n;
        for (int c = 0; c < 3; c++)
            ctx().check(kem_dev_upload(ctx().get(), x_[c], host_.data() + c * n, bytes));
        ctx().check(kem_field_eval_device(ctx().get(), field_.Container().get(), sigma_, n, x_[0], x_[1], x_[2],
                                          field_.MixedPrecision() ? KEM_MIXED : KEM_FP64, out_[3], out_[0], out_[1],
                                          out_[2]));
        for (int c = 0; c < 4; c++) ctx().check(kem_dev_download(ctx().get(), host_.data() + c * n, out_[c], bytes));
        for (int64_t k = 0; k < n; k++) {
            for (int c = 0; c < 3; c++) E[(size_t)k][(size_t)c] = host_[(size_t)(c * n + k)];
            phi[(size_t)k] = host_[(size_t)(3 * n + k)];
        }
    }

  private:
    friend class KSRootElectricField;
    const Context& ctx() const { return field_.Container().context(); }
    double* alloc(int64_t count)
    {
        void* p = nullptr;
        ctx().check(kem_dev_alloc(ctx().get(), (int64_t)sizeof(double) * count, &p));
        return (double*)p;
    }
    void reserve(int64_t n)
    {
        if (n <= cap_) return;
        for (double*& p : x_) (void)kem_dev_free(ctx().get(), p), p = nullptr;
        for (double*& p : out_) (void)kem_dev_free(ctx().get(), p), p = nullptr;
        cap_ = 0;
        for (double*& p : x_) p = alloc(n);
        for (double*& p : out_) p = alloc(n);
        cap_ = n;
    }
    void release()
    {
        if (sigma_) (void)kem_dev_free(ctx().get(), sigma_);
        sigma_ = nullptr;
        for (double*& p : x_)
            if (p) (void)kem_dev_free(ctx().get(), p), p = nullptr;
        for (double*& p : out_)
            if (p) (void)kem_dev_free(ctx().get(), p), p = nullptr;
        cap_ = 0;
    }
    KElectrostaticBoundaryField& field_;
    double* sigma_ = nullptr;
    double* x_[3] = {nullptr, nullptr, nullptr};
    double* out_[4] = {nullptr, nullptr, nullptr, nullptr};  // Ex, Ey, Ez, phi
    int64_t cap_ = 0;
    std::vector<double> host_;
};

// KSRootElectricField (Kassiopeia/Simulation/Source/KSRootElectricField.cxx:104-121): the sum
// of several electric-field components at the sample points of a step, added per point in the
// order the components were added, from zero.  Here a step's sample points are evaluated in
// one call for all components (kem_field_eval_fields: one pass per component, summed on the
// device in the same order).  Components must share one context and be initialized.
class KSRootElectricField {
  public:
    KSRootElectricField() = default;
    ~KSRootElectricField() { release(); }
    KSRootElectricField(const KSRootElectricField&) = delete;
    KSRootElectricField& operator=(const KSRootElectricField&) = delete;

    void AddElectricField(KSElectricKEMField* f) { fields_.push_back(f); }
    void RemoveElectricField(KSElectricKEMField* f)
    {
        for (size_t k = 0; k < fields_.size(); k++)
            if (fields_[k] == f) {
                fields_.erase(fields_.begin() + (long)k);
                return;
            }
        throw Error(KEM_E_ARG, "KSRootElectricField: field not found");
    }

    void CalculateFieldAndPotential(const Position& P, const double& t, FieldVector& E, double& phi)
    {
        std::vector<FieldVector> Ev;
        std::vector<double> pv;
        CalculateFieldAndPotential(std::vector<Position>{P}, t, Ev, pv);
        E = Ev[0];
        phi = pv[0];
    }
    void CalculateFieldAndPotential(const std::vector<Position>& P, const double& /*t*/, std::vector<FieldVector>& E,
                                    std::vector<double>& phi)
    {
        const int64_t n = (int64_t)P.size();
        E.assign(P.size(), FieldVector{0., 0., 0.});
        phi.assign(P.size(), 0.);
        if (n == 0 || fields_.empty()) return;
        const Context& c = fields_[0]->ctx();
        std::vector<const kem_mesh*> meshes;
        std::vector<const double*> sigmas;
        std::vector<int> prec;
        for (KSElectricKEMField* f : fields_) {
            if (!f->sigma_) throw Error(KEM_E_STATE, "KSRootElectricField: a component is not initialized");
            if (f->ctx().get() != c.get()) throw Error(KEM_E_STATE, "KSRootElectricField: components on two contexts");
            meshes.push_back(f->field_.Container().get());
            sigmas.push_back(f->sigma_);
            prec.push_back(f->field_.MixedPrecision() ? KEM_MIXED : KEM_FP64);
        }
        reserve(c, n);
        host_.resize((size_t)(4 * n));
        for (int64_t k = 0; k < n; k++)
            for (int q = 0; q < 3; q++) host_[(size_t)(q * n + k)] = P[(size_t)k][(size_t)q];
        const int64_t bytes = (int64_t)sizeof(double) * n;
        for (int q = 0; q < 3; q++) c.check(kem_dev_upload(c.get(), buf_[q], host_.data() + q * n, bytes));
        c.check(kem_field_eval_fields(c.get(), (int64_t)fields_.size(), meshes.data(), sigmas.data(), prec.data(), n,
                                      buf_[0], buf_[1], buf_[2], buf_[6], buf_[3], buf_[4], buf_[5]));
        for (int q = 0; q < 4; q++) c.check(kem_dev_download(c.get(), host_.data() + q * n, buf_[3 + q], bytes));
        for (int64_t k = 0; k < n; k++) {
            for (int q = 0; q < 3; q++) E[(size_t)k][(size_t)q] = host_[(size_t)(q * n + k)];
            phi[(size_t)k] = host_[(size_t)(3 * n + k)];
        }
    }

  private:
    void reserve(const Context& c, int64_t n)
    {
        if (n <= cap_ && ctx_ == c.get()) return;
        release();
        ctx_ = c.get();
        for (double*& p : buf_) {
            void* v = nullptr;
            c.check(kem_dev_alloc(ctx_, (int64_t)sizeof(double) * n, &v));
            p = (double*)v;
        }
        cap_ = n;
    }
    void release()
    {
        for (double*& p : buf_)
            if (p && ctx_) (void)kem_dev_free(ctx_, p), p = nullptr;
        cap_ = 0;
    }
    std::vector<KSElectricKEMField*> fields_;
    kem_ctx* ctx_ = nullptr;
    double* buf_[7] = {};  // x, y, z, Ex, Ey, Ez, phi
    int64_t cap_ = 0;
    std::vector<double> host_;
};

}  // namespace kemhip

#endif
