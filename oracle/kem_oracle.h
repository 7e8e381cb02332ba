/*
 * kem_oracle.h -- CPU FP64 restatement of KEMField's electrostatic boundary-element path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product path
 * (kassiopeia_amd/csrc).  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product never links, calls or falls back to it.
 *
 * Every function restates the algorithm of the reference file:line cited beside it,
 * keeping the reference's branch structure, constants and (where it matters for the
 * last bits) its order of operations.  Paths are relative to the KEMField source tree
 * of KATRIN-Experiment/Kassiopeia.
 *
 * Element records ("shape data"):
 *   triangle, rectangle: 11 doubles [A, B, P0x, P0y, P0z, N1x, N1y, N1z, N2x, N2y, N2z]
 *     (the layout of KElectrostaticCubatureTriangleIntegrator.cc:457-467)
 *   line segment:        7 doubles [P0x, P0y, P0z, P1x, P1y, P1z, diameter]
 *     (KElectrostaticQuadratureLineSegmentIntegrator.cc:197-203)
 *   conic section:       6 doubles [P0x, P0y, P0z, P1x, P1y, P1z] (radius = x, axis = z;
 *     Surfaces/src/KConicSection.cc:7-22)
 *   ring:                3 doubles [Px, Py, Pz] (radius = x, axis = z; Surfaces/src/KRing.cc:8-18)
 *   Records are stored with a stride of KO_REC doubles.
 *
 * Outputs of a single (element, point) evaluation are 4 doubles [Ex, Ey, Ez, phi] for a
 * unit surface charge density (sigma = 1 C/m^2; line: unit line charge per diameter as
 * in the reference).
 */
#ifndef KEM_ORACLE_H
#define KEM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KO_REC 11

enum ko_shape { KO_TRIANGLE = 0, KO_RECTANGLE = 1, KO_LINESEGMENT = 2, KO_CONICSECTION = 3, KO_RING = 4 };
enum ko_policy { KO_NUMERIC = 0, KO_RWG = 1, KO_ANALYTIC = 2, KO_REFERENCE = 3 };
enum ko_bc { KO_DIRICHLET = 0, KO_NEUMANN = 1 };

/* constants: KEMField/Source/Core/src/KEMConstants.cc:9-12 over Kommon KConst_2021.h:37,139-146 */
double ko_pi(void);
double ko_eps0(void);
double ko_one_over_4pi_eps0(void);

/* KTriangle::SetValues(p0,p1,p2)  Surfaces/src/KTriangle.cc:18-28 (rectangle: KRectangle::SetValues same rule) */
void ko_set_from_points(const double p0[3], const double p1[3], const double p2[3], double rec[KO_REC]);
/* Area/Centroid: Surfaces/include/KTriangle.hh:27-34, KRectangle.hh:27-34, KLineSegment */
double ko_area(int shape, const double* rec);
void ko_centroid(int shape, const double* rec, double c[3]);
void ko_normal(int shape, const double* rec, double n[3]);

/* triangle cubature nodes, n in {7,12,33}: Numeric/src/KElectrostaticCubatureTriangleIntegrator.cc:34,70,212 */
int ko_tri_nodes(int n, const double* rec, double* Q, double* w);
/* rectangle cubature nodes, n in {7,12,33}: Numeric/src/KElectrostaticCubatureRectangleIntegrator.cc:43,120,235 */
int ko_rect_nodes(int n, const double* rec, double* Q, double* w);
/* fixed-rule cubature for (E, phi): ..._TriNP :387 / ..._RectNP */
void ko_tri_cubature(int n, const double* rec, const double P[3], double out[4]);
void ko_rect_cubature(int n, const double* rec, const double P[3], double out[4]);

/* Van Oosterom-Strackee solid angle: Surfaces/include/KSolidAngle.hh:58-144 (triangle), :166-305 (rectangle) */
double ko_solid_angle_tri(const double* rec, const double P[3]);
double ko_solid_angle_rect(const double* rec, const double P[3]);

/* closed-form near field: RWG/src/KElectrostaticRWGTriangleIntegrator.cc:1013-1240,
   RWG/src/KElectrostaticRWGRectangleIntegrator.cc:1175-1430 */
void ko_tri_rwg(const double* rec, const double P[3], double out[4]);
void ko_rect_rwg(const double* rec, const double P[3], double out[4]);

/* line segment: Numeric/src/KElectrostaticQuadratureLineSegmentIntegrator.cc,
   Analytic/src/KElectrostaticAnalyticLineSegmentIntegrator.cc:15-83 */
void ko_line_quadrature(int nnodes, const double* rec, const double P[3], double out[4]);
void ko_line_analytic(const double* rec, const double P[3], double out[4]);

/* "analytic" policy: Analytic/src/KElectrostaticAnalyticTriangleIntegrator.cc:13-174 (helpers
   :176-506), Analytic/src/KElectrostaticAnalyticRectangleIntegrator.cc:25-205 */
void ko_tri_analytic(const double* rec, const double P[3], double out[4]);
void ko_rect_analytic(const double* rec, const double P[3], double out[4]);
/* "reference" policy: Reference/src/KElectrostaticBiQuadratureTriangleIntegrator.cc:24-150,
   ...BiQuadratureRectangleIntegrator.cc:18-104, ...256NodeQuadratureLineSegmentIntegrator.cc:192-270 */
void ko_tri_biquad(const double* rec, const double P[3], double out[4]);
void ko_rect_biquad(const double* rec, const double P[3], double out[4]);
void ko_line_256(const double* rec, const double P[3], double out[4]);
/* the "numeric" dispatch (default policy): cubature 7/12/33 vs closed form by distance ratio.
   Returns the regime used: 7, 12, 33 (cubature), 0 (closed form), 4/16 for line nodes. */
int ko_tri_numeric(const double* rec, const double P[3], double out[4]);
int ko_rect_numeric(const double* rec, const double P[3], double out[4]);
int ko_line_numeric(const double* rec, const double P[3], double out[4]);

/* axisymmetric elements (every policy): Analytic/src/KElectrostaticAnalyticRingIntegrator.cc,
   ...ConicSectionIntegrator.cc; the complete elliptic integrals K, E and (E - K)/k^2 as the
   reference evaluates them (Math/src/KEllipticIntegrals.cc) */
double ko_ell_K(double k);
double ko_ell_E(double k);
double ko_ell_EK(double k);
void ko_ring(const double* rec, const double P[3], double out[4]);
void ko_conic(const double* rec, const double P[3], double out[4]);

/* symmetry groups (KSymmetryGroup<Shape>): n groups, members [gs[g], gs[g+1]) of the records;
   a group's integral sums its members, its target is its first member */
void ko_assemble_groups(int policy, int64_t n, const int64_t* gs, const int32_t* shape, const double* recs,
                        const int32_t* bc, const double* flux, double* A);
void ko_field_groups(int policy, int64_t n, const int64_t* gs, const int32_t* shape, const double* recs,
                     const double* sigma, int64_t m, const double* P, double* phi, double* E);

/* policy dispatch over shapes: KElectrostaticBoundaryIntegratorFactory.cc:29-88 */
int ko_element_eval(int policy, int shape, const double* rec, const double P[3], double out[4]);

/* Collocation: Core/include/KElectrostaticBoundaryIntegrator.hh:260-280, Core/src/...cc:49-62.
   A is written row-major: A[i*n + j] = integral of source j at target i. */
void ko_assemble(int policy, int64_t n, const int32_t* shape, const double* recs, const int32_t* bc,
                 const double* flux, double* A);
double ko_matrix_element(int policy, int64_t i, int64_t j, const int32_t* shape, const double* recs,
                         const int32_t* bc, const double* flux);
/* boundary values b_i: KBoundaryIntegralVector.hh:60-75 */
void ko_rhs(int64_t n, const int32_t* bc, const double* value, double* b);

/* Kahan-summed direct field: FieldSolvers/Integrating/include/KElectrostaticIntegratingFieldSolver.hh:303-339.
   Points are AoS (m x 3); outputs phi[m], E[m*3]. */
void ko_field(int policy, int64_t n, const int32_t* shape, const double* recs, const double* sigma, int64_t m,
              const double* P, double* phi, double* E);
/* the same over nthreads OpenMP threads (points shared out; every point's sum as above) */
void ko_field_mt(int nthreads, int policy, int64_t n, const int32_t* shape, const double* recs, const double* sigma,
                 int64_t m, const double* P, double* phi, double* E);

/* GMRES (unrestarted unless restart>0), MGS + Givens, relative tolerance ||r||/||b||:
   LinearAlgebra/Solvers/include/KGeneralizedMinimalResidual.hh:178-337,
   KSimpleIterativeKrylovSolver.hh:140-194.  A row-major.  Returns iterations; *resid = final |p_{j+1}|/||b||. */
int ko_gmres(int64_t n, const double* A, const double* b, double* x, double tol, int maxit, int restart,
             double* resid, double* hist);
/* same with the orthogonalisation selectable: 0 = MGS (reference), 1 = CGS2 (the HIP path's) */
int ko_gmres_ortho(int64_t n, const double* A, const double* b, double* x, double tol, int maxit, int restart,
                   int ortho, double* resid, double* hist);
/* Robin Hood: KRobinHood.hh:62-104, KRobinHood_SingleThread.hh:53-120.  Returns sub-iterations. */
int64_t ko_robin_hood(int64_t n, const double* A, const double* b, double* x, double tol, int64_t check_interval,
                      int64_t max_sub, double* resid);
/* Gauss-Jordan without pivoting: LinearAlgebra/Solvers/include/KGaussianElimination.hh:24-58 */
int ko_gauss(int64_t n, const double* A, const double* b, double* x);

/* zonal-harmonic field solver of rings and conic sections (kem_oracle_zh.c): build over n
   elements (arrays must outlive the handle), evaluate (E, phi) at m points (out m x 4) with
   counts[3] = central, remote and direct evaluations, tree sizes, a central source point */
void* ko_zh_build(int64_t n, const int32_t* shape, const double* recs, const double* sigma, const double* params);
void ko_zh_eval(void* h, int64_t m, const double* P, double* out, int64_t* counts);
void ko_zh_info(void* h, int64_t* out);
int ko_zh_central_sp(void* h, int k, double* out);
void ko_zh_free(void* h);

#ifdef __cplusplus
}
#endif
#endif
