/*
 * kemhip.h -- C-ABI of the MI355X (gfx950) electrostatic boundary-element library.
 *
 * The drop-in boundary for KEMField's electrostatic BEM hot path.  Plain pointers and sizes
 * only; no exceptions cross this boundary: every call returns KEM_OK (0), a positive warning
 * or a negative KEM_E_* code, and kem_last_error() holds the message.  Each context owns (or
 * borrows) one HIP stream; calls on one context are not thread-safe, separate contexts are
 * independent.
 *
 * Each entry point names the KEMField interface it replaces (paths relative to
 * KEMField/Source of KATRIN-Experiment/Kassiopeia).
 *
 * Element records ("shape data", 11 doubles each, KSurfaceContainer global order):
 *   triangle / rectangle  [A, B, P0x, P0y, P0z, N1x, N1y, N1z, N2x, N2y, N2z]
 *       (KTriangle / KRectangle; layout of Numeric/src/KElectrostaticCubatureTriangleIntegrator.cc:457-467)
 *   line segment          [P0x, P0y, P0z, P1x, P1y, P1z, diameter, 0, 0, 0, 0]
 *       (KLineSegment; Numeric/src/KElectrostaticQuadratureLineSegmentIntegrator.cc:197-203)
 * Shape types: 0 triangle, 1 rectangle, 2 line segment.  Boundary types: 0 Dirichlet,
 * 1 Neumann (KBoundary.hh:82-264).
 */
#ifndef KEMHIP_H
#define KEMHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes */
#define KEM_OK 0
#define KEM_W_NOT_CONVERGED 1 /* iteration or time limit reached; x holds the best iterate */
#define KEM_E_ARG (-1)        /* invalid argument (null pointer, size, index) */
#define KEM_E_HIP (-2)        /* HIP runtime error (allocation, launch, no device) */
#define KEM_E_NAME (-3)       /* unknown integrator / solver name (KEMSimpleException) */
#define KEM_E_GEOM (-4)       /* degenerate or non-finite element */
#define KEM_E_STATE (-5)      /* object used with the wrong context / missing storage */

/* ---- integrator policies (KElectrostaticBoundaryIntegratorFactory.cc:29-88) */
#define KEM_POLICY_NUMERIC 0 /* default: cubature 7/12/33 by distance ratio, closed form near */
#define KEM_POLICY_RWG 1     /* closed form at every distance (segments: analytic) */
#define KEM_POLICY_ANALYTIC 2  /* Analytic triangle/rectangle/segment integrators */
#define KEM_POLICY_REFERENCE 3 /* nested 32-node Gauss-Legendre (segments: 8 x 32 nodes) */

/* ---- precision modes */
#define KEM_FP64 0  /* plain path: FP64 everywhere */
#define KEM_MIXED 1 /* FP32 far-field cubature / FP32 matrix; FP64 near field, refinement, sums */
#define KEM_MIXED_BF16 2 /* solves and block products only: BF16 matrix on BF16 MFMA (FP32 accumulation),
                            FP64 refinement */

/* ---- matrix storage flags for kem_matrix_assemble */
/* shape types and boundary types of kem_mesh_create (KTriangle / KRectangle / KLineSegment /
   KConicSection / KRing; KDirichletBoundary / KNeumannBoundary) */
#define KEM_SHAPE_TRIANGLE 0
#define KEM_SHAPE_RECTANGLE 1
#define KEM_SHAPE_LINESEGMENT 2
#define KEM_SHAPE_CONICSECTION 3 /* [P0x,P0y,P0z,P1x,P1y,P1z]: radius = x, axial position = z */
#define KEM_SHAPE_RING 4         /* [Px,Py,Pz]: radius = x > 0, axial position = z */
#define KEM_BC_DIRICHLET 0
#define KEM_BC_NEUMANN 1

#define KEM_STORE_FP64 1
#define KEM_STORE_FP32 2
#define KEM_STORE_BF16 4 /* entries rounded to FP32, then BF16 (nearest even); packed for the MFMA operand */

/* ---- solvers (KKrylovSolverFactory.hh:40-199, KRobinHood.hh, KGaussianElimination.hh) */
#define KEM_SOLVER_GMRES 0
#define KEM_SOLVER_ROBIN_HOOD 1
#define KEM_SOLVER_GAUSS 2 /* Gauss-Jordan without pivoting, KGaussianElimination.hh:24-58 (n <= 32768) */

typedef struct kem_ctx kem_ctx;
typedef struct kem_graph kem_graph;
typedef struct kem_mesh kem_mesh;
typedef struct kem_matrix kem_matrix;
typedef struct kem_zh kem_zh;

typedef struct kem_solve_opts {
    int method;             /* KEM_SOLVER_* */
    int precision;          /* KEM_FP64, or KEM_MIXED: FP32-matrix inner solves + FP64 refinement, or
                               KEM_MIXED_BF16: BF16-matrix inner solves (GMRES) + FP64 refinement */
    double tolerance;       /* relative: ||r||/||b|| (GMRES), ||r||_inf/||b||_inf (Robin Hood) */
    int max_iterations;     /* GMRES iterations (all restarts), Robin Hood norm checks */
    int restart;            /* GMRES iterations between restarts; 0 = never */
    int64_t check_interval; /* Robin Hood sub-iterations between norm checks; 0 = N */
    double time_limit_s;    /* wall-clock limit (KTimeTerminator); <= 0 = none */
    double inner_tolerance; /* KEM_MIXED: relative tolerance of each FP32 inner solve */
} kem_solve_opts;

typedef struct kem_solve_stats {
    int iterations;         /* GMRES iterations (sum over refinement steps), RH norm checks */
    int64_t sub_iterations; /* Robin Hood element corrections */
    int refinements;        /* KEM_MIXED: FP64 refinement steps */
    double residual;        /* the solver's convergence measure at exit */
    double true_residual;   /* ||b - A x|| / ||b|| recomputed with the FP64 matrix (if stored) */
    double seconds;         /* wall time of the solve */
} kem_solve_stats;

/* ---- library / context */
int kem_version(void);
/* create a context on a HIP device with a stream of its own (default priority) */
int kem_ctx_create(int device, kem_ctx** out);
void kem_ctx_destroy(kem_ctx* ctx);
/* run the context's work on a caller-owned HIP stream (hipStream_t); NULL restores its own */
int kem_ctx_set_stream(kem_ctx* ctx, void* stream);
void* kem_ctx_stream(kem_ctx* ctx);
int kem_ctx_sync(kem_ctx* ctx);
const char* kem_last_error(const kem_ctx* ctx);
/* KEM_MIXED field evaluation: far pairs with distance ratio <= ratio keep FP64 cubature, the
   rest run in FP32 (default 0: every far pair in FP32).  No reference counterpart: the
   reference has no mixed-precision mode. */
int kem_ctx_set_mixed_fp64_ratio(kem_ctx* ctx, double ratio);
/* Launch tuning (no reference counterpart; kassiopeia_amd/autotune.py measures and sets them):
   KEM_TUNE_FAR_RANGE_FP64 / _MIXED = elements per element range of the far pass over 262,144
   or more points of one charge-density set; the elements split into 8 R ranges, R per XCD, so
   each XCD's range stays in its L2 (0: no split; default 2560).  KEM_TUNE_GEMV_BLOCKS /
   KEM_TUNE_BMM_BLOCKS = target block counts of the split-K matrix-vector product (default 2048)
   and of the register-loaded multi-RHS block products in 64-row tiles times column splits
   (FP32 with n % 4 != 0, FP64 with n odd; default 2560);
   they set how many partial sums are reduced. */
#define KEM_TUNE_FAR_RANGE_FP64 1
#define KEM_TUNE_FAR_RANGE_MIXED 2
#define KEM_TUNE_GEMV_BLOCKS 3
#define KEM_TUNE_BMM_BLOCKS 4
#define KEM_TUNE_NEAR_BLOCKS 5 /* blocks of the near pass beside a single-set mixed far pass (0: one per 256 points; default 384) */
/* wavefront tile of the register-loaded FP32 MFMA block product: 64 rows (one float per lane
   and column), or 128 rows (four consecutive rows per lane in one 16-byte load; n % 4 == 0,
   which the streamed kernels take whenever the buffers are 16-byte aligned); default 64 */
#define KEM_TUNE_BMM_ROWS 6
/* Robin Hood sub-iterations: 0 = one workgroup up to 8,192 unknowns and several co-resident
   workgroups above (default), 1 = always one workgroup (the same corrections either way) */
#define KEM_TUNE_ROBIN_HOOD_GROUPS 7
/* target workgroup count of the block products of 33-64 right-hand sides that stream the
   matrix through LDS (FP32 matrices with n % 4 == 0, and BF16): 256 rows per workgroup times
   column splits; default 480 */
#define KEM_TUNE_BMM_STREAM_BLOCKS 8
/* largest workgroup count of the streamed block products of up to 32 right-hand sides (16 x
   16 MFMA tiles: FP32 matrices with n % 4 == 0, FP64 matrices with n even, also for 33-64; one
   32-column tile: BF16 matrices): 256 rows per workgroup times column splits, the split count
   within it chosen by a traffic model over rounds of two workgroups per compute unit; default
   4096 */
#define KEM_TUNE_BMM_TILE_BLOCKS 9
int kem_ctx_set_tuning(kem_ctx* ctx, int param, int64_t value);
int kem_ctx_get_tuning(const kem_ctx* ctx, int param, int64_t* value);
/* Instrumentation, no reference counterpart: with timing on, each field evaluation records HIP
   events on the context's stream around its far-field pass, each kem_matrix_assemble around
   its assembly kernels, each kem_matrix_multiply around its matrix-vector kernels and each
   64-column block of kem_matrix_multiply_batch around its block product (up to 256 timed
   regions between queries); kem_ctx_kernel_time waits for them
   and returns the summed time and the number of regions timed, then starts a new count. */
int kem_ctx_set_kernel_timing(kem_ctx* ctx, int on);
int kem_ctx_kernel_time(kem_ctx* ctx, double* total_ms, int* launches);
/* Device memory on the context's GPU for callers without an allocator of their own (no
   reference counterpart): allocate / free (waits for the stream first), and copies that wait
   for the stream to finish. */
int kem_dev_alloc(kem_ctx* ctx, int64_t bytes, void** ptr);
int kem_dev_free(kem_ctx* ctx, void* ptr);
int kem_dev_upload(kem_ctx* ctx, void* dst, const void* src, int64_t bytes);
int kem_dev_download(kem_ctx* ctx, void* dst, const void* src, int64_t bytes);
/* integrator policy by name: "numeric" | "default" | "analytic" | "rwg" | "reference"
   (KElectrostaticBoundaryIntegratorFactory::Make(name), .cc:78-88; unknown -> KEM_E_NAME) */
int kem_set_integrator(kem_ctx* ctx, const char* name);

/* ---- surfaces: KSurfaceContainer of KSurface<KElectrostaticBasis, boundary, shape>
   (Surfaces/src/KSurfaceContainer.cc:36-92).  Host arrays are copied.  bc_value: Dirichlet
   voltage (KDirichletBoundary); bc_flux: Neumann normal boundary flux (KNeumannBoundary).
   bc_type / bc_value / bc_flux may be NULL (all Dirichlet at 0 V). */
int kem_mesh_create(kem_ctx* ctx, int64_t n, const int32_t* shape_type, const double* shape_data,
                    const int32_t* bc_type, const double* bc_value, const double* bc_flux, kem_mesh** out);
/* Symmetry groups (KSymmetryGroup<Shape>, Surfaces/include/KSymmetryGroup.hh): ngroups
   unknowns over group_start[ngroups] elements; group g holds the elements [group_start[g],
   group_start[g+1]) of one shape type (its members, e.g. the reflections and rotations the
   reference generated) and one boundary condition, bc_type / bc_value / bc_flux per group.  A
   group's integral sums its members in order (Core/include/KElectrostaticElementIntegrator.hh:
   52-79); its collocation point, normal and area are its first member's.  Matrices, solves and
   fields then run over ngroups unknowns; element indices of kem_element_eval name groups.
   The derivative entry points reject grouped meshes. */
int kem_mesh_create_groups(kem_ctx* ctx, int64_t ngroups, const int64_t* group_start, const int32_t* shape_type,
                           const double* shape_data, const int32_t* bc_type, const double* bc_value,
                           const double* bc_flux, kem_mesh** out);
void kem_mesh_destroy(kem_mesh* mesh);
int64_t kem_mesh_size(const kem_mesh* mesh);
/* boundary values b_i (KBoundaryIntegralVector::operator(), BoundaryIntegrals/Core/include/
   KBoundaryIntegralVector.hh:60-75): V_i for Dirichlet rows, 0 for Neumann rows */
int kem_mesh_rhs(const kem_mesh* mesh, double* b);
/* replace the Dirichlet values (e.g. another electrode voltage set) */
int kem_mesh_set_values(kem_mesh* mesh, const double* bc_value);

/* ---- element integrals: KElectrostaticBoundaryIntegrator::ElectricFieldAndPotential(shape, P)
   (Electrostatic/Core/src/KElectrostaticBoundaryIntegrator.cc:101-312) for unit charge density.
   Batched: pair k evaluates element elem[k] at point P[3k..3k+2] (host arrays);
   out[4k..4k+3] = (Ex, Ey, Ez, phi). */
int kem_element_eval(kem_ctx* ctx, const kem_mesh* mesh, int64_t npairs, const int64_t* elem, const double* P,
                     double* out);

/* ---- direct field: KIntegratingFieldSolver<KElectrostaticBoundaryIntegrator, Kind>::
   ElectricFieldAndPotential (FieldSolvers/Integrating/include/KElectrostaticIntegratingFieldSolver.hh:303-339),
   fused (one pass) and batched over points.  Host arrays, synchronous: P is npts x 3 (AoS),
   sigma has mesh-size entries; phi (npts) and E (npts x 3) may be NULL to skip. */
int kem_field_eval(kem_ctx* ctx, const kem_mesh* mesh, const double* sigma, int64_t npts, const double* P,
                   int precision, double* phi, double* E);
/* Same on device arrays (SoA point streams), asynchronous on the context's stream.
   precision KEM_MIXED: far-field cubature in FP32 (node positions relative to the element
   centre, point-to-centre offset rounded once from FP64, accuracy-targeted rule ladder),
   closed-form near field and all sums over elements in FP64. */
int kem_field_eval_device(kem_ctx* ctx, const kem_mesh* mesh, const double* d_sigma, int64_t npts,
                          const double* d_x, const double* d_y, const double* d_z, int precision, double* d_phi,
                          double* d_ex, double* d_ey, double* d_ez);
/* Field of several charge-density sets in one pass (right-hand sides of kem_solve_batch, or
   uneven point batches each with its own solution): point i uses sigma[set[i]*n + j]
   (device arrays: sigma n x nsets column-major, set[npts] in [0, nsets), nsets <= 64).  No
   reference counterpart (one KIntegratingFieldSolver per solution there).  Validates the set
   indices (waits for the stream), then runs asynchronously like kem_field_eval_device. */
/* No-allocation variant of kem_field_eval_device, with the same arguments and results (the
   same kernels, element ranges and sums, so bit-equal outputs).  No counterpart in the
   reference: it is for callers that cannot allocate device memory, such as code under HIP
   graph capture or a tracking loop that forbids hipMalloc.  It calls no hipMalloc / hipFree:
   its partial sums, element weights and point ordering live in a static 48 MiB device scratch
   array, one per GPU, held until the process ends.  At most 65,536 points per call, and the
   point ordering, group charge densities and the workspace the eager path would allocate
   must fit the 48 MiB (about 1.5 M elements at 65,536 points); larger calls are refused with
   KEM_E_ARG before any device work.  All calls on one GPU share its scratch: calls that
   overlap on one GPU (several contexts or streams of one device) race, and must be
   serialised by the caller.  Asynchronous on the context's stream. */
int kem_field_eval_device_noalloc(kem_ctx* ctx, const kem_mesh* mesh, const double* d_sigma, int64_t npts,
                                  const double* d_x, const double* d_y, const double* d_z, int precision,
                                  double* d_phi, double* d_ex, double* d_ey, double* d_ez);
/* A field evaluation on fixed device buffers captured once as a HIP graph and replayed with
   kem_graph_launch (asynchronous on the context's stream): for repeated evaluations at points
   that change in place, e.g. the particles of a tracking step.  The buffers, sigma and the
   point count are fixed at capture; their contents may change between launches.  No
   reference counterpart.  The graph owns the workspace its kernels address (allocated here),
   so eager calls on the context between replays are allowed; it must be destroyed before its
   context and mesh.  phi, ex, ey, ez: all four required (also for kem_field_eval_device and
   kem_field_eval_sets). */
int kem_field_graph_create(kem_ctx* ctx, const kem_mesh* mesh, const double* sigma, int64_t npts, const double* x,
                           const double* y, const double* z, int precision, double* phi, double* ex, double* ey,
                           double* ez, kem_graph** out);
int kem_graph_launch(kem_graph* graph);
void kem_graph_destroy(kem_graph* graph);
/* Step-level field of several boundary fields: KSRootElectricField::CalculateFieldAndPotential
   (Kassiopeia/Simulation/Source/KSRootElectricField.cxx:104-121) sums every field's (E, phi) per
   sample point, in the order the fields were added, from zero.  Here nfields meshes of this
   context with their charge densities (device arrays, one per mesh) and precisions (KEM_FP64 /
   KEM_MIXED per field) are evaluated over all npts sample points of a step, one pass per field,
   and summed on the device in the same order (the same sums as per point).  Device SoA arrays,
   asynchronous on the context's stream; nfields = 0 gives zeros.  The reference rejects
   non-finite sample points (KSFieldError); here they give non-finite results. */
int kem_field_eval_fields(kem_ctx* ctx, int64_t nfields, const kem_mesh* const* meshes, const double* const* d_sigmas,
                          const int* precisions, int64_t npts, const double* d_x, const double* d_y,
                          const double* d_z, double* d_phi, double* d_ex, double* d_ey, double* d_ez);
int kem_field_eval_sets(kem_ctx* ctx, const kem_mesh* mesh, int64_t nsets, const double* sigma, const int32_t* set,
                        int64_t npts, const double* x, const double* y, const double* z, int precision, double* phi,
                        double* ex, double* ey, double* ez);

/* Instrumentation (no reference counterpart): number of (point, element) pairs of a field
   evaluation per quadrature regime of the numeric policy at the given precision (KEM_MIXED
   uses the FP32 far-field rule ladder), counts[12] = triangle 7/12/33-point/closed form,
   rectangle 7/12/33/closed form, segment 4-node/16-node/analytic, triangle 6-point (FP32
   ladder only).  Device points. */
int kem_field_census(kem_ctx* ctx, const kem_mesh* mesh, int64_t npts, const double* d_x, const double* d_y,
                     const double* d_z, int precision, int64_t* counts);

/* Instrumentation (no reference counterpart): the paths of the mixed-precision cluster far
   pass (single charge-density set, numeric policy) over device points, visited in the order a
   kem_field_eval_device pass of the same points visits them.  counts[10] = (point, element)
   pairs on the 3-point cluster path, on the 6-point record path, on the per-pair path by the
   rule each pair's own ladder picks (6, 7, 12, 33 points; segments 4, 16 nodes), near pairs
   left to the near pass, and the node evaluations the per-pair path runs (its wavefront rule
   for every far lane).  Synchronous. */
int kem_field_census_paths(kem_ctx* ctx, const kem_mesh* mesh, int64_t npts, const double* d_x, const double* d_y,
                           const double* d_z, int64_t* counts);

/* Instrumentation (no reference counterpart; host arithmetic only, no device access): how a
   field evaluation of npts points over nelem elements splits into element ranges and partial
   sums, for a mesh with (clustered = 1: triangles / rectangles / segments, numeric policy) or
   without element clusters, at a precision, with per-point set indices (multi = 1, as
   kem_field_eval_sets) or without, and a far-pass element range (KEM_TUNE_FAR_RANGE_*).
   ws_doubles < 0 plans for the workspace the library allocates.  out[8] = workspace doubles,
   element-weight doubles, partial-sum doubles written after them, element ranges, partial
   slices, points per chunk, XCD split (0 / 1), far kernel (0 FP64, 1 FP32 per pair, 2 FP32
   clusters).  The passes write only weights + partials <= workspace doubles. */
int kem_field_plan(int policy, int clustered, int precision, int multi, int64_t nelem, int64_t nsets, int64_t npts,
                   int64_t range_elems, int64_t ws_doubles, int64_t* out);

/* ---- dense influence matrix: KBoundaryIntegralMatrix<KElectrostaticBoundaryIntegrator, true>
   (BoundaryIntegrals/Core/include/KBoundaryIntegralMatrix.hh:16-99) as a device-resident
   KSquareMatrix<double> (LinearAlgebra/Core/include/KMatrix.hh:9-111).  A(i, j) is the
   collocation integral of source j at target i (Dirichlet: potential at the centroid of i;
   Neumann: E.n_i, diagonal (1+f)/(1-f)/(2 eps0)).  storage: KEM_STORE_FP64 and/or
   KEM_STORE_FP32. */
int kem_matrix_assemble(kem_ctx* ctx, const kem_mesh* mesh, int storage, kem_matrix** out);
/* Destroys a matrix (before its context).  Its FP64 / FP32 / BF16 storage returns to the
   context, which hands it to the next kem_matrix_assemble that needs a block of the same size
   (reallocating tens of GB right after freeing them waits for the freed memory to be cleared)
   and frees it in kem_ctx_destroy, or when a new allocation would otherwise fail. */
void kem_matrix_destroy(kem_matrix* A);
int64_t kem_matrix_dimension(const kem_matrix* A);
/* KSquareMatrix::operator()(i, j) */
int kem_matrix_get(kem_matrix* A, int64_t i, int64_t j, double* out);
/* copy the whole FP64 matrix to host, row-major (n x n) */
int kem_matrix_copy(kem_matrix* A, double* host_rowmajor);
/* KSquareMatrix::Multiply / MultiplyTranspose (KMatrix.hh:92-111): y = A x or y = A^T x, host vectors */
int kem_matrix_multiply(kem_matrix* A, int transpose, const double* x, double* y);

/* ---- charge-density solve: KKrylovChargeDensitySolver / KRobinHoodChargeDensitySolver
   (Interface/ChargeDensitySolvers/Electric/src/KKrylovChargeDensitySolver.cc:42-78,
   .../KRobinHoodChargeDensitySolver.cc:83-212) over KSimpleIterativeKrylovSolver<GMRES>
   (LinearAlgebra/Solvers/include/KSimpleIterativeKrylovSolver.hh:140-194) and KRobinHood
   (KRobinHood.hh:62-104).  x holds the start vector on entry (KBoundaryIntegralSolutionVector)
   and the solution on exit.  Returns KEM_OK, or KEM_W_NOT_CONVERGED at the iteration/time limit.
   The matrix keeps the GMRES workspace of its largest solve until kem_matrix_destroy (mostly
   the Krylov basis, (min(max_iterations, restart if set, 1000) + 1) x n doubles). */
void kem_solve_opts_default(kem_solve_opts* opts);
int kem_solve(kem_matrix* A, const kem_solve_opts* opts, const double* b, double* x, kem_solve_stats* stats);

/* Multi-right-hand-side solve (electrode voltage sets).  The reference solves one right-hand
   side per KChargeDensitySolver call (KKrylovChargeDensitySolver.cc:42-78); this runs GMRES
   for nrhs of them in lockstep, each with the reference's convergence rule
   (KSimpleIterativeKrylovSolver.hh:84-103), sharing one block matrix product per iteration
   (KEM_MIXED: FP32 matrix on MFMA, FP64 refinement against the FP64 matrix).  B, X: host,
   column-major n x nrhs (X holds the initial guesses and receives the solutions).
   residuals (optional, nrhs): final ||b - A x|| / ||b|| per right-hand side (FP64 matrix).
   stats: iterations / refinements = maximum over groups of 64, residual = worst.  The matrix
   keeps the lockstep workspace of its largest solve until kem_matrix_destroy (mostly the
   Krylov bases: (min(max_iterations, 300) + 1) x n x min(nrhs, 64) values, FP32 for the inner
   solves of KEM_MIXED / KEM_MIXED_BF16, FP64 for KEM_FP64). */
int kem_solve_batch(kem_matrix* A, const kem_solve_opts* opts, int64_t nrhs, const double* B, double* X,
                    double* residuals, kem_solve_stats* stats);
/* ---- adjoint (no reference counterpart: KEMField has no derivatives).  Gradients of a loss
   L(phi) of the potentials phi_k = sum_j G(P_k; T_j) sigma_j with respect to the charge
   densities and the triangle vertices, under the numeric policy with each pair's quadrature
   regime frozen.  Vertex gradients are per element and corner: dvert[j*9 + 3m + c] =
   d/d(v_m)_c with v_0 = P0, v_1 = P0 + A N1, v_2 = P0 + B N2 of element j's record.
   Triangles only (KEM_E_GEOM otherwise). */
/* A^T lambda = b by GMRES (FP64, or KEM_MIXED: FP32 A^T inner solves + FP64 refinement). */
int kem_solve_transposed(kem_matrix* A, const kem_solve_opts* opts, const double* b, double* x,
                         kem_solve_stats* stats);
/* gphi[k] = dL/dphi_k at npts points P (AoS): dsigma[j] = sum_k gphi_k G(P_k; T_j) and dvert
   (n x 9) = sum_k gphi_k sigma_j dG(P_k; T_j)/dvertices (host arrays, synchronous). */
int kem_field_vjp(kem_ctx* ctx, const kem_mesh* mesh, const double* sigma, int64_t npts, const double* P,
                  const double* gphi, double* dsigma, double* dvert);
/* dvert (n x 9) = -sum_ij lambda_i sigma_j dA_ij/dvertices for the Dirichlet collocation
   matrix A_ij = G(c_i; T_j), c_i = centroid of T_i (the term lambda^T (dA) sigma of the adjoint
   of A sigma = b).  KEM_E_STATE for meshes with Neumann rows. */
int kem_assemble_vjp(kem_ctx* ctx, const kem_mesh* mesh, const double* lambda, const double* sigma, double* dvert);

/* ---- charge-density solution cache (host code, no device): KChargeDensitySolver::FindSolution
   / SaveSolution (Interface/ChargeDensitySolvers/Electric/src/KChargeDensitySolver.cc:51-280)
   over KEMFileInterface's directory of .kbd files (IO/FileManipulation/include/
   KEMFileInterface.hh:23-81, KEMFile.hh), keyed by masked-bit MD5 hashes of the surfaces
   (IO/HashGenerator/include/KMD5HashGenerator.hh).  Surfaces as for kem_mesh_create (element
   records, boundary types and values, KSurfaceContainer order; no symmetry groups).  Errors
   return KEM_E_ARG with the message in kem_cache_error(). */
#define KEM_HASH_OMIT_BASIS 1     /* leave out the charge densities (KElectrostaticBasis) */
#define KEM_HASH_OMIT_DIRICHLET 2 /* leave out the Dirichlet values (KBoundaryType<.., KDirichletBoundary>) */
const char* kem_cache_error(void);
/* MD5 (RFC 1321) of bytes: 32 lowercase hex digits and a NUL into hex33 (md5.hh hex_digest) */
int kem_md5_hex(const void* data, int64_t bytes, char* hex33);
/* KMD5HashGenerator over n doubles streamed in order: each rounded half up at its highest
   masked bit and the masked_bits low bits cleared, values of magnitude below threshold left out
   (MaskedBits / Threshold; the reference's defaults are 1 and 1e-12, its solvers' 20 and 1e-14) */
int kem_hash_doubles(const double* x, int64_t n, int masked_bits, double threshold, char* hex33);
/* KMD5HashGenerator::GenerateHash(KSurfaceContainer) of the surfaces with charge densities sigma
   (nullable when omitted), leaving out the policies in omit (KEM_HASH_OMIT_*) */
int kem_hash_surfaces(int64_t n, const int32_t* shape_type, const double* shape_data, const int32_t* bc_type,
                      const double* bc_value, const double* bc_flux, const double* sigma, int omit, int masked_bits,
                      double threshold, char* hex33);
/* SaveSolution: appends to directory/file (created if missing) the residual threshold record,
   labelled with the shape, shape + boundary and solution hashes, and the solved container */
int kem_cache_save(const char* directory, const char* file, int64_t n, const int32_t* shape_type,
                   const double* shape_data, const int32_t* bc_type, const double* bc_value, const double* bc_flux,
                   const double* sigma, double residual_threshold, int masked_bits, double hash_threshold);
/* FindSolution: *found = 1 when a stored solution of these surfaces and boundary values has a
   residual threshold <= tolerance (the smallest such), 2 when stored solutions of the same
   surfaces with other boundary values superpose to these (KSuperpositionSolver), 0 otherwise;
   sigma (n) receives the charge densities, file_out (nullable, file_cap bytes) the file read */
int kem_cache_find(const char* directory, int64_t n, const int32_t* shape_type, const double* shape_data,
                   const int32_t* bc_type, const double* bc_value, const double* bc_flux, double tolerance,
                   int masked_bits, double hash_threshold, double* sigma, int* found, char* file_out,
                   int64_t file_cap);

/* Y = A X for nrhs vectors (host, column-major n x nrhs): KSquareMatrix::Multiply
   (KMatrix.hh:92-111) over a block; KEM_MIXED uses the FP32 matrix on FP32 MFMA, KEM_MIXED_BF16
   the BF16 matrix on BF16 MFMA (X rounded to BF16, FP32 accumulation, FP64 sum of the split-K
   partials). */
int kem_matrix_multiply_batch(kem_matrix* A, int precision, int64_t nrhs, const double* X, double* Y);

/* ---- zonal harmonics: the field of axisymmetric elements (conic sections, rings) about the z
   axis as Legendre expansions about source points on the axis.
   KZonalHarmonicContainer<KElectrostaticBasis> + ComputeCoefficients (FieldSolvers/ZonalHarmonic/
   Generator/include/KZonalHarmonicContainer.hh:246-345, KZonalHarmonicCoefficientGenerator.hh,
   KZHCoefficientGeneratorConicSection.cc, KZHCoefficientGeneratorRing.cc) -> kem_zh_create;
   KZonalHarmonicFieldSolver<KElectrostaticBasis>::ElectricFieldAndPotential (Solver/src/
   KElectrostaticZonalHarmonicFieldSolver.cc:98-124) -> kem_zh_field_eval(_device).  The
   parameters are KZonalHarmonicParameters' (Generator/include/KZonalHarmonicParameters.hh). */
typedef struct kem_zh_params {
    int32_t n_bifurcations;              /* 5 */
    double convergence_ratio;            /* 0.99 */
    double proximity_to_source_point;    /* 1e-14 */
    double convergence_parameter;        /* 1e-15 */
    int32_t n_central_coefficients;      /* 500 */
    int32_t central_fractional_spacing;  /* 1 */
    double central_fractional_distance;  /* 0.2 */
    double central_delta_z;              /* 0.01 */
    double central_z1, central_z2;       /* 0, 0: the elements' axial extent */
    int32_t n_remote_coefficients;       /* 500 */
    int32_t n_remote_source_points;      /* 3 */
    double remote_z1, remote_z2;         /* 0, 0 */
} kem_zh_params;
void kem_zh_params_default(kem_zh_params* params);
/* The expansion of a mesh (without symmetry groups) with charge densities sigma (host, mesh
   size; rings: total charge).  Elements that are not axisymmetric are integrated directly
   wherever the expansion is evaluated.  params may be NULL (the defaults).  The mesh must
   outlive the expansion. */
int kem_zh_create(kem_ctx* ctx, const kem_mesh* mesh, const double* sigma, const kem_zh_params* params,
                  kem_zh** out);
void kem_zh_destroy(kem_zh* zh);
/* out[4]: tree nodes, central source points, remote source points, tree depth */
int kem_zh_info(const kem_zh* zh, int64_t* out);
/* (E, phi) at device points (SoA; asynchronous unless counts is given).  counts (host, 3,
   nullable): central expansions, remote expansions and direct integrations the call took
   (KZonalHarmonicComputer's execution counts); with counts the call waits for the stream. */
int kem_zh_field_eval_device(kem_zh* zh, int64_t npts, const double* d_x, const double* d_y, const double* d_z,
                             double* d_phi, double* d_ex, double* d_ey, double* d_ez, int64_t* counts);
/* Same on host arrays (P npts x 3; phi, E nullable), synchronous. */
int kem_zh_field_eval(kem_zh* zh, int64_t npts, const double* P, double* phi, double* E, int64_t* counts);

#ifdef __cplusplus
}
#endif
#endif
