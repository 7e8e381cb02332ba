// kem_kernels.h -- host launchers of the HIP kernels (kem_kernels.hip).
#ifndef KEM_KERNELS_H
#define KEM_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace kem {

// Device arrays of a mesh that the field passes read.
struct FieldMesh {
    int64_t n = 0;
    const double* geom = nullptr;
    const double* nodes = nullptr;
    const double* recs = nullptr;      // raw shape records (analytic / reference policies)
    const float* nodes32 = nullptr;    // FP32 node table (mixed-precision far field)
    const double* clusters = nullptr;  // element clusters (kem::build_clusters); null: element scan
    const int32_t* order = nullptr;    // elements in cluster order
    const float* far6 = nullptr;       // 6-point far records in cluster order (kem::build_far6)
    const float* far3 = nullptr;       // 3-point far records in cluster order
    const double* far6_cl = nullptr;   // and the clusters' 6- and 3-point reach (KEM_FAR_CL_STRIDE)
    bool axisym = false;               // holds conic sections or rings (kernels with their integrators)
};

// A second stream on which the culled near pass runs beside the far pass (the near pass waits
// on latency, the far pass on the vector ALUs), with the two events that fork it from and join
// it to the calling stream.  Default priority, created by the caller; null st: one stream.
struct SideStream {
    hipStream_t st = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// Doubles of the ws argument of launch_field for npts points and nsets charge-density sets
// (multi: per-point set indices given) with the element ranges range_elems asks for.
int64_t field_workspace_doubles(int policy, const FieldMesh& m, bool mixed, double fp64_ratio, bool multi,
                                int64_t nsets, int64_t npts, int64_t range_elems);

// How launch_field splits one evaluation (host arithmetic, no device access).
struct FieldPlan {
    int kind = 0;          // far kernel: 0 FP64, 1 FP32 per pair, 2 FP32 over element clusters
    bool xcd = false;      // 8 R element ranges, R per XCD, over chunks of points
    int64_t ranges = 1;    // element ranges of the far pass
    int64_t slices = 1;    // partial-sum slices: the ranges, and one for a culled near pass
    int64_t per = 0;       // elements per range
    int64_t chunk = 0;     // points per chunk
    int64_t weights = 0;   // doubles of per-call element weights at the start of ws
    int64_t partials = 0;  // doubles of partial sums written after them (0: straight to the outputs)
};
// The plan for a workspace of ws_doubles doubles (< 0: field_workspace_doubles); launch_field
// refuses to run when weights + partials exceed its ws_doubles.
FieldPlan plan_field(int policy, const FieldMesh& m, bool mixed, double fp64_ratio, bool multi, int64_t nsets,
                     int64_t npts, int64_t range_elems, int64_t ws_doubles);

// phi/E at npts points (SoA) from the mesh's elements with charge densities sigma.  pset
// (nullable): per-point index of the charge-density set, point i using
// sigma[pset[i]*n + j], 0 <= pset[i] < nsets.  mixed selects the mixed-precision far field
// (FP32 node sums beyond distance ratio fp64_ratio, FP64 below it, FP64 near field and
// accumulation); with fp64_ratio = 0 and far6 given it runs over element clusters.  clusters /
// order select the culled near pass under the numeric policy; null scans every element.
// range_elems > 0: from 262,144 points on, a single-set numeric-policy far pass splits the
// elements into 8 R ranges of at most range_elems elements, R per XCD (0: no split).
// ws: scratch of ws_doubles doubles (element-range partial sums, per-set element weights); a
// smaller workspace than field_workspace_doubles asks for takes fewer element ranges, and one
// smaller than the plan's weights + partials returns hipErrorInvalidValue.  phi, ex, ey, ez:
// npts doubles each, all four required.  far_begin / far_end (nullable): events recorded on
// st around the far-field pass.  side (nullable): the culled near pass runs on side->st; the
// sums are the same with and without it (the near pass fills its own partial slice).
hipError_t launch_field(hipStream_t st, int policy, const FieldMesh& m, bool mixed, double fp64_ratio,
                        const double* sigma, const int32_t* pset, int64_t nsets, const double* px, const double* py,
                        const double* pz, int64_t npts, double* phi, double* ex, double* ey, double* ez, double* ws,
                        int64_t ws_doubles, int64_t range_elems, hipEvent_t far_begin = nullptr,
                        hipEvent_t far_end = nullptr, const SideStream* side = nullptr);

// Same as launch_field with the points visited in Morton order (kem_sort.hip): sort keys,
// gather (with the set indices), evaluate, scatter back.  sort_ws: sort_workspace_bytes(npts) bytes.
size_t sort_workspace_bytes(int64_t npts);
hipError_t launch_field_sorted(hipStream_t st, int policy, const FieldMesh& m, bool mixed, double fp64_ratio,
                               const double* sigma, const int32_t* pset, int64_t nsets, const double* px,
                               const double* py, const double* pz, int64_t npts, double* phi, double* ex, double* ey,
                               double* ez, void* sort_ws, size_t sort_ws_bytes, double* ws, int64_t ws_doubles,
                               int64_t range_elems, hipEvent_t far_begin = nullptr, hipEvent_t far_end = nullptr,
                               const SideStream* side = nullptr);

// *bad = 1 when a set index lies outside [0, nsets) (*bad must be 0 on entry)
hipError_t launch_check_sets(hipStream_t st, const int32_t* pset, int64_t n, int64_t nsets, int* bad);

// pairs per quadrature regime (11 counters, see k_census)
hipError_t launch_census(hipStream_t st, const double* geom, int64_t nelem, const double* px, const double* py,
                         const double* pz, int64_t npts, int mixed, double fp64_ratio, unsigned long long* counts);

// path census of the mixed cluster far pass (single set) over points in visiting order
// (10 counters, see k_census_far32c); launch_census_paths_sorted visits them in the Morton
// order of a pass of npts >= 4096 points (sort_ws: sort_workspace_bytes(npts) bytes)
hipError_t launch_census_paths(hipStream_t st, const FieldMesh& m, const double* px, const double* py,
                               const double* pz, int64_t npts, unsigned long long* counts);
hipError_t launch_census_paths_sorted(hipStream_t st, const FieldMesh& m, const double* px, const double* py,
                                      const double* pz, int64_t npts, void* sort_ws, size_t sort_ws_bytes,
                                      unsigned long long* counts);

hipError_t launch_pairs(hipStream_t st, int policy, const double* geom, const double* nodes, const double* recs,
                        int64_t npairs, const int64_t* elem, const double* P, double* out, bool axisym = false,
                        const int64_t* gstart = nullptr);

// collocation matrix; with element clusters (centre, radius per KEM_CLUSTER elements in the
// Morton order `order`) the numeric policy splits it into cubature and closed-form kernels, and
// without Neumann rows (field = false) that split needs anodes, assemble_nodes_doubles(n)
// doubles of workspace for the potential node table
int64_t assemble_nodes_doubles(int64_t n);
hipError_t launch_assemble(hipStream_t st, int policy, bool field, const double* geom, const double* nodes,
                           const double* recs, const double* targets, int64_t n, double* A64, float* A32,
                           bool axisym = false, const int64_t* gstart = nullptr, const double* clusters = nullptr,
                           const int32_t* order = nullptr, double* anodes = nullptr);
// phi[i] += t[i], ex[i] += t[npts + i], ey[i] += t[2 npts + i], ez[i] += t[3 npts + i]
hipError_t launch_accumulate4(hipStream_t st, const double* t, int64_t npts, double* phi, double* ex, double* ey,
                              double* ez);
// symmetry groups: out[s*nm + i] = sigma[s*n + gof[i]] (each member takes its group's density)
hipError_t launch_expand_groups(hipStream_t st, const int32_t* gof, int64_t nm, int64_t n, int64_t nsets,
                                const double* sigma, double* out);

// Launch shapes of the calling thread, set from the context at every C-ABI entry point:
// target block counts of the split-K matrix-vector product and of the FP32 / FP64 block product
// (register-loaded kernels in 64-row tiles; the LDS-streamed ones in 256-row workgroups).
struct LaunchTuning {
    int64_t gemv_blocks = 2048;
    int64_t bmm_blocks = 2560;
    int64_t bmm_stream_blocks = 480;  // workgroups of the LDS-streamed 64-RHS block products
    int64_t bmm_tile_blocks = 4096;   // most workgroups of the 16 x 16-tile products (FP32 <= 32 RHS, FP64)
    int64_t cus = 256;                // compute units of the device (set from the device at context creation)
    int64_t near_blocks = 384;  // grid of the near pass beside the mixed far pass (0: one block per point block)
    int64_t bmm_rows = 64;      // rows per wavefront tile of the FP32 block product (64, or 128 when n % 4 == 0)
    int64_t rh_groups = 0;      // workgroups of the Robin Hood sub-iterations (0: by size, 1: one)
};
void set_launch_tuning(const LaunchTuning& t);
const LaunchTuning& launch_tuning();

int gemv_splits(int64_t nrows, int64_t ncols);
// y = alpha * A x + beta * y, A column-major nrows x ncols (leading dimension lda);
// part: scratch of gemv_splits(nrows, ncols) * nrows doubles.  skip (nullable): device flag;
// the kernels do nothing while it is nonzero (a finished solve in a replayed graph).
template <typename T>
hipError_t launch_gemv(hipStream_t st, const T* A, int64_t lda, int64_t nrows, int64_t ncols, const double* x,
                       double alpha, double beta, double* y, double* part, const int* skip = nullptr);
// y = A^T x
template <typename T>
hipError_t launch_gemv_t(hipStream_t st, const T* A, int64_t lda, int64_t nrows, int64_t ncols, const double* x,
                         double* y, const int* skip = nullptr);

// ---- GMRES Arnoldi step on device state (one step per graph replay)
// State of one GMRES cycle: basis V (kmax+1 columns of n), Hessenberg H (column-major, leading
// dimension ld = kmax+1), rotations cs / sn, rotated right-hand side p, projections h1 / h2,
// norm scalar, the column index j and the stop flag done (ints), the cycle's column limit
// jlimit, and the tolerance tol * ||b|| against which |p[j+1]| is tested.
constexpr int kArnProjSplits = 16;  // most row ranges of a GMRES projection
int arn_proj_splits(int64_t n);
struct ArnoldiState {
    int64_t n = 0;
    int kmax = 0;
    double* V = nullptr;
    double* w = nullptr;
    double* vcur = nullptr;  // the current basis vector (input of the matrix-vector product)
    double* H = nullptr;
    double* cs = nullptr;
    double* sn = nullptr;
    double* p = nullptr;
    double* h1 = nullptr;
    double* h2 = nullptr;
    double* nrm = nullptr;
    double* part = nullptr;  // 256 dot partials + gemv partials (gemv_splits(n, kmax + 1) * n)
    double* hp = nullptr;    // projection partials, (kmax + 1) * kArnProjSplits
    int* ints = nullptr;     // [0] j, [1] done, [2] jlimit, [3] converged
    double* scal = nullptr;  // [0] tol * ||b||, [1] |p[j+1]| after the last step
};
// After the matrix-vector product w = A vcur: CGS2 against columns 0..j of V, normalisation into
// V[j+1] and vcur, Givens rotation of column j (KGeneralizedMinimalResidual.hh:178-271), j += 1;
// done = 1 when |p[j+1]| <= tol * ||b||, j reaches jlimit or the new norm is 0.  Every kernel
// returns at once while done is set.
hipError_t launch_arnoldi_step(hipStream_t st, const ArnoldiState& a);

// *out = a.b (or its square root); part: scratch of 256 doubles
hipError_t launch_dot(hipStream_t st, const double* a, const double* b, int64_t n, double* part, double* out,
                      bool want_sqrt);
// y += sign * (*alpha) * x
hipError_t launch_axpy_dev(hipStream_t st, const double* alpha, double sign, const double* x, double* y, int64_t n);
// y = x * (1 / *d)
hipError_t launch_div_dev(hipStream_t st, const double* d, const double* x, double* y, int64_t n);
// z = a x + b y
hipError_t launch_axpby(hipStream_t st, double a, const double* x, double b, const double* y, double* z, int64_t n);

// Gauss-Jordan without pivoting on M = [A | b] (column-major, n x (n+1), overwritten);
// f: scratch of n doubles; x: n doubles
hipError_t launch_gauss_jordan(hipStream_t st, double* M, int64_t n, double* f, double* x);

// Robin Hood sub-iterations: one workgroup up to 8,192 unknowns, above that robin_hood_groups(n)
// co-resident workgroups (cooperative launch) with slots (2 x 256 x 3 doubles) and flags (2
// uints: arrival counter, fault; flags[1] != 0 after the call: a workgroup timed out waiting)
int robin_hood_groups(int64_t n);
hipError_t launch_robin_hood(hipStream_t st, const double* A, const double* b, double* x, double* bi, int64_t n,
                             int64_t steps, double* rnorm, double* slots, unsigned int* flags);

// ---- multi-right-hand-side kernels (kem_batch.hip)
// Block-product precisions: FP64 matrix (FP64 MFMA; the VALU when n is odd), FP32 matrix (FP32
// MFMA), BF16 matrix (BF16 MFMA), each accumulating in its matrix's precision (BF16: FP32) and
// summing the split-K partials in FP64.
constexpr int kBmmFp64 = 0, kBmmFp32 = 1, kBmmBf16 = 2;
// Workspace of launch_bmm / launch_bmm16 for nrhs <= 64 right-hand sides.
size_t bmm_workspace_bytes(int64_t n, int nrhs, int prec);
// panel columns of a block product of nrhs right-hand sides (16, 32 or 64): the row stride of the
// FP32 / FP64 panels and the column count of the BF16 panel groups
int bmm_panel_cols(int64_t n, int nrhs, int prec);
// BF16 matrix storage of an n x n matrix: uint16_t elements (columns padded to a multiple of 8),
// packed from the FP64 (or, when A64 is null, FP32) column-major matrix: A16[(g*n + i)*8 + m] =
// A(i, 8g + m), each entry rounded to FP32 and then to BF16 to nearest even.
size_t bf16_matrix_elements(int64_t n);
hipError_t launch_pack_bf16(hipStream_t st, const double* A64, const float* A32, int64_t n, uint16_t* A16);
// Y[c*ldy + i] = sum_j A16(i, j) X(j, c) for c < nrhs <= 64 on BF16 MFMA (X rounded to BF16,
// FP32 accumulation, FP64 sum of the split-K partials).  x_is_panel: X is the packed BF16 panel
// (uint16_t, (n + 7) / 8 groups of 8 rows x (nrhs <= 32 ? 32 : 64) columns); else X is FP64,
// column-major with leading dimension ldx.
hipError_t launch_bmm16(hipStream_t st, const uint16_t* A16, int64_t n, const void* X, int64_t ldx, int nrhs,
                        double* Y, int64_t ldy, void* ws, size_t ws_bytes, bool x_is_panel,
                        const int* skip = nullptr);
// Y[c*ldy + i] = sum_j A(i, j) X(j, c) for c < nrhs <= 64: A32 != null selects the FP32 MFMA
// product (X rounded to FP32), else the FP64 product with A64.  x_is_panel: X already is the
// row-major panel (FP32: n x (nrhs <= 32 ? 32 : 64) floats; FP64: n x (nrhs <= 16 ? 16 : 64)).
hipError_t launch_bmm(hipStream_t st, const float* A32, const double* A64, int64_t n, const double* X, int64_t ldx,
                      int nrhs, double* Y, int64_t ldy, void* ws, size_t ws_bytes, bool x_is_panel,
                      const int* skip = nullptr);
// batched Krylov-basis operations, V[(r*(kmax+1) + q)*n + i] with vstride = (kmax+1)*n
hipError_t launch_bproj(hipStream_t st, const double* V, int64_t vstride, const double* W, int64_t n, int nq,
                        int nrhs, double* h, int ldh);
hipError_t launch_bupdate(hipStream_t st, const double* V, int64_t vstride, const double* h, int ldh, const int* cnt,
                          double sign, int64_t n, int nrhs, double* Y);
hipError_t launch_bupdate(hipStream_t st, const float* V, int64_t vstride, const double* h, int ldh, const int* cnt,
                          double sign, int64_t n, int nrhs, double* Y);
hipError_t launch_bnorm(hipStream_t st, const double* W, int64_t n, int nrhs, double* nrm, int stride);
hipError_t launch_bnormalize(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                             int64_t n, int nrhs, double* V, int64_t vstride, int q, int kp, float* Vt32,
                             double* Vt64, uint16_t* Vt16 = nullptr);
hipError_t launch_bnormalize(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                             int64_t n, int nrhs, float* V, int64_t vstride, int q, int kp, float* Vt32,
                             double* Vt64, uint16_t* Vt16 = nullptr);
hipError_t launch_baxpby(hipStream_t st, double a, const double* X, double b, const double* Y, double* Z, int64_t len);

// Lockstep GMRES over nrhs <= 64 right-hand sides on device state: after one step the
// Hessenberg columns, rotations and residuals of every active right-hand side are updated on
// the device (each with the single-RHS algorithm) and the step is captured as a HIP graph.
struct BatchArnoldi {
    int64_t n = 0;
    int nrhs = 0, kmax = 0, ldh = 0, kp = 0;
    bool use32 = false;
    const float* A32 = nullptr;
    const double* A64 = nullptr;
    const uint16_t* A16 = nullptr;  // non-null: BF16 block products (panel P16)
    double* V = nullptr;  // nrhs bases of kmax + 1 columns, vstride apart
    float* V32 = nullptr;  // the same in FP32 (used instead of V when set: inner solves of mixed precision)
    int64_t vstride = 0;
    double* W = nullptr;  // n x nrhs work block
    float* P32 = nullptr;  // row-major panel of the current basis vectors (n x kp)
    double* P64 = nullptr;
    uint16_t* P16 = nullptr;  // packed BF16 panel (groups of 8 rows, zero past n)
    void* ws = nullptr;  // block-product workspace
    size_t ws_bytes = 0;
    double *h1 = nullptr, *h2 = nullptr;  // projections, ldh x nrhs
    double* nrm = nullptr;                // nrhs
    double* H = nullptr;                  // nrhs x ldh x ldh
    double *cs = nullptr, *sn = nullptr, *p = nullptr;  // ldh x nrhs
    double *bnorm = nullptr, *tol = nullptr, *resid = nullptr;  // nrhs
    int *active = nullptr, *done = nullptr, *jr = nullptr, *cnt = nullptr;  // nrhs
    int* ints = nullptr;  // [0] j, [1] stop, [2] jlimit, [3] steps
    double* proj = nullptr;  // projection partials, batch_proj_doubles(n, nrhs, ldh)
};
// partial sums of the Arnoldi projections: row chunks x nrhs x ldh doubles
int64_t batch_proj_doubles(int64_t n, int nrhs, int ldh);
hipError_t launch_batch_arnoldi_step(hipStream_t st, const BatchArnoldi& a);
// Y (ldh x nrhs) = the least-squares coefficients of every right-hand side; cnt = jr
hipError_t launch_batch_trisolve(hipStream_t st, const BatchArnoldi& a, double* Y);

// ---- adjoint (kem_adjoint.hip): vertex derivatives under the numeric policy, triangles only
size_t vjp_workspace_doubles(int64_t n, int64_t npts);
// dsig[j] = sum_k gk[k] G(P_k; T_j);  dvert[j*9 + 3m + c] = sum_k gk[k] sigma_j dG(P_k; T_j)/d(v_m)_c
hipError_t launch_field_vjp(hipStream_t st, const double* geom, const double* sigma, const double* rule, int64_t n,
                            const double* pts, const double* gk, int64_t npts, double* dsig, double* dvert,
                            double* ws, size_t ws_doubles);
// dvert = d/dV sum_ij lw_i A_ij u_j (Dirichlet collocation matrix, centroids c_i move with T_i)
hipError_t launch_matrix_vjp(hipStream_t st, const double* geom, const double* targets, const double* rule,
                             int64_t n, const double* lw, const double* u, double* dvert, double* ws,
                             size_t ws_doubles);

}  // namespace kem

#endif
