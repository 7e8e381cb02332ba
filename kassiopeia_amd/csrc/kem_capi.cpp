// kem_capi.cpp -- C-ABI of the MI355X electrostatic BEM library (include/kemhip.h).
//
// Host orchestration in C++: context/stream ownership, mesh upload, matrix storage and the
// iterative solvers (GMRES, Robin Hood, FP32-matrix solves with FP64 iterative refinement),
// all driving the HIP kernels of kem_kernels.hip on device-resident data.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/kemhip.h"
#include "kem_kernels.h"
#include "kem_layout.h"
#include "kem_prep.h"
#include "kem_zh.h"

// mixed precision: far pairs up to this distance ratio keep FP64 cubature; 0 = every far pair
// in FP32 (the FP32 rule ladder alone keeps the global field error near 3e-8)
constexpr double kMixedFp64RatioDefault = 0.0;
// the inner tolerance of BF16-matrix solves: the BF16 product rounds its input vector to 8
// significant bits, so an inner solve cannot reduce its true residual much below 2^-8
constexpr double kBf16InnerTolerance = 1.0 / 256.0;
// elements per XCD range of the large single-set far passes (kem::launch_field); 0 = no split
constexpr int64_t kFarRangeDefault = 2560;

// Device scratch of the no-allocation field entry point (kem_field_eval_device_noalloc): one
// 48 MiB array per GPU, held until the process ends.  Host code never passes the array's name
// to a kernel (on the host it is a stand-in, not a device address): the address of each
// device's copy is looked up once, at the first context created on that device, and kept in
// g_scratch.  Calls on one GPU share it, so calls that overlap on one GPU race.
constexpr size_t kScratchBytes = (size_t)48 << 20;
constexpr int64_t kNoAllocMaxPoints = 65536;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) char scratch[kScratchBytes];
struct ScratchSlot {
    bool looked_up = false;
    hipError_t status = hipSuccess;
    void* ptr = nullptr;
};
std::mutex g_scratch_mu;
ScratchSlot g_scratch[kMaxDevices];

// Workspace of one field evaluation: partial sums and per-call element weights (ws), the
// charge densities of symmetry-group members (gsig) and the point ordering (sort).  A context
// owns one that grows with its calls; a captured graph owns one sized at capture, since a later
// call may reallocate the context's while the graph still holds its addresses; the
// no-allocation entry point uses one carved from the device's static scratch (fixed: never
// reallocated, a call that needs more is refused).
struct FieldWs {
    double* ws = nullptr;
    int64_t ws_doubles = 0;
    double* gsig = nullptr;
    int64_t gsig_doubles = 0;
    void* sort = nullptr;
    size_t sort_bytes = 0;
    bool fixed = false;
};

struct kem_ctx {
    int device = 0;
    hipStream_t own = nullptr;
    hipStream_t stream = nullptr;
    kem::SideStream side;  // the near pass beside the far pass (own stream, fork/join events)
    int policy = KEM_POLICY_NUMERIC;
    double mixed_fp64_ratio = kMixedFp64RatioDefault;
    int64_t far_range_fp64 = kFarRangeDefault, far_range_mixed = kFarRangeDefault;
    kem::LaunchTuning tuning;
    std::string err;
    double* ws = nullptr;  // matrix-vector partial sums (grow-only)
    int64_t ws_doubles = 0;
    double* scal = nullptr;  // small device scalars (dot products, Hessenberg column)
    int64_t scal_doubles = 0;
    FieldWs fws;  // workspace of eager field evaluations (grow-only)
    double* acc4 = nullptr;  // one field's (phi, Ex, Ey, Ez) of a step-level sum over fields (grow-only)
    int64_t acc4_doubles = 0;
    // storage of destroyed matrices, kept for the next matrix of the same size: freeing and
    // allocating tens of GB again costs up to a second on this hardware (the freed memory is
    // cleared before it is handed out again), as much as a 100,000-triangle assembly ten times
    std::vector<std::pair<void*, size_t>> store_cache;
    // far-kernel timing: one event pair per field evaluation since the last query
    static constexpr int kTimedMax = 256;
    bool timing = false;
    hipEvent_t ev[2 * kTimedMax] = {};
    int ev_used = 0;
};

struct kem_mesh {
    kem_ctx* ctx = nullptr;
    int64_t n = 0;   // unknowns: elements, or symmetry groups
    int64_t nm = 0;  // elements the device arrays hold (the members of all groups)
    int64_t* gstart = nullptr;  // symmetry groups: members [gstart[g], gstart[g+1]) of group g (null: none)
    int32_t* gof = nullptr;     // and the group of each member
    double* geom = nullptr;
    double* nodes = nullptr;
    double* targets = nullptr;
    float* nodes32 = nullptr;  // mixed-precision far-field node table
    double* recs = nullptr;      // raw shape records (analytic / reference policies)
    double* rule = nullptr;      // triangle rule barycentrics and weights (adjoint)
    bool all_triangles = true;
    bool has_axisym = false;  // conic sections or rings
    double* clusters = nullptr;  // near-pass element clusters (centre, radius)
    int32_t* order = nullptr;    // elements in cluster order
    float* far6 = nullptr;       // mixed far field: 6-point records in cluster order
    float* far3 = nullptr;       // 3-point records in cluster order
    double* far6_cl = nullptr;   // and the clusters' 6- and 3-point reach
    std::vector<int32_t> shape, bc;
    std::vector<double> value, flux;
    bool has_neumann = false;
};

struct kem_graph {
    kem_ctx* ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    FieldWs fws;  // the buffers the captured kernels address, owned by the graph
};

// zonal-harmonic expansion of a mesh's axisymmetric elements: the tree and coefficients on
// the device, the charge densities of the direct fallback, and the counts of the last call
struct kem_zh {
    kem_ctx* ctx = nullptr;
    const kem_mesh* mesh = nullptr;
    kem::ZHDevice dev;
    int32_t *node = nullptr, *child = nullptr;
    int64_t* elem = nullptr;
    double *sp = nullptr, *coeff = nullptr, *leg = nullptr, *sigma = nullptr;
    unsigned long long* counts = nullptr;
    int64_t nnodes = 0, ncentral = 0, nremote = 0;
};

struct kem_matrix {
    kem_ctx* ctx = nullptr;
    int64_t n = 0;
    double* A64 = nullptr;
    float* A32 = nullptr;
    uint16_t* A16 = nullptr;  // BF16, packed in groups of 8 columns (kem::launch_pack_bf16)
    int trans = 0;  // solves run on A^T (kem_solve_transposed)
    // the lockstep solver's device workspace (Krylov bases, Hessenberg matrices, panel, split
    // partials), kept from one solve to the next and grown when a solve needs more
    void* bws = nullptr;
    size_t bws_bytes = 0;
    // the single-right-hand-side solver's workspace (Krylov basis, Hessenberg matrix, partials)
    void* sws = nullptr;
    size_t sws_bytes = 0;
};

namespace {

int fail(kem_ctx* ctx, int code, const std::string& msg)
{
    if (ctx) ctx->err = msg;
    return code;
}

#define KEM_CK(ctx, expr)                                                                               \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail((ctx), KEM_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_));           \
    } while (0)

template <typename T>
hipError_t dalloc(T** p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) return hipSuccess;
    return hipMalloc((void**)p, sizeof(T) * (size_t)count);
}

template <typename T>
void dfree(T*& p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

// matrix storage: a cached block of exactly this size, else a new allocation (after releasing
// the cache if the first attempt fails)
template <typename T>
hipError_t store_alloc(kem_ctx* c, T** p, int64_t count)
{
    *p = nullptr;
    if (count <= 0) return hipSuccess;
    const size_t bytes = sizeof(T) * (size_t)count;
    for (size_t k = 0; k < c->store_cache.size(); k++)
        if (c->store_cache[k].second == bytes) {
            *p = (T*)c->store_cache[k].first;
            c->store_cache.erase(c->store_cache.begin() + (std::ptrdiff_t)k);
            return hipSuccess;
        }
    hipError_t e = hipMalloc((void**)p, bytes);
    if (e != hipSuccess && !c->store_cache.empty()) {
        (void)hipGetLastError();
        for (auto& b : c->store_cache) (void)hipFree(b.first);
        c->store_cache.clear();
        e = hipMalloc((void**)p, bytes);
    }
    return e;
}

template <typename T>
void store_free(kem_ctx* c, T*& p, int64_t count)
{
    if (p) c->store_cache.emplace_back((void*)p, sizeof(T) * (size_t)count);
    p = nullptr;
}

hipError_t ensure_ws(kem_ctx* c, int64_t doubles)
{
    if (c->ws_doubles >= doubles) return hipSuccess;
    dfree(c->ws);
    c->ws_doubles = 0;
    hipError_t e = dalloc(&c->ws, doubles);
    if (e == hipSuccess) c->ws_doubles = doubles;
    return e;
}

// grow w to at least the given sizes (a fixed workspace is never reallocated: a call that needs
// more than it holds is refused with hipErrorInvalidValue)
hipError_t ws_reserve(FieldWs& w, int64_t ws_doubles, int64_t gsig_doubles, size_t sort_bytes)
{
    if (w.fixed)
        return (w.ws_doubles >= ws_doubles && w.gsig_doubles >= gsig_doubles && w.sort_bytes >= sort_bytes)
                   ? hipSuccess
                   : hipErrorInvalidValue;
    hipError_t e = hipSuccess;
    if (w.ws_doubles < ws_doubles) {
        dfree(w.ws);
        w.ws_doubles = 0;
        e = dalloc(&w.ws, ws_doubles);
        if (e != hipSuccess) return e;
        w.ws_doubles = ws_doubles;
    }
    if (w.gsig_doubles < gsig_doubles) {
        dfree(w.gsig);
        w.gsig_doubles = 0;
        e = dalloc(&w.gsig, gsig_doubles);
        if (e != hipSuccess) return e;
        w.gsig_doubles = gsig_doubles;
    }
    if (w.sort_bytes < sort_bytes) {
        if (w.sort) (void)hipFree(w.sort);
        w.sort = nullptr;
        w.sort_bytes = 0;
        e = hipMalloc(&w.sort, sort_bytes);
        if (e != hipSuccess) return e;
        w.sort_bytes = sort_bytes;
    }
    return hipSuccess;
}

void ws_release(FieldWs& w)
{
    if (!w.fixed) {
        dfree(w.ws);
        dfree(w.gsig);
        if (w.sort) (void)hipFree(w.sort);
    }
    w = FieldWs();
}

hipError_t ensure_scal(kem_ctx* c, int64_t doubles)
{
    if (c->scal_doubles >= doubles) return hipSuccess;
    dfree(c->scal);
    c->scal_doubles = 0;
    hipError_t e = dalloc(&c->scal, doubles);
    if (e == hipSuccess) c->scal_doubles = doubles;
    return e;
}

// make the context's device current and drop a stale error another caller left behind, so
// that the hipGetLastError() after each launch reports this call's own launches
hipError_t enter(kem_ctx* c)
{
    (void)hipGetLastError();
    kem::set_launch_tuning(c->tuning);
    return hipSetDevice(c->device);
}

// the next timing event pair when kernel timing is on (else null events)
void timing_pair(kem_ctx* c, hipEvent_t* t0, hipEvent_t* t1)
{
    *t0 = *t1 = nullptr;
    if (c->timing && c->ev_used < kem_ctx::kTimedMax) {
        *t0 = c->ev[2 * c->ev_used];
        *t1 = c->ev[2 * c->ev_used + 1];
        c->ev_used++;
    }
}

// RAII device vector for solver temporaries
struct DVec {
    double* p = nullptr;
    ~DVec() { dfree(p); }
};

// point counts from which field evaluation visits points in Morton order
constexpr int64_t kSortMinPoints = 4096;

// Workspace sizes of one field evaluation with `range` elements per XCD range
struct WsNeed {
    int64_t ws = 0, gsig = 0;
    size_t sort = 0;
};
WsNeed field_ws_need(const kem_ctx* c, const kem_mesh* m, const kem::FieldMesh& fm, bool multi, int64_t nsets,
                     int64_t npts, bool mixed, int64_t range)
{
    WsNeed n;
    n.ws = kem::field_workspace_doubles(c->policy, fm, mixed, c->mixed_fp64_ratio, multi, nsets, npts, range);
    n.gsig = m->gstart ? nsets * m->nm : 0;
    if (npts >= kSortMinPoints && npts <= INT32_MAX) n.sort = kem::sort_workspace_bytes(npts);
    return n;
}

double now_s()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

int kem_version(void) { return 1; }

int kem_ctx_create(int device, kem_ctx** out)
{
    if (!out) return KEM_E_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return KEM_E_HIP;
    if (device < 0 || device >= ndev) return KEM_E_ARG;
    if (hipSetDevice(device) != hipSuccess) return KEM_E_HIP;
    kem_ctx* c = new kem_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->own, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&c->side.st, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->side.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->side.join, hipEventDisableTiming) != hipSuccess) {
        kem_ctx_destroy(c);
        return KEM_E_HIP;
    }
    c->stream = c->own;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cus > 0)
        c->tuning.cus = cus;
    (void)hipGetLastError();
    if (device < kMaxDevices) {
        // the device's copy of the no-allocation scratch, once per device, with it current
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        ScratchSlot& s = g_scratch[device];
        if (!s.looked_up) {
            void* p = nullptr;
            s.status = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            s.ptr = s.status == hipSuccess ? p : nullptr;
            s.looked_up = true;
            (void)hipGetLastError();
        }
    }
    *out = c;
    return KEM_OK;
}

void kem_ctx_destroy(kem_ctx* c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    dfree(c->ws);
    dfree(c->scal);
    dfree(c->acc4);
    ws_release(c->fws);
    for (auto& b : c->store_cache) (void)hipFree(b.first);
    c->store_cache.clear();
    for (hipEvent_t& e : c->ev)
        if (e) (void)hipEventDestroy(e);
    if (c->side.st) (void)hipStreamSynchronize(c->side.st);
    if (c->side.fork) (void)hipEventDestroy(c->side.fork);
    if (c->side.join) (void)hipEventDestroy(c->side.join);
    if (c->side.st) (void)hipStreamDestroy(c->side.st);
    if (c->own) (void)hipStreamDestroy(c->own);
    delete c;
}

int kem_ctx_set_stream(kem_ctx* c, void* stream)
{
    if (!c) return KEM_E_ARG;
    c->stream = stream ? (hipStream_t)stream : c->own;
    return KEM_OK;
}

void* kem_ctx_stream(kem_ctx* c) { return c ? (void*)c->stream : nullptr; }

int kem_ctx_set_mixed_fp64_ratio(kem_ctx* c, double ratio)
{
    if (!c || !(ratio >= 0.)) return KEM_E_ARG;
    c->mixed_fp64_ratio = ratio;
    return KEM_OK;
}

int kem_ctx_set_tuning(kem_ctx* c, int param, int64_t value)
{
    if (!c) return KEM_E_ARG;
    switch (param) {
    case KEM_TUNE_FAR_RANGE_FP64:
    case KEM_TUNE_FAR_RANGE_MIXED:
        if (value < 0 || value > ((int64_t)1 << 31)) return fail(c, KEM_E_ARG, "tuning: range size out of [0, 2^31]");
        (param == KEM_TUNE_FAR_RANGE_FP64 ? c->far_range_fp64 : c->far_range_mixed) = value;
        return KEM_OK;
    case KEM_TUNE_GEMV_BLOCKS:
    case KEM_TUNE_BMM_BLOCKS:
        if (value < 1 || value > (1 << 20)) return fail(c, KEM_E_ARG, "tuning: block count out of [1, 2^20]");
        (param == KEM_TUNE_GEMV_BLOCKS ? c->tuning.gemv_blocks : c->tuning.bmm_blocks) = value;
        return KEM_OK;
    case KEM_TUNE_NEAR_BLOCKS:
        if (value < 0 || value > (1 << 20)) return fail(c, KEM_E_ARG, "tuning: block count out of [0, 2^20]");
        c->tuning.near_blocks = value;
        return KEM_OK;
    case KEM_TUNE_BMM_ROWS:
        if (value != 64 && value != 128) return fail(c, KEM_E_ARG, "tuning: block-product rows per wavefront 64 or 128");
        c->tuning.bmm_rows = value;
        return KEM_OK;
    case KEM_TUNE_ROBIN_HOOD_GROUPS:
        if (value != 0 && value != 1) return fail(c, KEM_E_ARG, "tuning: Robin Hood workgroups 0 (by size) or 1");
        c->tuning.rh_groups = value;
        return KEM_OK;
    case KEM_TUNE_BMM_STREAM_BLOCKS:
    case KEM_TUNE_BMM_TILE_BLOCKS:
        if (value < 1 || value > (1 << 20)) return fail(c, KEM_E_ARG, "tuning: block count out of [1, 2^20]");
        (param == KEM_TUNE_BMM_STREAM_BLOCKS ? c->tuning.bmm_stream_blocks : c->tuning.bmm_tile_blocks) = value;
        return KEM_OK;
    default:
        return fail(c, KEM_E_ARG, "tuning: unknown parameter");
    }
}

int kem_ctx_get_tuning(const kem_ctx* c, int param, int64_t* value)
{
    if (!c || !value) return KEM_E_ARG;
    switch (param) {
    case KEM_TUNE_FAR_RANGE_FP64: *value = c->far_range_fp64; return KEM_OK;
    case KEM_TUNE_FAR_RANGE_MIXED: *value = c->far_range_mixed; return KEM_OK;
    case KEM_TUNE_GEMV_BLOCKS: *value = c->tuning.gemv_blocks; return KEM_OK;
    case KEM_TUNE_BMM_BLOCKS: *value = c->tuning.bmm_blocks; return KEM_OK;
    case KEM_TUNE_NEAR_BLOCKS: *value = c->tuning.near_blocks; return KEM_OK;
    case KEM_TUNE_BMM_ROWS: *value = c->tuning.bmm_rows; return KEM_OK;
    case KEM_TUNE_ROBIN_HOOD_GROUPS: *value = c->tuning.rh_groups; return KEM_OK;
    case KEM_TUNE_BMM_STREAM_BLOCKS: *value = c->tuning.bmm_stream_blocks; return KEM_OK;
    case KEM_TUNE_BMM_TILE_BLOCKS: *value = c->tuning.bmm_tile_blocks; return KEM_OK;
    default: return KEM_E_ARG;
    }
}

int kem_ctx_set_kernel_timing(kem_ctx* c, int on)
{
    if (!c) return KEM_E_ARG;
    KEM_CK(c, enter(c));
    if (on)
        for (hipEvent_t& e : c->ev)
            if (!e) KEM_CK(c, hipEventCreate(&e));
    c->timing = on != 0;
    c->ev_used = 0;
    return KEM_OK;
}

int kem_ctx_kernel_time(kem_ctx* c, double* total_ms, int* launches)
{
    if (!c || !total_ms || !launches) return KEM_E_ARG;
    KEM_CK(c, enter(c));
    double t = 0.;
    for (int k = 0; k < c->ev_used; k++) {
        KEM_CK(c, hipEventSynchronize(c->ev[2 * k + 1]));
        float ms = 0.f;
        KEM_CK(c, hipEventElapsedTime(&ms, c->ev[2 * k], c->ev[2 * k + 1]));
        t += ms;
    }
    *total_ms = t;
    *launches = c->ev_used;
    c->ev_used = 0;
    return KEM_OK;
}

int kem_dev_alloc(kem_ctx* c, int64_t bytes, void** p)
{
    if (!c || !p || bytes < 0) return KEM_E_ARG;
    *p = nullptr;
    if (bytes == 0) return KEM_OK;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipMalloc(p, (size_t)bytes));
    return KEM_OK;
}

int kem_dev_free(kem_ctx* c, void* p)
{
    if (!c) return KEM_E_ARG;
    if (!p) return KEM_OK;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    KEM_CK(c, hipFree(p));
    return KEM_OK;
}

int kem_dev_upload(kem_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || (bytes > 0 && (!dst || !src)) || bytes < 0) return KEM_E_ARG;
    if (bytes == 0) return KEM_OK;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyHostToDevice, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

int kem_dev_download(kem_ctx* c, void* dst, const void* src, int64_t bytes)
{
    if (!c || (bytes > 0 && (!dst || !src)) || bytes < 0) return KEM_E_ARG;
    if (bytes == 0) return KEM_OK;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

int kem_ctx_sync(kem_ctx* c)
{
    if (!c) return KEM_E_ARG;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

const char* kem_last_error(const kem_ctx* c) { return c ? c->err.c_str() : "null context"; }

int kem_set_integrator(kem_ctx* c, const char* name)
{
    if (!c || !name) return KEM_E_ARG;
    std::string s(name);
    if (s == "numeric" || s == "default")
        c->policy = KEM_POLICY_NUMERIC;
    else if (s == "rwg")
        c->policy = KEM_POLICY_RWG;
    else if (s == "analytic")
        c->policy = KEM_POLICY_ANALYTIC;
    else if (s == "reference")
        c->policy = KEM_POLICY_REFERENCE;
    else
        return fail(c, KEM_E_NAME, "unknown integrator '" + s + "' (known: numeric, analytic, rwg, reference, default)");
    return KEM_OK;
}

// ------------------------------------------------------------------ mesh
namespace {
// A mesh of n unknowns over nm elements: gs (nullable) gives the members [gs[g], gs[g+1]) of
// each symmetry group; without it every element is its own unknown (nm = n).
int mesh_build(kem_ctx* c, int64_t n, const int64_t* gs, int64_t nm, const int32_t* shape, const double* data,
               const int32_t* bc, const double* value, const double* flux, kem_mesh** out)
{
    KEM_CK(c, enter(c));
    std::vector<double> g((size_t)nm * KG_STRIDE), nd((size_t)nm * KEM_NODE_STRIDE), t((size_t)n * KT_STRIDE);
    std::vector<float> nd32((size_t)nm * KEM_NODE32_STRIDE);
    kem_mesh* m = new kem_mesh();
    m->ctx = c;
    m->n = n;
    m->nm = nm;
    m->shape.assign(shape, shape + nm);
    m->bc.assign((size_t)n, 0);
    m->value.assign((size_t)n, 0.);
    m->flux.assign((size_t)n, 0.);
    for (int64_t i = 0; i < n; i++) {
        if (bc) m->bc[i] = bc[i];
        if (value) m->value[i] = value[i];
        if (flux) m->flux[i] = flux[i];
        if (m->bc[i] != 0 && m->bc[i] != 1) {
            delete m;
            return fail(c, KEM_E_ARG, "mesh: boundary type must be 0 (Dirichlet) or 1 (Neumann)");
        }
        if (m->bc[i] == 1) m->has_neumann = true;
        // the collocation target of a group is its first member (KSymmetryGroup::Centroid /
        // Normal, Surfaces/include/KSymmetryGroup.hh:48-59)
        const int64_t f = gs ? gs[i] : i;
        kem::prep_target(shape[f], data + 11 * f, m->bc[i], m->flux[i], t.data() + (size_t)i * KT_STRIDE);
    }
    for (int64_t i = 0; i < nm; i++) {
        if (kem::prep_element(shape[i], data + 11 * i, g.data() + (size_t)i * KG_STRIDE,
                              nd.data() + (size_t)i * KEM_NODE_STRIDE) != 0) {
            delete m;
            return fail(c, KEM_E_GEOM, "mesh: degenerate or non-finite element " + std::to_string(i));
        }
        kem::prep_nodes32(g.data() + (size_t)i * KG_STRIDE, nd.data() + (size_t)i * KEM_NODE_STRIDE,
                          nd32.data() + (size_t)i * KEM_NODE32_STRIDE);
    }
    std::vector<int32_t> order((size_t)nm);
    std::vector<double> clusters((size_t)(4 * ((nm + KEM_CLUSTER - 1) / KEM_CLUSTER)));
    kem::build_clusters(nm, g.data(), order.data(), clusters.data());
    std::vector<float> far6((size_t)nm * KEM_FAR6_STRIDE), far3((size_t)nm * KEM_FAR3_STRIDE);
    std::vector<double> far6_cl(clusters.size() / 4 * KEM_FAR_CL_STRIDE);
    kem::build_far6(nm, g.data(), nd32.data(), order.data(), clusters.data(), far6.data(), far3.data(),
                    far6_cl.data());
    std::vector<double> recs((size_t)nm * KEM_REC, 0.);
    for (int64_t i = 0; i < nm; i++)
        for (int k = 0; k < 11; k++) recs[(size_t)i * KEM_REC + k] = data[11 * i + k];
    std::vector<double> rule(4 * KEM_NODE_SLOTS, 0.);
    kem::triangle_rule_table(rule.data());
    for (int64_t i = 0; i < nm; i++) {
        m->all_triangles = m->all_triangles && shape[i] == KEM_TRIANGLE;
        m->has_axisym = m->has_axisym || shape[i] == KEM_CONICSECTION || shape[i] == KEM_RING;
    }
    hipError_t e = dalloc(&m->geom, (int64_t)g.size());
    if (e == hipSuccess) e = dalloc(&m->rule, (int64_t)rule.size());
    if (e == hipSuccess) e = hipMemcpy(m->rule, rule.data(), sizeof(double) * rule.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->recs, (int64_t)recs.size());
    if (e == hipSuccess) e = hipMemcpy(m->recs, recs.data(), sizeof(double) * recs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->clusters, (int64_t)clusters.size());
    if (e == hipSuccess) e = dalloc(&m->order, nm);
    if (e == hipSuccess && gs) {
        std::vector<int32_t> gof((size_t)nm);
        for (int64_t k = 0; k < n; k++)
            for (int64_t i = gs[k]; i < gs[k + 1]; i++) gof[(size_t)i] = (int32_t)k;
        e = dalloc(&m->gstart, n + 1);
        if (e == hipSuccess) e = hipMemcpy(m->gstart, gs, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = dalloc(&m->gof, nm);
        if (e == hipSuccess) e = hipMemcpy(m->gof, gof.data(), sizeof(int32_t) * nm, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess)
        e = hipMemcpy(m->clusters, clusters.data(), sizeof(double) * clusters.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->order, order.data(), sizeof(int32_t) * order.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->far6, (int64_t)far6.size());
    if (e == hipSuccess) e = hipMemcpy(m->far6, far6.data(), sizeof(float) * far6.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->far3, (int64_t)far3.size());
    if (e == hipSuccess) e = hipMemcpy(m->far3, far3.data(), sizeof(float) * far3.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->far6_cl, (int64_t)far6_cl.size());
    if (e == hipSuccess)
        e = hipMemcpy(m->far6_cl, far6_cl.data(), sizeof(double) * far6_cl.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = dalloc(&m->nodes, (int64_t)nd.size());
    if (e == hipSuccess) e = dalloc(&m->targets, (int64_t)t.size());
    if (e == hipSuccess) e = dalloc(&m->nodes32, (int64_t)nd32.size());
    if (e == hipSuccess)
        e = hipMemcpy(m->nodes32, nd32.data(), sizeof(float) * nd32.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->geom, g.data(), sizeof(double) * g.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->nodes, nd.data(), sizeof(double) * nd.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(m->targets, t.data(), sizeof(double) * t.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        kem_mesh_destroy(m);
        return fail(c, KEM_E_HIP, std::string("mesh upload: ") + hipGetErrorString(e));
    }
    *out = m;
    return KEM_OK;
}
}  // namespace

int kem_mesh_create(kem_ctx* c, int64_t n, const int32_t* shape, const double* data, const int32_t* bc,
                    const double* value, const double* flux, kem_mesh** out)
{
    if (!c || !out) return KEM_E_ARG;
    *out = nullptr;
    if (n <= 0 || !shape || !data) return fail(c, KEM_E_ARG, "mesh: empty or null arrays");
    return mesh_build(c, n, nullptr, n, shape, data, bc, value, flux, out);
}

int kem_mesh_create_groups(kem_ctx* c, int64_t ngroups, const int64_t* group_start, const int32_t* shape,
                           const double* data, const int32_t* bc, const double* value, const double* flux,
                           kem_mesh** out)
{
    if (!c || !out) return KEM_E_ARG;
    *out = nullptr;
    if (ngroups <= 0 || !group_start || !shape || !data) return fail(c, KEM_E_ARG, "mesh: empty or null arrays");
    if (group_start[0] != 0) return fail(c, KEM_E_ARG, "mesh: group_start[0] must be 0");
    for (int64_t k = 0; k < ngroups; k++) {
        if (group_start[k + 1] <= group_start[k]) return fail(c, KEM_E_ARG, "mesh: every group needs a member");
        for (int64_t i = group_start[k] + 1; i < group_start[k + 1]; i++)
            if (shape[i] != shape[group_start[k]])
                return fail(c, KEM_E_ARG, "mesh: the members of a group share one shape type");
    }
    if (group_start[ngroups] > INT32_MAX) return fail(c, KEM_E_ARG, "mesh: too many members");
    return mesh_build(c, ngroups, group_start, group_start[ngroups], shape, data, bc, value, flux, out);
}

void kem_mesh_destroy(kem_mesh* m)
{
    if (!m) return;
    if (m->ctx) (void)hipSetDevice(m->ctx->device);
    dfree(m->geom);
    dfree(m->nodes);
    dfree(m->targets);
    dfree(m->nodes32);
    dfree(m->clusters);
    dfree(m->order);
    dfree(m->far6);
    dfree(m->far3);
    dfree(m->far6_cl);
    dfree(m->recs);
    dfree(m->rule);
    dfree(m->gstart);
    dfree(m->gof);
    delete m;
}

int64_t kem_mesh_size(const kem_mesh* m) { return m ? m->n : 0; }

int kem_mesh_rhs(const kem_mesh* m, double* b)
{
    if (!m || !b) return KEM_E_ARG;
    for (int64_t i = 0; i < m->n; i++) b[i] = m->bc[i] == 0 ? m->value[i] : 0.;
    return KEM_OK;
}

int kem_mesh_set_values(kem_mesh* m, const double* value)
{
    if (!m || !value) return KEM_E_ARG;
    m->value.assign(value, value + m->n);
    return KEM_OK;
}

// ------------------------------------------------------------------ evaluation
int kem_element_eval(kem_ctx* c, const kem_mesh* m, int64_t np, const int64_t* elem, const double* P, double* out)
{
    if (!c || !m || (np > 0 && (!elem || !P || !out))) return KEM_E_ARG;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (np <= 0) return KEM_OK;
    for (int64_t k = 0; k < np; k++)
        if (elem[k] < 0 || elem[k] >= m->n) return fail(c, KEM_E_ARG, "element index out of range");
    KEM_CK(c, enter(c));
    int64_t* de = nullptr;
    double *dP = nullptr, *dO = nullptr;
    hipError_t e = dalloc(&de, np);
    if (e == hipSuccess) e = dalloc(&dP, 3 * np);
    if (e == hipSuccess) e = dalloc(&dO, 4 * np);
    if (e == hipSuccess) e = hipMemcpyAsync(de, elem, sizeof(int64_t) * np, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dP, P, sizeof(double) * 3 * np, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess) e = kem::launch_pairs(c->stream, c->policy, m->geom, m->nodes, m->recs, np, de, dP, dO, m->has_axisym,
                                                m->gstart);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dO, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(de);
    dfree(dP);
    dfree(dO);
    if (e != hipSuccess) return fail(c, KEM_E_HIP, std::string("element_eval: ") + hipGetErrorString(e));
    return KEM_OK;
}

namespace {
// field evaluation on device arrays; pset (nullable): per-point charge-density set
void field_mesh(const kem_mesh* m, kem::FieldMesh& f)
{
    f.n = m->nm;
    f.geom = m->geom;
    f.nodes = m->nodes;
    f.recs = m->recs;
    f.nodes32 = m->nodes32;
    f.clusters = m->clusters;
    f.order = m->order;
    f.far6 = m->far6;
    f.far3 = m->far3;
    f.far6_cl = m->far6_cl;
    f.axisym = m->has_axisym;
}

// Field evaluation on device arrays into workspace w; pset (nullable): per-point
// charge-density set.  side: run the near pass on the context's side stream (not under graph
// capture).  range < 0: the context's element ranges for the precision.
int field_eval(kem_ctx* c, const kem_mesh* m, const double* sigma, const int32_t* pset, int64_t nsets, int64_t npts,
               const double* x, const double* y, const double* z, int precision, double* phi, double* ex, double* ey,
               double* ez, FieldWs& w, bool side, int64_t range = -1)
{
    const kem::SideStream* ss = side ? &c->side : nullptr;
    kem::FieldMesh fm;
    field_mesh(m, fm);
    const bool mixed = precision == KEM_MIXED;
    if (range < 0) range = mixed ? c->far_range_mixed : c->far_range_fp64;
    const WsNeed need = field_ws_need(c, m, fm, pset != nullptr, nsets, npts, mixed, range);
    // a fixed workspace may hold less than the preferred size: the plan then takes fewer
    // element ranges, and launch_field refuses one that does not fit
    KEM_CK(c, ws_reserve(w, w.fixed ? 0 : need.ws, need.gsig, need.sort));
    if (m->gstart) {
        // symmetry groups: every member carries its group's charge density
        KEM_CK(c, kem::launch_expand_groups(c->stream, m->gof, m->nm, m->n, nsets, sigma, w.gsig));
        sigma = w.gsig;
    }
    hipEvent_t t0, t1;
    timing_pair(c, &t0, &t1);
    if (need.sort) {
        // visit the points in Morton order so the lanes of a wavefront share quadrature rules
        KEM_CK(c, kem::launch_field_sorted(c->stream, c->policy, fm, mixed, c->mixed_fp64_ratio, sigma, pset, nsets,
                                           x, y, z, npts, phi, ex, ey, ez, w.sort, w.sort_bytes, w.ws, w.ws_doubles,
                                           range, t0, t1, ss));
        return KEM_OK;
    }
    KEM_CK(c, kem::launch_field(c->stream, c->policy, fm, mixed, c->mixed_fp64_ratio, sigma, pset, nsets, x, y, z,
                                npts, phi, ex, ey, ez, w.ws, w.ws_doubles, range, t0, t1, ss));
    return KEM_OK;
}
}  // namespace

int kem_field_eval_device(kem_ctx* c, const kem_mesh* m, const double* sigma, int64_t npts, const double* x,
                          const double* y, const double* z, int precision, double* phi, double* ex, double* ey,
                          double* ez)
{
    if (!c || !m || !sigma || (npts > 0 && (!x || !y || !z || !phi || !ex || !ey || !ez))) return KEM_E_ARG;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (precision != KEM_FP64 && precision != KEM_MIXED) return fail(c, KEM_E_ARG, "field: unknown precision mode");
    if (npts <= 0) return KEM_OK;
    KEM_CK(c, enter(c));
    return field_eval(c, m, sigma, nullptr, 1, npts, x, y, z, precision, phi, ex, ey, ez, c->fws, true);
}

int kem_field_eval_device_noalloc(kem_ctx* c, const kem_mesh* m, const double* sigma, int64_t npts,
                                  const double* x, const double* y, const double* z, int precision, double* phi,
                                  double* ex, double* ey, double* ez)
{
    if (!c || !m || !sigma || (npts > 0 && (!x || !y || !z || !phi || !ex || !ey || !ez))) return KEM_E_ARG;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (precision != KEM_FP64 && precision != KEM_MIXED) return fail(c, KEM_E_ARG, "field: unknown precision mode");
    if (npts > kNoAllocMaxPoints)
        return fail(c, KEM_E_ARG, "no-allocation field: at most " + std::to_string(kNoAllocMaxPoints) +
                                      " points per call (" + std::to_string(npts) + " given)");
    if (npts <= 0) return KEM_OK;
    void* base = nullptr;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        if (c->device < kMaxDevices && g_scratch[c->device].looked_up) {
            if (g_scratch[c->device].status != hipSuccess)
                return fail(c, KEM_E_HIP, std::string("no-allocation field: scratch address: ") +
                                              hipGetErrorString(g_scratch[c->device].status));
            base = g_scratch[c->device].ptr;
        }
    }
    if (!base) return fail(c, KEM_E_STATE, "no-allocation field: no scratch on this device");
    KEM_CK(c, enter(c));
    // the scratch holds the point ordering, the group members' charge densities and the
    // partial sums + element weights, each 256-byte aligned; the partial sums get what is left,
    // which must hold the workspace the eager path would allocate (the same plan, the same sums)
    kem::FieldMesh fm;
    field_mesh(m, fm);
    const bool mixed = precision == KEM_MIXED;
    const int64_t range = mixed ? c->far_range_mixed : c->far_range_fp64;
    const WsNeed need = field_ws_need(c, m, fm, false, 1, npts, mixed, range);
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t sort_b = up(need.sort), gsig_b = up(sizeof(double) * (size_t)need.gsig);
    const size_t ws_b = sizeof(double) * (size_t)need.ws;
    if (sort_b + gsig_b + ws_b > kScratchBytes)
        return fail(c, KEM_E_ARG, "no-allocation field: " + std::to_string(npts) + " points over " +
                                      std::to_string(m->nm) + " elements need " +
                                      std::to_string((sort_b + gsig_b + ws_b) >> 20) + " MiB of scratch; it holds " +
                                      std::to_string(kScratchBytes >> 20) + " MiB");
    char* p = (char*)base;
    FieldWs w;
    w.fixed = true;
    w.sort = need.sort ? (void*)p : nullptr;
    w.sort_bytes = need.sort;
    p += sort_b;
    w.gsig = need.gsig ? (double*)p : nullptr;
    w.gsig_doubles = need.gsig;
    p += gsig_b;
    w.ws = (double*)p;
    w.ws_doubles = (int64_t)((kScratchBytes - sort_b - gsig_b) / sizeof(double));
    return field_eval(c, m, sigma, nullptr, 1, npts, x, y, z, precision, phi, ex, ey, ez, w, true);
}

int kem_field_graph_create(kem_ctx* c, const kem_mesh* m, const double* sigma, int64_t npts, const double* x,
                           const double* y, const double* z, int precision, double* phi, double* ex, double* ey,
                           double* ez, kem_graph** out)
{
    if (!c || !m || !sigma || !out || npts <= 0 || !x || !y || !z || !phi || !ex || !ey || !ez) return KEM_E_ARG;
    *out = nullptr;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (precision != KEM_FP64 && precision != KEM_MIXED) return fail(c, KEM_E_ARG, "field: unknown precision mode");
    KEM_CK(c, enter(c));
    kem_graph* g = new kem_graph();
    g->ctx = c;
    // the graph's own workspace reaches its final size before the capture (no allocation
    // inside a graph); later calls on the context cannot move it
    {
        kem::FieldMesh fm;
        field_mesh(m, fm);
        const bool mixed = precision == KEM_MIXED;
        const WsNeed need =
            field_ws_need(c, m, fm, false, 1, npts, mixed, mixed ? c->far_range_mixed : c->far_range_fp64);
        const hipError_t e = ws_reserve(g->fws, need.ws, need.gsig, need.sort);
        if (e != hipSuccess) {
            kem_graph_destroy(g);
            return fail(c, KEM_E_HIP, std::string("field graph workspace: ") + hipGetErrorString(e));
        }
    }
    KEM_CK(c, hipStreamSynchronize(c->stream));
    const bool timing = c->timing;
    c->timing = false;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    int rc = KEM_OK;
    if (e == hipSuccess) {
        rc = field_eval(c, m, sigma, nullptr, 1, npts, x, y, z, precision, phi, ex, ey, ez, g->fws, false);
        hipGraph_t graph = nullptr;
        const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
        g->graph = graph;
        if (rc == KEM_OK && e2 != hipSuccess) e = e2;
    }
    c->timing = timing;
    if (e == hipSuccess && rc == KEM_OK) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess || rc != KEM_OK) {
        kem_graph_destroy(g);
        if (rc != KEM_OK) return rc;
        return fail(c, KEM_E_HIP, std::string("field graph capture: ") + hipGetErrorString(e));
    }
    *out = g;
    return KEM_OK;
}

int kem_graph_launch(kem_graph* g)
{
    if (!g || !g->exec) return KEM_E_ARG;
    kem_ctx* c = g->ctx;
    KEM_CK(c, enter(c));
    KEM_CK(c, hipGraphLaunch(g->exec, c->stream));
    return KEM_OK;
}

void kem_graph_destroy(kem_graph* g)
{
    if (!g) return;
    if (g->ctx) {
        (void)hipSetDevice(g->ctx->device);
        // a replay still queued on the context's stream reads the graph's workspace
        (void)hipStreamSynchronize(g->ctx->stream);
    }
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
    if (g->graph) (void)hipGraphDestroy(g->graph);
    ws_release(g->fws);
    delete g;
}

int kem_field_eval_fields(kem_ctx* c, int64_t nfields, const kem_mesh* const* meshes, const double* const* sigmas,
                          const int* precisions, int64_t npts, const double* x, const double* y, const double* z,
                          double* phi, double* ex, double* ey, double* ez)
{
    if (!c || nfields < 0 || (nfields > 0 && (!meshes || !sigmas || !precisions))) return KEM_E_ARG;
    if (npts > 0 && (!x || !y || !z || !phi || !ex || !ey || !ez)) return KEM_E_ARG;
    for (int64_t f = 0; f < nfields; f++) {
        if (!meshes[f] || !sigmas[f]) return fail(c, KEM_E_ARG, "fields: null mesh or charge densities");
        if (meshes[f]->ctx != c) return fail(c, KEM_E_STATE, "fields: mesh belongs to another context");
        if (precisions[f] != KEM_FP64 && precisions[f] != KEM_MIXED)
            return fail(c, KEM_E_ARG, "fields: unknown precision mode");
    }
    if (npts <= 0) return KEM_OK;
    KEM_CK(c, enter(c));
    if (nfields == 0) {
        // no field: zero field and potential (the reference's sums start from zero)
        for (double* o : {phi, ex, ey, ez}) KEM_CK(c, hipMemsetAsync(o, 0, sizeof(double) * npts, c->stream));
        return KEM_OK;
    }
    // the first field writes the outputs, each further one is evaluated aside and added
    int rc = field_eval(c, meshes[0], sigmas[0], nullptr, 1, npts, x, y, z, precisions[0], phi, ex, ey, ez, c->fws,
                        true);
    if (rc != KEM_OK) return rc;
    if (nfields > 1 && c->acc4_doubles < 4 * npts) {
        KEM_CK(c, hipStreamSynchronize(c->stream));
        dfree(c->acc4);
        c->acc4_doubles = 0;
        KEM_CK(c, dalloc(&c->acc4, 4 * npts));
        c->acc4_doubles = 4 * npts;
    }
    double* t = c->acc4;
    for (int64_t f = 1; f < nfields; f++) {
        rc = field_eval(c, meshes[f], sigmas[f], nullptr, 1, npts, x, y, z, precisions[f], t, t + npts, t + 2 * npts,
                        t + 3 * npts, c->fws, true);
        if (rc != KEM_OK) return rc;
        KEM_CK(c, kem::launch_accumulate4(c->stream, t, npts, phi, ex, ey, ez));
    }
    return KEM_OK;
}

int kem_field_eval_sets(kem_ctx* c, const kem_mesh* m, int64_t nsets, const double* sigma, const int32_t* set,
                        int64_t npts, const double* x, const double* y, const double* z, int precision, double* phi,
                        double* ex, double* ey, double* ez)
{
    if (!c || !m || !sigma || nsets <= 0 || (npts > 0 && (!x || !y || !z || !set || !phi || !ex || !ey || !ez)))
        return KEM_E_ARG;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (precision != KEM_FP64 && precision != KEM_MIXED) return fail(c, KEM_E_ARG, "field: unknown precision mode");
    if (nsets > (1 << 6)) return fail(c, KEM_E_ARG, "field: at most 64 charge-density sets per call");
    if (npts <= 0) return KEM_OK;
    KEM_CK(c, enter(c));
    KEM_CK(c, ensure_scal(c, 1));
    int* bad = (int*)c->scal;
    int hbad = 0;
    KEM_CK(c, hipMemsetAsync(bad, 0, sizeof(int), c->stream));
    KEM_CK(c, kem::launch_check_sets(c->stream, set, npts, nsets, bad));
    KEM_CK(c, hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    if (hbad) return fail(c, KEM_E_ARG, "field: set index outside [0, nsets)");
    return field_eval(c, m, sigma, set, nsets, npts, x, y, z, precision, phi, ex, ey, ez, c->fws, true);
}

int kem_field_plan(int policy, int clustered, int precision, int multi, int64_t nelem, int64_t nsets, int64_t npts,
                   int64_t range_elems, int64_t ws_doubles, int64_t* out)
{
    if (!out || nelem <= 0 || nsets < 1 || npts < 0 || range_elems < 0) return KEM_E_ARG;
    if (precision != KEM_FP64 && precision != KEM_MIXED) return KEM_E_ARG;
    // the plan only tests which mesh arrays exist: stand-in addresses, never dereferenced
    static const char present = 0;
    kem::FieldMesh fm;
    fm.n = nelem;
    if (clustered) {
        fm.clusters = (const double*)&present;
        fm.order = (const int32_t*)&present;
        fm.nodes32 = (const float*)&present;
        fm.far6 = fm.far3 = (const float*)&present;
        fm.far6_cl = (const double*)&present;
    }
    const bool mixed = precision == KEM_MIXED;
    const int64_t need =
        kem::field_workspace_doubles(policy, fm, mixed, kMixedFp64RatioDefault, multi != 0, nsets, npts, range_elems);
    const kem::FieldPlan p = kem::plan_field(policy, fm, mixed, kMixedFp64RatioDefault, multi != 0, nsets, npts,
                                             range_elems, ws_doubles < 0 ? need : ws_doubles);
    out[0] = ws_doubles < 0 ? need : ws_doubles;
    out[1] = p.weights;
    out[2] = p.partials;
    out[3] = p.ranges;
    out[4] = p.slices;
    out[5] = p.chunk;
    out[6] = p.xcd ? 1 : 0;
    out[7] = p.kind;
    return KEM_OK;
}

int kem_field_census(kem_ctx* c, const kem_mesh* m, int64_t npts, const double* x, const double* y, const double* z,
                     int precision, int64_t* counts)
{
    if (!c || !m || !counts || (npts > 0 && (!x || !y || !z))) return KEM_E_ARG;
    KEM_CK(c, enter(c));
    unsigned long long* d = nullptr;
    KEM_CK(c, dalloc(&d, 12));
    hipError_t e = hipMemsetAsync(d, 0, sizeof(unsigned long long) * 12, c->stream);
    if (e == hipSuccess)
        e = kem::launch_census(c->stream, m->geom, m->nm, x, y, z, npts, precision == KEM_MIXED ? 1 : 0,
                               c->mixed_fp64_ratio, d);
    unsigned long long h[12];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d);
    if (e != hipSuccess) return fail(c, KEM_E_HIP, std::string("census: ") + hipGetErrorString(e));
    for (int k = 0; k < 12; k++) counts[k] = (int64_t)h[k];
    return KEM_OK;
}

int kem_field_census_paths(kem_ctx* c, const kem_mesh* m, int64_t npts, const double* x, const double* y,
                           const double* z, int64_t* counts)
{
    if (!c || !m || !counts || (npts > 0 && (!x || !y || !z))) return KEM_E_ARG;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    for (int k = 0; k < 10; k++) counts[k] = 0;
    if (npts <= 0) return KEM_OK;
    kem::FieldMesh fm;
    field_mesh(m, fm);
    if (c->policy != KEM_POLICY_NUMERIC || !fm.far6 || !fm.far6_cl || !fm.order || c->mixed_fp64_ratio != 0.)
        return fail(c, KEM_E_STATE, "path census: the mesh's mixed passes do not take the cluster far pass");
    KEM_CK(c, enter(c));
    unsigned long long* d = nullptr;
    void* sort = nullptr;
    const bool sorted = npts >= kSortMinPoints && npts <= INT32_MAX;
    const size_t sort_bytes = sorted ? kem::sort_workspace_bytes(npts) : 0;
    hipError_t e = dalloc(&d, 10);
    if (e == hipSuccess && sorted) e = hipMalloc(&sort, sort_bytes);
    if (e == hipSuccess) e = hipMemsetAsync(d, 0, sizeof(unsigned long long) * 10, c->stream);
    if (e == hipSuccess)
        e = sorted ? kem::launch_census_paths_sorted(c->stream, fm, x, y, z, npts, sort, sort_bytes, d)
                   : kem::launch_census_paths(c->stream, fm, x, y, z, npts, d);
    unsigned long long h[10];
    if (e == hipSuccess) e = hipMemcpyAsync(h, d, sizeof(h), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    dfree(d);
    if (sort) (void)hipFree(sort);
    if (e != hipSuccess) return fail(c, KEM_E_HIP, std::string("path census: ") + hipGetErrorString(e));
    for (int k = 0; k < 10; k++) counts[k] = (int64_t)h[k];
    return KEM_OK;
}

int kem_field_eval(kem_ctx* c, const kem_mesh* m, const double* sigma, int64_t npts, const double* P,
                   int precision, double* phi, double* E)
{
    if (!c || !m || !sigma || (npts > 0 && !P)) return KEM_E_ARG;
    if (npts <= 0) return KEM_OK;
    KEM_CK(c, enter(c));
    std::vector<double> soa((size_t)npts * 3);
    for (int64_t i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) soa[(size_t)k * npts + i] = P[3 * i + k];
    double *dS = nullptr, *dP = nullptr, *dO = nullptr;
    hipError_t e = dalloc(&dS, m->n);
    if (e == hipSuccess) e = dalloc(&dP, 3 * npts);
    if (e == hipSuccess) e = dalloc(&dO, 4 * npts);
    if (e == hipSuccess) e = hipMemcpyAsync(dS, sigma, sizeof(double) * m->n, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dP, soa.data(), sizeof(double) * 3 * npts, hipMemcpyHostToDevice, c->stream);
    int rc = KEM_OK;
    if (e == hipSuccess) {
        rc = kem_field_eval_device(c, m, dS, npts, dP, dP + npts, dP + 2 * npts, precision, dO, dO + npts,
                                   dO + 2 * npts, dO + 3 * npts);
        if (rc == KEM_OK) {
            std::vector<double> o((size_t)npts * 4);
            e = hipMemcpyAsync(o.data(), dO, sizeof(double) * 4 * npts, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess) {
                // device layout: phi, Ex, Ey, Ez planes of npts each
                for (int64_t i = 0; i < npts; i++) {
                    if (phi) phi[i] = o[(size_t)i];
                    if (E)
                        for (int k = 0; k < 3; k++) E[3 * i + k] = o[(size_t)(k + 1) * npts + i];
                }
            }
        }
    }
    dfree(dS);
    dfree(dP);
    dfree(dO);
    if (rc != KEM_OK) return rc;
    if (e != hipSuccess) return fail(c, KEM_E_HIP, std::string("field_eval: ") + hipGetErrorString(e));
    return KEM_OK;
}

// ------------------------------------------------------------------ matrix
int kem_matrix_assemble(kem_ctx* c, const kem_mesh* m, int storage, kem_matrix** out)
{
    if (!c || !m || !out) return KEM_E_ARG;
    *out = nullptr;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (!(storage & (KEM_STORE_FP64 | KEM_STORE_FP32 | KEM_STORE_BF16)))
        return fail(c, KEM_E_ARG, "matrix: no storage selected");
    KEM_CK(c, enter(c));
    kem_matrix* A = new kem_matrix();
    A->ctx = c;
    A->n = m->n;
    const int64_t nn = m->n * m->n;
    hipError_t e = hipSuccess;
    if (storage & KEM_STORE_FP64) e = store_alloc(c, &A->A64, nn);
    // BF16 storage is packed from the FP64 matrix, else from the FP32 one (assembled for it
    // and dropped again when FP32 storage was not asked for)
    const bool need32 = (storage & KEM_STORE_FP32) || ((storage & KEM_STORE_BF16) && !(storage & KEM_STORE_FP64));
    if (e == hipSuccess && need32) e = store_alloc(c, &A->A32, nn);
    if (e == hipSuccess && (storage & KEM_STORE_BF16))
        e = store_alloc(c, &A->A16, (int64_t)kem::bf16_matrix_elements(m->n));
    // the potential node table of the split numeric-policy assembly (Dirichlet rows)
    DVec an;
    if (e == hipSuccess && c->policy == 0 && !m->has_neumann && !m->has_axisym && !m->gstart && m->clusters &&
        m->order)
        e = dalloc(&an.p, kem::assemble_nodes_doubles(m->n));
    hipEvent_t t0, t1;
    timing_pair(c, &t0, &t1);
    if (e == hipSuccess && t0) e = hipEventRecord(t0, c->stream);
    if (e == hipSuccess)
        e = kem::launch_assemble(c->stream, c->policy, m->has_neumann, m->geom, m->nodes, m->recs, m->targets, m->n, A->A64,
                                 A->A32, m->has_axisym, m->gstart, m->clusters, m->order, an.p);
    if (e == hipSuccess && t1) e = hipEventRecord(t1, c->stream);
    if (e == hipSuccess && A->A16) e = kem::launch_pack_bf16(c->stream, A->A64, A->A32, m->n, A->A16);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && A->A32 && !(storage & KEM_STORE_FP32)) store_free(c, A->A32, nn);
    if (e != hipSuccess) {
        kem_matrix_destroy(A);
        return fail(c, KEM_E_HIP, std::string("assemble: ") + hipGetErrorString(e));
    }
    *out = A;
    return KEM_OK;
}

void kem_matrix_destroy(kem_matrix* A)
{
    if (!A) return;
    if (A->ctx) (void)hipSetDevice(A->ctx->device);
    // the storage goes back to the context for the next matrix of the same size
    const int64_t nn = A->n * A->n;
    if (A->ctx) {
        store_free(A->ctx, A->A64, nn);
        store_free(A->ctx, A->A32, nn);
        store_free(A->ctx, A->A16, (int64_t)kem::bf16_matrix_elements(A->n));
    } else {
        dfree(A->A64);
        dfree(A->A32);
        dfree(A->A16);
    }
    if (A->bws) (void)hipFree(A->bws);
    if (A->sws) (void)hipFree(A->sws);
    delete A;
}

int64_t kem_matrix_dimension(const kem_matrix* A) { return A ? A->n : 0; }

int kem_matrix_get(kem_matrix* A, int64_t i, int64_t j, double* out)
{
    if (!A || !out) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    if (i < 0 || j < 0 || i >= A->n || j >= A->n) return fail(c, KEM_E_ARG, "matrix index out of range");
    KEM_CK(c, enter(c));
    const int64_t off = j * A->n + i;
    if (A->A64) {
        KEM_CK(c, hipMemcpy(out, A->A64 + off, sizeof(double), hipMemcpyDeviceToHost));
    } else {
        float f;
        KEM_CK(c, hipMemcpy(&f, A->A32 + off, sizeof(float), hipMemcpyDeviceToHost));
        *out = f;
    }
    return KEM_OK;
}

int kem_matrix_copy(kem_matrix* A, double* host)
{
    if (!A || !host) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    if (!A->A64) return fail(c, KEM_E_STATE, "matrix has no FP64 storage");
    KEM_CK(c, enter(c));
    const int64_t n = A->n;
    std::vector<double> cm((size_t)(n * n));
    KEM_CK(c, hipMemcpy(cm.data(), A->A64, sizeof(double) * n * n, hipMemcpyDeviceToHost));
    for (int64_t i = 0; i < n; i++)
        for (int64_t j = 0; j < n; j++) host[i * n + j] = cm[(size_t)(j * n + i)];
    return KEM_OK;
}

namespace {

// y = A x (or A^T x) on device with the matrix's FP64 storage (or FP32 when use32); skip:
// optional device flag that turns the kernels into no-ops (a finished solve in a graph replay)
hipError_t matvec_dev(kem_matrix* A, bool use32, int transpose, const double* x, double* y,
                      const int* skip = nullptr)
{
    kem_ctx* c = A->ctx;
    const int64_t n = A->n;
    if (transpose) {
        if (use32) return kem::launch_gemv_t<float>(c->stream, A->A32, n, n, n, x, y, skip);
        return kem::launch_gemv_t<double>(c->stream, A->A64, n, n, n, x, y, skip);
    }
    hipError_t e = ensure_ws(c, (int64_t)kem::gemv_splits(n, n) * n);
    if (e != hipSuccess) return e;
    if (use32) return kem::launch_gemv<float>(c->stream, A->A32, n, n, n, x, 1.0, 0.0, y, c->ws, skip);
    return kem::launch_gemv<double>(c->stream, A->A64, n, n, n, x, 1.0, 0.0, y, c->ws, skip);
}

}  // namespace

int kem_matrix_multiply(kem_matrix* A, int transpose, const double* x, double* y)
{
    if (!A || !x || !y) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    KEM_CK(c, enter(c));
    const int64_t n = A->n;
    DVec dx, dy;
    KEM_CK(c, dalloc(&dx.p, n));
    KEM_CK(c, dalloc(&dy.p, n));
    KEM_CK(c, hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    KEM_CK(c, ensure_ws(c, (int64_t)kem::gemv_splits(n, n) * n));
    hipEvent_t t0, t1;
    timing_pair(c, &t0, &t1);
    if (t0) KEM_CK(c, hipEventRecord(t0, c->stream));
    KEM_CK(c, matvec_dev(A, A->A64 == nullptr, transpose, dx.p, dy.p));
    if (t1) KEM_CK(c, hipEventRecord(t1, c->stream));
    KEM_CK(c, hipMemcpyAsync(y, dy.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

// ------------------------------------------------------------------ solvers
void kem_solve_opts_default(kem_solve_opts* o)
{
    if (!o) return;
    o->method = KEM_SOLVER_GMRES;
    o->precision = KEM_FP64;
    o->tolerance = 1e-8;  // KIterativeSolver default (KIterativeSolver.hh:14)
    o->max_iterations = 1 << 30;
    o->restart = 0;
    o->check_interval = 0;
    o->time_limit_s = 0.;
    o->inner_tolerance = 1e-5;
}

namespace {

struct SolveCtl {
    double t0 = 0., limit = 0.;
    bool expired() const { return limit > 0. && now_s() - t0 > limit; }
};

// GMRES with Givens rotations and a triangular least-squares solve
// (KGeneralizedMinimalResidual.hh:178-337), convergence |p_{j+1}| / ||b|| <= tol
// (KSimpleIterativeKrylovSolver.hh:84-103).  The Arnoldi step orthogonalises with classical
// Gram-Schmidt plus one reorthogonalisation pass (CGS2): two projection / update pairs over the
// device-resident basis per iteration instead of 2(j+1) dependent dot/axpy launches, and
// orthogonality to working precision (the reference's MGS loses it on BEM matrices, so its
// iteration counts near the tolerance are rounding-dependent).  One iteration -- the
// matrix-vector product, CGS2, normalisation and the Givens rotation -- runs on device state
// (kem::launch_arnoldi_step) and is captured once per solve as a HIP graph; the host replays it
// kGmresReplays times between reads of the column index and stop flag, and a finished cycle
// turns the remaining replays into no-ops.  Device vectors d_b, d_x (x holds x0).  On restart
// (or at the basis cap) x <- x0 + V y.
constexpr int kGmresReplays = 4;

int gmres_batch(kem_matrix* A, int prec, int nrhs, const double* d_B, double* d_X, const std::vector<double>& tol,
                int maxit, int restart, const SolveCtl& ctl, int* iters, std::vector<double>& resid,
                std::vector<int>& conv_out);

int gmres_dev(kem_matrix* A, bool use32, const double* d_b, double* d_x, double tol, int maxit, int restart,
              const SolveCtl& ctl, int* iters, double* resid)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    int kmax = (int)std::min<int64_t>(std::min<int64_t>(maxit, 1000), n + 1);
    if (restart > 0) kmax = std::min(kmax, restart);
    if (kmax < 1) kmax = 1;
    const int ld = kmax + 1;
    const int64_t gparts = 256 + (int64_t)kem::gemv_splits(n, ld) * n;
    // one allocation, carved into 256-byte aligned pieces, kept on the matrix between solves
    // (the Krylov basis of a 100,000-unknown solve is 800 MB)
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    const size_t oV = carve(sizeof(double) * (size_t)ld * n);
    const size_t ow = carve(sizeof(double) * (size_t)n);
    const size_t ovc = carve(sizeof(double) * (size_t)n);
    const size_t oH = carve(sizeof(double) * (size_t)ld * ld);
    const size_t orot = carve(sizeof(double) * 3 * (size_t)(kmax + 2));
    const size_t ohh = carve(sizeof(double) * (2 * (size_t)(kmax + 8) + 8));
    const size_t opart = carve(sizeof(double) * (size_t)(gparts + (int64_t)ld * kem::kArnProjSplits));
    const size_t oscal = carve(sizeof(double) * 8);
    const size_t oints = carve(sizeof(int) * 4);
    if (A->sws_bytes < off) {
        if (A->sws) (void)hipFree(A->sws);
        A->sws = nullptr;
        A->sws_bytes = 0;
        KEM_CK(c, hipMalloc(&A->sws, off));
        A->sws_bytes = off;
    }
    char* base = (char*)A->sws;
    struct {
        double* p;
    } V{(double*)(base + oV)}, w{(double*)(base + ow)}, vcur{(double*)(base + ovc)}, H{(double*)(base + oH)},
        rot{(double*)(base + orot)}, hh{(double*)(base + ohh)}, part{(double*)(base + opart)},
        scal{(double*)(base + oscal)};
    int* const d_ints = (int*)(base + oints);
    KEM_CK(c, ensure_ws(c, (int64_t)kem::gemv_splits(n, n) * n));
    kem::ArnoldiState a;
    a.n = n;
    a.kmax = kmax;
    a.V = V.p;
    a.w = w.p;
    a.vcur = vcur.p;
    a.H = H.p;
    a.cs = rot.p;
    a.sn = rot.p + (kmax + 2);
    a.p = rot.p + 2 * (kmax + 2);
    a.h1 = hh.p;
    a.h2 = hh.p + (kmax + 8);
    a.nrm = hh.p + 2 * (kmax + 8);
    a.part = part.p;
    a.hp = part.p + gparts;
    a.ints = d_ints;
    a.scal = scal.p;
    double bnorm;
    KEM_CK(c, kem::launch_dot(st, d_b, d_b, n, part.p, scal.p + 2, true));
    KEM_CK(c, hipMemcpyAsync(&bnorm, scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, st));
    KEM_CK(c, hipStreamSynchronize(st));
    if (bnorm == 0.) bnorm = 1.;

    // one iteration as a graph (launched directly if the stream cannot be captured)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct GraphGuard {
        hipGraph_t* g;
        hipGraphExec_t* e;
        ~GraphGuard()
        {
            if (*e) (void)hipGraphExecDestroy(*e);
            if (*g) (void)hipGraphDestroy(*g);
        }
    } graph_guard{&graph, &exec};
    auto launch_step = [&]() -> hipError_t {
        hipError_t e = matvec_dev(A, use32, A->trans, vcur.p, w.p, d_ints + 1);
        if (e == hipSuccess) e = kem::launch_arnoldi_step(st, a);
        return e;
    };
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const hipError_t e1 = launch_step();
        const hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (e1 != hipSuccess || e2 != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
        }
    }
    (void)hipGetLastError();

    std::vector<double> Hh((size_t)ld * ld), p(kmax + 2), y(kmax + 1);
    int it = 0;
    double rn = 0.;
    int status = KEM_OK;
    for (;;) {
        // r = b - A x ; v0 = r / ||r||
        KEM_CK(c, matvec_dev(A, use32, A->trans, d_x, w.p));
        KEM_CK(c, kem::launch_axpby(st, 1.0, d_b, -1.0, w.p, w.p, n));
        KEM_CK(c, kem::launch_dot(st, w.p, w.p, n, part.p, scal.p + 2, true));
        KEM_CK(c, kem::launch_div_dev(st, scal.p + 2, w.p, V.p, n));
        KEM_CK(c, hipMemcpyAsync(vcur.p, V.p, sizeof(double) * n, hipMemcpyDeviceToDevice, st));
        KEM_CK(c, hipMemcpyAsync(&rn, scal.p + 2, sizeof(double), hipMemcpyDeviceToHost, st));
        KEM_CK(c, hipStreamSynchronize(st));
        if (!(rn > 0.)) break;  // x is already exact
        const int jlimit = std::min(kmax, maxit - it);
        int hints[4] = {0, 0, jlimit, 0};
        const double hs[2] = {tol * bnorm, rn};
        KEM_CK(c, hipMemcpyAsync(d_ints, hints, sizeof(hints), hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(scal.p, hs, sizeof(hs), hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.p, &rn, sizeof(double), hipMemcpyHostToDevice, st));
        bool expired = false;
        for (;;) {
            for (int k = 0; k < kGmresReplays; k++) KEM_CK(c, exec ? hipGraphLaunch(exec, st) : launch_step());
            double last[2];
            KEM_CK(c, hipMemcpyAsync(hints, d_ints, sizeof(hints), hipMemcpyDeviceToHost, st));
            KEM_CK(c, hipMemcpyAsync(last, scal.p, sizeof(last), hipMemcpyDeviceToHost, st));
            KEM_CK(c, hipStreamSynchronize(st));
            rn = last[1];
            if (hints[1]) break;
            if (ctl.expired()) {
                expired = true;
                break;
            }
        }
        const int j = hints[0];
        it += j;
        const bool conv = hints[3] != 0;
        // x += V y with H y = p (upper triangular)
        if (j > 0) {
            KEM_CK(c, hipMemcpyAsync(Hh.data(), H.p, sizeof(double) * (size_t)ld * j, hipMemcpyDeviceToHost, st));
            KEM_CK(c, hipMemcpyAsync(p.data(), a.p, sizeof(double) * j, hipMemcpyDeviceToHost, st));
            KEM_CK(c, hipStreamSynchronize(st));
            for (int r = j - 1; r >= 0; r--) {
                double s = p[r];
                for (int q = r + 1; q < j; q++) s -= Hh[(size_t)q * ld + r] * y[q];
                y[r] = s / Hh[(size_t)r * ld + r];
            }
            KEM_CK(c, hipMemcpyAsync(hh.p, y.data(), sizeof(double) * j, hipMemcpyHostToDevice, st));
            KEM_CK(c, kem::launch_gemv<double>(st, V.p, n, n, j, hh.p, 1.0, 1.0, d_x, c->ws));
            KEM_CK(c, hipStreamSynchronize(st));
        }
        if (conv) break;
        // stopped without convergence: iteration limit, time limit or breakdown; otherwise the
        // basis is full and the cycle restarts from the new x
        if (expired || it >= maxit || j < jlimit) {
            status = KEM_W_NOT_CONVERGED;
            break;
        }
    }
    *iters = it;
    *resid = rn / bnorm;
    return status;
}

// ||b - A64 x|| / ||b|| on device
int true_residual(kem_matrix* A, const double* d_b, const double* d_x, double* d_tmp, double* out)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    KEM_CK(c, ensure_scal(c, 260));
    KEM_CK(c, matvec_dev(A, A->A64 == nullptr, A->trans, d_x, d_tmp));
    KEM_CK(c, kem::launch_axpby(st, 1.0, d_b, -1.0, d_tmp, d_tmp, n));
    KEM_CK(c, kem::launch_dot(st, d_tmp, d_tmp, n, c->scal, c->scal + 256, true));
    KEM_CK(c, kem::launch_dot(st, d_b, d_b, n, c->scal, c->scal + 257, true));
    double v[2];
    KEM_CK(c, hipMemcpyAsync(v, c->scal + 256, sizeof(double) * 2, hipMemcpyDeviceToHost, st));
    KEM_CK(c, hipStreamSynchronize(st));
    *out = v[0] / (v[1] > 0. ? v[1] : 1.);
    return KEM_OK;
}

int robin_hood_dev(kem_matrix* A, const double* d_b, double* d_x, const std::vector<double>& hb,
                   const std::vector<double>& hx, const kem_solve_opts* o, const SolveCtl& ctl, kem_solve_stats* s)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    if (!A->A64) return fail(c, KEM_E_STATE, "Robin Hood needs FP64 matrix storage");
    DVec bi;
    KEM_CK(c, dalloc(&bi.p, n));
    KEM_CK(c, hipMemsetAsync(bi.p, 0, sizeof(double) * n, st));
    double xinf = 0., binf = 0.;
    for (int64_t i = 0; i < n; i++) {
        xinf = std::max(xinf, std::fabs(hx[i]));
        binf = std::max(binf, std::fabs(hb[i]));
    }
    if (xinf > 1.e-16) KEM_CK(c, matvec_dev(A, false, 0, d_x, bi.p));
    if (binf == 0.) binf = 1.;
    const int64_t interval = o->check_interval > 0 ? o->check_interval : n;
    KEM_CK(c, ensure_scal(c, 260));
    // the several-workgroup sub-iterations' slots and barrier flags
    DVec slots;
    unsigned int* flags = nullptr;
    const bool grid = kem::robin_hood_groups(n) > 1;
    if (grid) {
        KEM_CK(c, dalloc(&slots.p, 2 * 256 * 3));
        KEM_CK(c, dalloc(&flags, 2));
    }
    struct Flags {
        unsigned int*& p;
        ~Flags() { dfree(p); }
    } flags_guard{flags};
    double rnorm = 1.;  // KIterativeSolver::fResidualNorm starts at 1
    int checks = 0;
    int status = KEM_OK;
    while (rnorm > o->tolerance) {
        if (checks >= o->max_iterations || ctl.expired()) {
            status = KEM_W_NOT_CONVERGED;
            break;
        }
        KEM_CK(c, kem::launch_robin_hood(st, A->A64, d_b, d_x, bi.p, n, interval, c->scal + 256, slots.p, flags));
        double v;
        unsigned int hf[2] = {0, 0};
        KEM_CK(c, hipMemcpyAsync(&v, c->scal + 256, sizeof(double), hipMemcpyDeviceToHost, st));
        if (grid) KEM_CK(c, hipMemcpyAsync(hf, flags, sizeof(hf), hipMemcpyDeviceToHost, st));
        KEM_CK(c, hipStreamSynchronize(st));
        if (hf[1]) return fail(c, KEM_E_HIP, "Robin Hood: a workgroup timed out at the grid barrier");
        rnorm = v / binf;
        checks++;
    }
    s->iterations = checks;
    s->sub_iterations = (int64_t)checks * interval;
    s->residual = rnorm;
    return status;
}

}  // namespace

int kem_solve(kem_matrix* A, const kem_solve_opts* opts, const double* b, double* x, kem_solve_stats* stats)
{
    if (!A || !b || !x) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    kem_solve_opts o;
    kem_solve_opts_default(&o);
    if (opts) o = *opts;
    kem_solve_stats s{};
    if (o.max_iterations <= 0) return fail(c, KEM_E_ARG, "max_iterations must be positive");
    if (!(o.tolerance > 0.)) return fail(c, KEM_E_ARG, "tolerance must be positive");
    KEM_CK(c, enter(c));
    const int64_t n = A->n;
    SolveCtl ctl;
    ctl.t0 = now_s();
    ctl.limit = o.time_limit_s;
    DVec db, dx, dr, dd;
    KEM_CK(c, dalloc(&db.p, n));
    KEM_CK(c, dalloc(&dx.p, n));
    KEM_CK(c, hipMemcpyAsync(db.p, b, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    KEM_CK(c, hipMemcpyAsync(dx.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    int status = KEM_OK;
    if (o.method == KEM_SOLVER_ROBIN_HOOD) {
        std::vector<double> hb(b, b + n), hx(x, x + n);
        status = robin_hood_dev(A, db.p, dx.p, hb, hx, &o, ctl, &s);
    } else if (o.method == KEM_SOLVER_GMRES) {
        if (o.precision == KEM_FP64) {
            if (!A->A64) return fail(c, KEM_E_STATE, "FP64 solve needs FP64 matrix storage");
            status = gmres_dev(A, false, db.p, dx.p, o.tolerance, o.max_iterations, o.restart, ctl, &s.iterations,
                               &s.residual);
        } else if (o.precision == KEM_MIXED || o.precision == KEM_MIXED_BF16) {
            // FP32- (or BF16-) matrix GMRES inner solves, FP64 residual refinement
            const bool bf16 = o.precision == KEM_MIXED_BF16;
            if (!bf16 && !A->A32) return fail(c, KEM_E_STATE, "mixed solve needs FP32 matrix storage");
            if (bf16 && !A->A16) return fail(c, KEM_E_STATE, "BF16 solve needs BF16 matrix storage");
            if (bf16 && A->trans) return fail(c, KEM_E_ARG, "BF16 solves run on A, not on its transpose");
            if (!A->A64) return fail(c, KEM_E_STATE, "mixed solve needs FP64 matrix storage for the residual");
            KEM_CK(c, dalloc(&dr.p, n));
            KEM_CK(c, dalloc(&dd.p, n));
            double rel = 0.;
            for (int ref = 0; ref < 50; ref++) {
                int rc = true_residual(A, db.p, dx.p, dr.p, &rel);
                if (rc != KEM_OK) return rc;
                // dr holds r = b - A64 x
                if (rel <= o.tolerance) break;
                if (s.iterations >= o.max_iterations || ctl.expired()) {
                    status = KEM_W_NOT_CONVERGED;
                    break;
                }
                KEM_CK(c, hipMemsetAsync(dd.p, 0, sizeof(double) * n, c->stream));
                int its = 0;
                double res = 0.;
                double inner = std::max(o.inner_tolerance, 0.5 * o.tolerance / std::max(rel, 1e-300));
                if (bf16) inner = std::max(inner, kBf16InnerTolerance);
                if (bf16) {
                    // the lockstep solver with one right-hand side on BF16 MFMA block products
                    std::vector<double> tol1(1, std::min(inner, 0.5)), res1;
                    std::vector<int> conv1;
                    rc = gmres_batch(A, kem::kBmmBf16, 1, dr.p, dd.p, tol1, o.max_iterations - s.iterations, 0,
                                     ctl, &its, res1, conv1);
                } else {
                    rc = gmres_dev(A, true, dr.p, dd.p, std::min(inner, 0.5), o.max_iterations - s.iterations, 0,
                                   ctl, &its, &res);
                }
                if (rc < 0) return rc;
                s.iterations += its;
                s.refinements++;
                KEM_CK(c, kem::launch_axpby(c->stream, 1.0, dx.p, 1.0, dd.p, dx.p, n));
            }
            s.residual = rel;
            if (rel > o.tolerance) status = KEM_W_NOT_CONVERGED;
        } else {
            return fail(c, KEM_E_ARG, "unknown precision mode");
        }
    } else if (o.method == KEM_SOLVER_GAUSS) {
        // dense direct solve on a copy of the FP64 matrix (KGaussianEliminationChargeDensitySolver)
        if (!A->A64) return fail(c, KEM_E_STATE, "Gaussian elimination needs FP64 matrix storage");
        if (n > 32768) return fail(c, KEM_E_ARG, "Gaussian elimination is limited to n <= 32768");
        DVec M, f;
        KEM_CK(c, dalloc(&M.p, n * (n + 1)));
        KEM_CK(c, dalloc(&f.p, n));
        KEM_CK(c, hipMemcpyAsync(M.p, A->A64, sizeof(double) * n * n, hipMemcpyDeviceToDevice, c->stream));
        KEM_CK(c, hipMemcpyAsync(M.p + n * n, db.p, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
        KEM_CK(c, kem::launch_gauss_jordan(c->stream, M.p, n, f.p, dx.p));
        KEM_CK(c, hipStreamSynchronize(c->stream));  // before M and f are released
        s.iterations = 1;
    } else {
        return fail(c, KEM_E_NAME, "unknown solver method");
    }
    if (status < 0) return status;
    KEM_CK(c, hipMemcpyAsync(x, dx.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    if (A->A64) {
        DVec tmp;
        KEM_CK(c, dalloc(&tmp.p, n));
        int rc = true_residual(A, db.p, dx.p, tmp.p, &s.true_residual);
        if (rc != KEM_OK) return rc;
    }
    KEM_CK(c, hipStreamSynchronize(c->stream));
    s.seconds = now_s() - ctl.t0;
    if (stats) *stats = s;
    return status;
}


// ------------------------------------------------------------------ adjoint
int kem_solve_transposed(kem_matrix* A, const kem_solve_opts* opts, const double* b, double* x,
                         kem_solve_stats* stats)
{
    if (!A) return KEM_E_ARG;
    kem_solve_opts o;
    kem_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.method != KEM_SOLVER_GMRES) return fail(A->ctx, KEM_E_ARG, "transposed solve supports GMRES only");
    A->trans = 1;
    const int rc = kem_solve(A, &o, b, x, stats);
    A->trans = 0;
    return rc;
}

namespace {
int vjp_checks(kem_ctx* c, const kem_mesh* m)
{
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (c->policy != KEM_POLICY_NUMERIC) return fail(c, KEM_E_STATE, "derivatives need the numeric integrator policy");
    if (!m->all_triangles || m->gstart)
        return fail(c, KEM_E_GEOM, "derivatives are implemented for triangles without symmetry groups only");
    return KEM_OK;
}
}  // namespace

int kem_field_vjp(kem_ctx* c, const kem_mesh* m, const double* sigma, int64_t npts, const double* P,
                  const double* gphi, double* dsigma, double* dvert)
{
    if (!c || !m || !sigma || !dsigma || !dvert || npts < 0 || (npts > 0 && (!P || !gphi))) return KEM_E_ARG;
    int rc = vjp_checks(c, m);
    if (rc != KEM_OK) return rc;
    KEM_CK(c, enter(c));
    const int64_t n = m->n;
    DVec ds, dP, dg, o1, o2, ws;
    const size_t wsd = kem::vjp_workspace_doubles(n, npts);
    KEM_CK(c, dalloc(&ds.p, n));
    KEM_CK(c, dalloc(&dP.p, 3 * std::max<int64_t>(npts, 1)));
    KEM_CK(c, dalloc(&dg.p, std::max<int64_t>(npts, 1)));
    KEM_CK(c, dalloc(&o1.p, n));
    KEM_CK(c, dalloc(&o2.p, 9 * n));
    KEM_CK(c, dalloc(&ws.p, (int64_t)wsd));
    KEM_CK(c, hipMemcpyAsync(ds.p, sigma, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    if (npts > 0) {
        KEM_CK(c, hipMemcpyAsync(dP.p, P, sizeof(double) * 3 * npts, hipMemcpyHostToDevice, c->stream));
        KEM_CK(c, hipMemcpyAsync(dg.p, gphi, sizeof(double) * npts, hipMemcpyHostToDevice, c->stream));
    }
    KEM_CK(c, kem::launch_field_vjp(c->stream, m->geom, ds.p, m->rule, n, dP.p, dg.p, npts, o1.p, o2.p, ws.p, wsd));
    KEM_CK(c, hipMemcpyAsync(dsigma, o1.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipMemcpyAsync(dvert, o2.p, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

int kem_assemble_vjp(kem_ctx* c, const kem_mesh* m, const double* lambda, const double* sigma, double* dvert)
{
    if (!c || !m || !lambda || !sigma || !dvert) return KEM_E_ARG;
    int rc = vjp_checks(c, m);
    if (rc != KEM_OK) return rc;
    if (m->has_neumann) return fail(c, KEM_E_STATE, "matrix derivatives are implemented for Dirichlet rows only");
    KEM_CK(c, enter(c));
    const int64_t n = m->n;
    std::vector<double> neg(lambda, lambda + n);
    for (auto& v : neg) v = -v;
    DVec dl, du, o2, ws;
    const size_t wsd = kem::vjp_workspace_doubles(n, 0);
    KEM_CK(c, dalloc(&dl.p, n));
    KEM_CK(c, dalloc(&du.p, n));
    KEM_CK(c, dalloc(&o2.p, 9 * n));
    KEM_CK(c, dalloc(&ws.p, (int64_t)wsd));
    KEM_CK(c, hipMemcpyAsync(dl.p, neg.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    KEM_CK(c, hipMemcpyAsync(du.p, sigma, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    KEM_CK(c, kem::launch_matrix_vjp(c->stream, m->geom, m->targets, m->rule, n, dl.p, du.p, o2.p, ws.p, wsd));
    KEM_CK(c, hipMemcpyAsync(dvert, o2.p, sizeof(double) * 9 * n, hipMemcpyDeviceToHost, c->stream));
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

// ------------------------------------------------------------------ multi-RHS
namespace {

// panel width of the block product for nrhs right-hand sides

// Y = A X for up to 64 right-hand sides (device, column-major with leading dimension n)
// the block-product precision of a solve precision: FP64 matrix, FP32 matrix or BF16 matrix
int bmm_prec(int precision)
{
    return precision == KEM_MIXED_BF16 ? kem::kBmmBf16 : precision == KEM_MIXED ? kem::kBmmFp32 : kem::kBmmFp64;
}

// Y = A X into a workspace of at least kem::bmm_workspace_bytes(n, nrhs, prec) bytes
int block_product_in(kem_matrix* A, int prec, int nrhs, const double* d_X, double* d_Y, void* ws, size_t ws_bytes)
{
    kem_ctx* c = A->ctx;
    if (prec == kem::kBmmBf16)
        KEM_CK(c, kem::launch_bmm16(c->stream, A->A16, A->n, d_X, A->n, nrhs, d_Y, A->n, ws, ws_bytes, false));
    else
        KEM_CK(c, kem::launch_bmm(c->stream, prec == kem::kBmmFp32 ? A->A32 : nullptr,
                                  prec == kem::kBmmFp32 ? nullptr : A->A64, A->n, d_X, A->n, nrhs, d_Y, A->n, ws,
                                  ws_bytes, false));
    return KEM_OK;
}

// as block_product_in, growing the caller's workspace *ws (*ws_bytes bytes) first if needed
int block_product(kem_matrix* A, int prec, int nrhs, const double* d_X, double* d_Y, void** ws, size_t* ws_bytes)
{
    kem_ctx* c = A->ctx;
    const size_t need = kem::bmm_workspace_bytes(A->n, nrhs, prec);
    if (*ws_bytes < need) {
        if (*ws) (void)hipFree(*ws);
        *ws = nullptr;
        *ws_bytes = 0;
        KEM_CK(c, hipMalloc(ws, need));
        *ws_bytes = need;
    }
    return block_product_in(A, prec, nrhs, d_X, d_Y, *ws, *ws_bytes);
}

struct DBuf {
    void* p = nullptr;
    size_t bytes = 0;
    ~DBuf()
    {
        if (p) (void)hipFree(p);
    }
};

// Lockstep GMRES over nrhs <= 64 right-hand sides: each keeps its own Krylov space, Hessenberg
// matrix and Givens rotations (the single-RHS algorithm of gmres_dev, CGS2, x <- x0 + V y), and
// one block product A [v_1 .. v_k] per iteration serves all of them (FP32 matrix: MFMA).  A
// right-hand side leaves the lockstep when it converges (|p_{j+1}| <= tol_r ||b_r||) or breaks
// down.  d_B, d_X: n x nrhs column-major; X holds x0 on entry.
int gmres_batch(kem_matrix* A, int prec, int nrhs, const double* d_B, double* d_X, const std::vector<double>& tol,
                int maxit, int restart, const SolveCtl& ctl, int* iters, std::vector<double>& resid,
                std::vector<int>& conv_out)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    int kmax = (int)std::min<int64_t>(std::min<int64_t>(maxit, 300), n + 1);
    if (restart > 0) kmax = std::min(kmax, restart);
    if (kmax < 1) kmax = 1;
    const int ldh = kmax + 2;
    const int kp = kem::bmm_panel_cols(n, nrhs, prec);
    const int64_t vstride = (int64_t)(kmax + 1) * n;
    // the panel's rows past n (BF16: the last group of 8) stay zero
    const size_t panel_bytes = prec == kem::kBmmBf16 ? sizeof(uint16_t) * (size_t)((n + 7) / 8) * 8 * kp
                                                     : (prec == kem::kBmmFp32 ? sizeof(float) : sizeof(double)) *
                                                           (size_t)n * kp;
    const size_t wsb = kem::bmm_workspace_bytes(n, nrhs, prec);
    // one allocation, carved into 256-byte aligned pieces, kept on the matrix between solves
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
        const size_t at = off;
        off += (bytes + 255) & ~(size_t)255;
        return at;
    };
    // the Krylov bases of the FP32 / BF16 inner solves in FP32 (they solve to >= 1e-5; the
    // orthogonalisation passes move half the bytes), of FP64 solves in FP64
    const bool v32 = prec != kem::kBmmFp64;
    const size_t oV = carve((v32 ? sizeof(float) : sizeof(double)) * (size_t)(vstride * nrhs));
    const size_t oW = carve(sizeof(double) * (size_t)(n * nrhs));
    const size_t oh1 = carve(sizeof(double) * (size_t)ldh * nrhs);
    const size_t oh2 = carve(sizeof(double) * (size_t)ldh * nrhs);
    const size_t oY = carve(sizeof(double) * (size_t)ldh * nrhs);
    const size_t oH = carve(sizeof(double) * (size_t)nrhs * ldh * ldh);
    const size_t orot = carve(sizeof(double) * 3 * (size_t)ldh * nrhs);
    const size_t ovec = carve(sizeof(double) * 4 * (size_t)nrhs);
    const size_t opanel = carve(panel_bytes);
    const size_t oints = carve(sizeof(int) * (4 * (size_t)nrhs + 4));
    const size_t ows = carve(wsb);
    const size_t oproj = carve(sizeof(double) * (size_t)kem::batch_proj_doubles(n, nrhs, ldh));
    if (A->bws_bytes < off) {
        if (A->bws) (void)hipFree(A->bws);
        A->bws = nullptr;
        A->bws_bytes = 0;
        KEM_CK(c, hipMalloc(&A->bws, off));
        A->bws_bytes = off;
    }
    char* base = (char*)A->bws;
    double* const V = v32 ? nullptr : (double*)(base + oV);
    float* const V32 = v32 ? (float*)(base + oV) : nullptr;
    double* const W = (double*)(base + oW);
    double* const h1 = (double*)(base + oh1);
    double* const h2 = (double*)(base + oh2);
    double* const Y = (double*)(base + oY);
    double* const H = (double*)(base + oH);
    double* const rot = (double*)(base + orot);
    double* const vec = (double*)(base + ovec);
    void* const panel = base + opanel;
    void* const ints = base + oints;
    void* const ws = base + ows;
    KEM_CK(c, hipMemsetAsync(panel, 0, panel_bytes, st));
    kem::BatchArnoldi a;
    a.n = n;
    a.nrhs = nrhs;
    a.kmax = kmax;
    a.ldh = ldh;
    a.kp = kp;
    a.use32 = prec == kem::kBmmFp32;
    a.A32 = A->A32;
    a.A64 = A->A64;
    a.V = V;
    a.V32 = V32;
    a.vstride = vstride;
    a.W = W;
    a.A16 = prec == kem::kBmmBf16 ? A->A16 : nullptr;
    a.P32 = prec == kem::kBmmFp32 ? (float*)panel : nullptr;
    a.P64 = prec == kem::kBmmFp64 ? (double*)panel : nullptr;
    a.P16 = prec == kem::kBmmBf16 ? (uint16_t*)panel : nullptr;
    a.ws = ws;
    a.ws_bytes = wsb;
    a.h1 = h1;
    a.h2 = h2;
    a.H = H;
    a.cs = rot;
    a.sn = rot + (int64_t)ldh * nrhs;
    a.p = rot + 2 * (int64_t)ldh * nrhs;
    a.nrm = vec;
    a.bnorm = vec + nrhs;
    a.tol = vec + 2 * nrhs;
    a.resid = vec + 3 * nrhs;
    a.active = (int*)ints;
    a.done = a.active + nrhs;
    a.jr = a.done + nrhs;
    a.cnt = a.jr + nrhs;
    a.ints = a.cnt + nrhs;
    a.proj = (double*)(base + oproj);

    std::vector<double> bnorm(nrhs), hp((size_t)ldh * nrhs, 0.), hh1(nrhs);
    std::vector<int> active(nrhs), done(nrhs, 0), ones(nrhs, 1);
    // the norms land ldh apart (h1's layout); one contiguous copy of that span reads them back (a
    // strided 2D copy into pageable memory took 2.6 ms per call)
    const size_t span = (size_t)(nrhs - 1) * ldh + 1;
    std::vector<double> stage(span);
    KEM_CK(c, kem::launch_bnorm(st, d_B, n, nrhs, h1, ldh));
    KEM_CK(c, hipMemcpyAsync(stage.data(), h1, sizeof(double) * span, hipMemcpyDeviceToHost, st));
    KEM_CK(c, hipStreamSynchronize(st));
    for (int r = 0; r < nrhs; r++) bnorm[r] = stage[(size_t)r * ldh];
    for (auto& b : bnorm)
        if (b == 0.) b = 1.;
    KEM_CK(c, hipMemcpyAsync(a.bnorm, bnorm.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, st));
    KEM_CK(c, hipMemcpyAsync(a.tol, tol.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, st));
    KEM_CK(c, hipMemcpyAsync(a.done, done.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));

    // one lockstep iteration as a graph (launched directly if the stream cannot be captured)
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    struct GraphGuard {
        hipGraph_t* g;
        hipGraphExec_t* e;
        ~GraphGuard()
        {
            if (*e) (void)hipGraphExecDestroy(*e);
            if (*g) (void)hipGraphDestroy(*g);
        }
    } graph_guard{&graph, &exec};
    if (hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal) == hipSuccess) {
        const hipError_t e1 = kem::launch_batch_arnoldi_step(st, a);
        const hipError_t e2 = hipStreamEndCapture(st, &graph);
        if (e1 != hipSuccess || e2 != hipSuccess || hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (exec) (void)hipGraphExecDestroy(exec);
            if (graph) (void)hipGraphDestroy(graph);
            exec = nullptr;
            graph = nullptr;
        }
    }
    (void)hipGetLastError();

    resid.assign(nrhs, 0.);
    int it = 0, status = KEM_OK;
    for (;;) {
        // W = B - A X; V_0 = W / ||W||
        int rc = block_product_in(A, prec, nrhs, d_X, W, ws, wsb);
        if (rc != KEM_OK) return rc;
        KEM_CK(c, kem::launch_baxpby(st, 1.0, d_B, -1.0, W, W, n * nrhs));
        KEM_CK(c, kem::launch_bnorm(st, W, n, nrhs, h1, ldh));
        KEM_CK(c, hipMemcpyAsync(stage.data(), h1, sizeof(double) * span, hipMemcpyDeviceToHost, st));
        KEM_CK(c, hipStreamSynchronize(st));
        for (int r = 0; r < nrhs; r++) hh1[r] = stage[(size_t)r * ldh];
        int nact = 0;
        for (int r = 0; r < nrhs; r++) {
            hp[(size_t)r * ldh] = hh1[r];
            resid[r] = hh1[r] / bnorm[r];
            active[r] = !done[r] && hh1[r] > 0. && resid[r] > tol[r];
            if (!active[r]) done[r] = 1;
            nact += active[r];
        }
        if (nact == 0) break;
        const int jlimit = std::min(kmax, maxit - it);
        const int hints[4] = {0, 0, jlimit, 0};
        KEM_CK(c, hipMemcpyAsync(a.active, active.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.done, done.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.cnt, ones.data(), sizeof(int) * nrhs, hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.ints, hints, sizeof(hints), hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.p, hp.data(), sizeof(double) * ldh * nrhs, hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemcpyAsync(a.resid, resid.data(), sizeof(double) * nrhs, hipMemcpyHostToDevice, st));
        KEM_CK(c, hipMemsetAsync(a.jr, 0, sizeof(int) * nrhs, st));
        KEM_CK(c, v32 ? kem::launch_bnormalize(st, W, h1, ldh, a.active, n, nrhs, V32, vstride, 0, kp, a.P32, a.P64,
                                               a.P16)
                      : kem::launch_bnormalize(st, W, h1, ldh, a.active, n, nrhs, V, vstride, 0, kp, a.P32, a.P64,
                                               a.P16));
        int got[4] = {0, 0, 0, 0};
        bool expired = false;
        for (;;) {
            for (int k = 0; k < kGmresReplays; k++)
                KEM_CK(c, exec ? hipGraphLaunch(exec, st) : kem::launch_batch_arnoldi_step(st, a));
            KEM_CK(c, hipMemcpyAsync(got, a.ints, sizeof(got), hipMemcpyDeviceToHost, st));
            KEM_CK(c, hipStreamSynchronize(st));
            if (got[1]) break;
            if (ctl.expired()) {
                expired = true;
                break;
            }
        }
        it += got[3];
        // x_r += V_r y_r with H_r y_r = p_r (upper triangular, jr[r] columns)
        KEM_CK(c, kem::launch_batch_trisolve(st, a, Y));
        KEM_CK(c, v32 ? kem::launch_bupdate(st, V32, vstride, Y, ldh, a.cnt, 1.0, n, nrhs, d_X)
                      : kem::launch_bupdate(st, V, vstride, Y, ldh, a.cnt, 1.0, n, nrhs, d_X));
        KEM_CK(c, hipMemcpyAsync(done.data(), a.done, sizeof(int) * nrhs, hipMemcpyDeviceToHost, st));
        KEM_CK(c, hipMemcpyAsync(resid.data(), a.resid, sizeof(double) * nrhs, hipMemcpyDeviceToHost, st));
        KEM_CK(c, hipStreamSynchronize(st));
        bool all = true;
        for (int r = 0; r < nrhs; r++) all = all && done[r];
        if (all) break;
        if (expired || it >= maxit) {
            status = KEM_W_NOT_CONVERGED;
            break;
        }
    }
    conv_out.assign(done.begin(), done.end());
    *iters = it;
    return status;
}

// ||b_r - A64 x_r|| / ||b_r|| for all r; R receives the residual vectors
int block_residual(kem_matrix* A, int nrhs, const double* d_B, const double* d_X, double* d_R, std::vector<double>& rel,
                   void** ws, size_t* wsb)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    int rc = block_product(A, A->A64 ? kem::kBmmFp64 : kem::kBmmFp32, nrhs, d_X, d_R, ws, wsb);
    if (rc != KEM_OK) return rc;
    KEM_CK(c, kem::launch_baxpby(st, 1.0, d_B, -1.0, d_R, d_R, n * nrhs));
    KEM_CK(c, ensure_scal(c, 2 * (int64_t)nrhs));
    KEM_CK(c, kem::launch_bnorm(st, d_R, n, nrhs, c->scal, 1));
    KEM_CK(c, kem::launch_bnorm(st, d_B, n, nrhs, c->scal + nrhs, 1));
    std::vector<double> v(2 * (size_t)nrhs);
    KEM_CK(c, hipMemcpyAsync(v.data(), c->scal, sizeof(double) * 2 * nrhs, hipMemcpyDeviceToHost, st));
    KEM_CK(c, hipStreamSynchronize(st));
    rel.resize(nrhs);
    for (int r = 0; r < nrhs; r++) rel[r] = v[r] / (v[nrhs + r] > 0. ? v[nrhs + r] : 1.);
    return KEM_OK;
}

// one group of <= 64 right-hand sides (device B, X)
int solve_group(kem_matrix* A, const kem_solve_opts& o, int nrhs, const double* d_B, double* d_X,
                const SolveCtl& ctl, kem_solve_stats* s, std::vector<double>& rel)
{
    kem_ctx* c = A->ctx;
    hipStream_t st = c->stream;
    const int64_t n = A->n;
    int its = 0;
    std::vector<int> conv;
    if (o.precision == KEM_FP64) {
        if (!A->A64) return fail(c, KEM_E_STATE, "FP64 solve needs FP64 matrix storage");
        std::vector<double> tol(nrhs, o.tolerance);
        int status =
            gmres_batch(A, kem::kBmmFp64, nrhs, d_B, d_X, tol, o.max_iterations, o.restart, ctl, &its, rel, conv);
        if (status < 0) return status;
        s->iterations = std::max(s->iterations, its);
        return status;
    }
    const int inner = bmm_prec(o.precision);
    if (inner == kem::kBmmFp32 && !A->A32) return fail(c, KEM_E_STATE, "mixed solve needs FP32 matrix storage");
    if (inner == kem::kBmmBf16 && !A->A16) return fail(c, KEM_E_STATE, "BF16 solve needs BF16 matrix storage");
    if (!A->A64) return fail(c, KEM_E_STATE, "mixed solve needs FP64 matrix storage for the residual");
    DVec R, D;
    DBuf ws;
    KEM_CK(c, dalloc(&R.p, n * nrhs));
    KEM_CK(c, dalloc(&D.p, n * nrhs));
    int status = KEM_OK, total = 0, refinements = 0;
    for (int ref = 0; ref < 50; ref++) {
        int rc = block_residual(A, nrhs, d_B, d_X, R.p, rel, &ws.p, &ws.bytes);
        if (rc != KEM_OK) return rc;
        double worst = 0.;
        std::vector<double> tol(nrhs);
        for (int r = 0; r < nrhs; r++) {
            worst = std::max(worst, rel[r]);
            // inner tolerance relative to the residual right-hand side; converged ones are skipped
            tol[r] = rel[r] <= o.tolerance ? 2.0
                                            : std::min(0.5, std::max(inner == kem::kBmmBf16 ? kBf16InnerTolerance
                                                                                              : o.inner_tolerance,
                                                                     0.5 * o.tolerance / rel[r]));
        }
        if (worst <= o.tolerance) break;
        if (total >= o.max_iterations || ctl.expired()) {
            status = KEM_W_NOT_CONVERGED;
            break;
        }
        KEM_CK(c, hipMemsetAsync(D.p, 0, sizeof(double) * n * nrhs, st));
        std::vector<double> inner_rel;
        rc = gmres_batch(A, inner, nrhs, R.p, D.p, tol, o.max_iterations - total, 0, ctl, &its, inner_rel, conv);
        if (rc < 0) return rc;
        total += its;
        refinements++;
        KEM_CK(c, kem::launch_baxpby(st, 1.0, d_X, 1.0, D.p, d_X, n * nrhs));
    }
    s->iterations = std::max(s->iterations, total);
    s->refinements = std::max(s->refinements, refinements);
    for (double v : rel)
        if (v > o.tolerance) status = KEM_W_NOT_CONVERGED;
    return status;
}

}  // namespace

int kem_matrix_multiply_batch(kem_matrix* A, int precision, int64_t nrhs, const double* X, double* Y)
{
    if (!A || nrhs < 0 || (nrhs > 0 && (!X || !Y))) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    if (precision != KEM_FP64 && precision != KEM_MIXED && precision != KEM_MIXED_BF16)
        return fail(c, KEM_E_ARG, "unknown precision mode");
    const int prec = bmm_prec(precision);
    if (prec == kem::kBmmFp32 && !A->A32) return fail(c, KEM_E_STATE, "FP32 product needs FP32 matrix storage");
    if (prec == kem::kBmmBf16 && !A->A16) return fail(c, KEM_E_STATE, "BF16 product needs BF16 matrix storage");
    if (prec == kem::kBmmFp64 && !A->A64) return fail(c, KEM_E_STATE, "FP64 product needs FP64 matrix storage");
    if (nrhs == 0) return KEM_OK;
    KEM_CK(c, enter(c));
    const int64_t n = A->n;
    const int g = (int)std::min<int64_t>(nrhs, 64);
    DVec dX, dY;
    DBuf ws;
    KEM_CK(c, dalloc(&dX.p, n * g));
    KEM_CK(c, dalloc(&dY.p, n * g));
    ws.bytes = kem::bmm_workspace_bytes(n, g, prec);
    KEM_CK(c, hipMalloc(&ws.p, ws.bytes));
    for (int64_t r0 = 0; r0 < nrhs; r0 += 64) {
        const int k = (int)std::min<int64_t>(64, nrhs - r0);
        KEM_CK(c, hipMemcpyAsync(dX.p, X + r0 * n, sizeof(double) * n * k, hipMemcpyHostToDevice, c->stream));
        hipEvent_t t0, t1;
        timing_pair(c, &t0, &t1);
        if (t0) KEM_CK(c, hipEventRecord(t0, c->stream));
        int rc = block_product(A, prec, k, dX.p, dY.p, &ws.p, &ws.bytes);
        if (rc != KEM_OK) return rc;
        if (t1) KEM_CK(c, hipEventRecord(t1, c->stream));
        KEM_CK(c, hipMemcpyAsync(Y + r0 * n, dY.p, sizeof(double) * n * k, hipMemcpyDeviceToHost, c->stream));
    }
    KEM_CK(c, hipStreamSynchronize(c->stream));
    return KEM_OK;
}

int kem_solve_batch(kem_matrix* A, const kem_solve_opts* opts, int64_t nrhs, const double* B, double* X,
                    double* residuals, kem_solve_stats* stats)
{
    if (!A || nrhs < 0 || (nrhs > 0 && (!B || !X))) return KEM_E_ARG;
    kem_ctx* c = A->ctx;
    kem_solve_opts o;
    kem_solve_opts_default(&o);
    if (opts) o = *opts;
    if (o.method != KEM_SOLVER_GMRES) return fail(c, KEM_E_ARG, "batched solve supports GMRES only");
    if (o.max_iterations <= 0) return fail(c, KEM_E_ARG, "max_iterations must be positive");
    if (!(o.tolerance > 0.)) return fail(c, KEM_E_ARG, "tolerance must be positive");
    if (o.precision != KEM_FP64 && o.precision != KEM_MIXED && o.precision != KEM_MIXED_BF16)
        return fail(c, KEM_E_ARG, "unknown precision mode");
    kem_solve_stats s{};
    if (nrhs == 0) {
        if (stats) *stats = s;
        return KEM_OK;
    }
    KEM_CK(c, enter(c));
    const int64_t n = A->n;
    SolveCtl ctl;
    ctl.t0 = now_s();
    ctl.limit = o.time_limit_s;
    const int g = (int)std::min<int64_t>(nrhs, 64);
    DVec dB, dX, dT;
    KEM_CK(c, dalloc(&dB.p, n * g));
    KEM_CK(c, dalloc(&dX.p, n * g));
    int status = KEM_OK;
    double worst = 0., worst_true = 0.;
    for (int64_t r0 = 0; r0 < nrhs; r0 += 64) {
        const int k = (int)std::min<int64_t>(64, nrhs - r0);
        KEM_CK(c, hipMemcpyAsync(dB.p, B + r0 * n, sizeof(double) * n * k, hipMemcpyHostToDevice, c->stream));
        KEM_CK(c, hipMemcpyAsync(dX.p, X + r0 * n, sizeof(double) * n * k, hipMemcpyHostToDevice, c->stream));
        std::vector<double> rel;
        int rc = solve_group(A, o, k, dB.p, dX.p, ctl, &s, rel);
        if (rc < 0) return rc;
        if (rc == KEM_W_NOT_CONVERGED) status = rc;
        KEM_CK(c, hipMemcpyAsync(X + r0 * n, dX.p, sizeof(double) * n * k, hipMemcpyDeviceToHost, c->stream));
        if (A->A64) {
            if (!dT.p) KEM_CK(c, dalloc(&dT.p, n * g));
            DBuf ws;
            std::vector<double> tr;
            rc = block_residual(A, k, dB.p, dX.p, dT.p, tr, &ws.p, &ws.bytes);
            if (rc != KEM_OK) return rc;
            for (int r = 0; r < k; r++) {
                worst_true = std::max(worst_true, tr[r]);
                if (residuals) residuals[r0 + r] = tr[r];
            }
        } else if (residuals) {
            for (int r = 0; r < k; r++) residuals[r0 + r] = rel[r];
        }
        for (double v : rel) worst = std::max(worst, v);
    }
    KEM_CK(c, hipStreamSynchronize(c->stream));
    s.residual = worst;
    s.true_residual = worst_true;
    s.seconds = now_s() - ctl.t0;
    if (stats) *stats = s;
    return status;
}

// ------------------------------------------------------------------ zonal harmonics
void kem_zh_params_default(kem_zh_params* p)
{
    if (!p) return;
    const kem::ZHParams d;
    p->n_bifurcations = d.nbif;
    p->convergence_ratio = d.conv_ratio;
    p->proximity_to_source_point = d.prox;
    p->convergence_parameter = d.conv_param;
    p->n_central_coefficients = d.ncen;
    p->central_fractional_spacing = d.frac_spacing ? 1 : 0;
    p->central_fractional_distance = d.frac_dist;
    p->central_delta_z = d.cen_dz;
    p->central_z1 = d.cen_z1;
    p->central_z2 = d.cen_z2;
    p->n_remote_coefficients = d.nrem;
    p->n_remote_source_points = d.nrem_sp;
    p->remote_z1 = d.rem_z1;
    p->remote_z2 = d.rem_z2;
}

void kem_zh_destroy(kem_zh* z)
{
    if (!z) return;
    if (z->ctx) {
        (void)hipSetDevice(z->ctx->device);
        (void)hipStreamSynchronize(z->ctx->stream);
    }
    dfree(z->node);
    dfree(z->child);
    dfree(z->elem);
    dfree(z->sp);
    dfree(z->coeff);
    dfree(z->leg);
    dfree(z->sigma);
    dfree(z->counts);
    delete z;
}

int kem_zh_create(kem_ctx* c, const kem_mesh* m, const double* sigma, const kem_zh_params* params, kem_zh** out)
{
    if (!c || !m || !sigma || !out) return KEM_E_ARG;
    *out = nullptr;
    if (m->ctx != c) return fail(c, KEM_E_STATE, "mesh belongs to another context");
    if (m->gstart) return fail(c, KEM_E_ARG, "zonal harmonics: meshes with symmetry groups are not supported");
    kem_zh_params dp;
    kem_zh_params_default(&dp);
    const kem_zh_params& q = params ? *params : dp;
    kem::ZHParams p;
    p.nbif = q.n_bifurcations;
    p.conv_ratio = q.convergence_ratio;
    p.prox = q.proximity_to_source_point;
    p.conv_param = q.convergence_parameter;
    p.ncen = q.n_central_coefficients;
    p.frac_spacing = q.central_fractional_spacing != 0;
    p.frac_dist = q.central_fractional_distance;
    p.cen_dz = q.central_delta_z;
    p.cen_z1 = q.central_z1;
    p.cen_z2 = q.central_z2;
    p.nrem = q.n_remote_coefficients;
    p.nrem_sp = q.n_remote_source_points;
    p.rem_z1 = q.remote_z1;
    p.rem_z2 = q.remote_z2;
    if (p.nbif < 0 || p.nbif > 20 || p.ncen < 3 || p.nrem < 3 || p.nrem_sp < 1 || !(p.conv_ratio > 0.) ||
        !(p.frac_dist > 0.) || !(p.cen_dz > 0.) || !(p.conv_param > 0.))
        return fail(c, KEM_E_ARG, "zonal harmonics: invalid parameters");
    KEM_CK(c, enter(c));
    // the elements as the generators see them (radius = x, axial position = z)
    const int64_t n = m->n;
    std::vector<double> g((size_t)n * KG_STRIDE);
    KEM_CK(c, hipMemcpy(g.data(), m->geom, sizeof(double) * g.size(), hipMemcpyDeviceToHost));
    std::vector<kem::ZHElement> el((size_t)n);
    for (int64_t i = 0; i < n; i++) {
        const double* d = g.data() + i * KG_STRIDE + KG_V;
        kem::ZHElement& e = el[(size_t)i];
        e.sigma = sigma[i];
        if (m->shape[(size_t)i] == KEM_SHAPE_RING) {
            e.type = kem::ZHElement::kRing;
            e.r0 = d[0];
            e.z0 = d[2];
        } else if (m->shape[(size_t)i] == KEM_SHAPE_CONICSECTION) {
            e.type = kem::ZHElement::kConic;
            e.r0 = d[0];
            e.z0 = d[2];
            e.r1 = d[3];
            e.z1 = d[5];
        }
    }
    const kem::ZHTables t = kem::zh_build(el, p);
    kem_zh* z = new kem_zh();
    z->ctx = c;
    z->mesh = m;
    std::vector<double> spf(4 * t.sp.size());
    for (size_t i = 0; i < t.sp.size(); i++) {
        spf[4 * i + 0] = t.sp[i].z0;
        spf[4 * i + 1] = t.sp[i].rho;
        spf[4 * i + 2] = (double)t.sp[i].offset;
        spf[4 * i + 3] = (double)t.sp[i].ncoef;
    }
    hipError_t e = hipSuccess;
    auto up = [&](auto** dst, const auto& v) {
        if (e != hipSuccess || v.empty()) return;
        e = dalloc(dst, (int64_t)v.size());
        if (e == hipSuccess) e = hipMemcpy(*dst, v.data(), sizeof(v[0]) * v.size(), hipMemcpyHostToDevice);
    };
    up(&z->node, t.node);
    up(&z->child, t.child);
    up(&z->elem, t.elem);
    up(&z->sp, spf);
    up(&z->coeff, t.coeff);
    up(&z->leg, t.leg);
    up(&z->sigma, std::vector<double>(sigma, sigma + n));
    if (e == hipSuccess) e = dalloc(&z->counts, 3);
    if (e != hipSuccess) {
        kem_zh_destroy(z);
        return fail(c, KEM_E_HIP, std::string("zonal harmonics: ") + hipGetErrorString(e));
    }
    z->dev.node = z->node;
    z->dev.child = z->child;
    z->dev.elem = z->elem;
    z->dev.sp = z->sp;
    z->dev.coeff = z->coeff;
    z->dev.leg = z->leg;
    z->dev.nleg = t.nleg;
    z->dev.depth = t.depth;
    z->dev.conv_ratio = p.conv_ratio;
    z->dev.prox = p.prox;
    z->dev.conv_param = p.conv_param;
    z->nnodes = (int64_t)t.node.size() / 8;
    for (size_t k = 0; k < t.node.size(); k += 8) {
        z->ncentral += t.node[k + 1];
        z->nremote += t.node[k + 3];
    }
    *out = z;
    return KEM_OK;
}

int kem_zh_info(const kem_zh* z, int64_t* out)
{
    if (!z || !out) return KEM_E_ARG;
    out[0] = z->nnodes;
    out[1] = z->ncentral;
    out[2] = z->nremote;
    out[3] = z->dev.depth;
    return KEM_OK;
}

int kem_zh_field_eval_device(kem_zh* z, int64_t npts, const double* x, const double* y, const double* zz,
                             double* phi, double* ex, double* ey, double* ez, int64_t* counts)
{
    if (!z || (npts > 0 && (!x || !y || !zz || !phi || !ex || !ey || !ez))) return KEM_E_ARG;
    kem_ctx* c = z->ctx;
    if (counts) counts[0] = counts[1] = counts[2] = 0;
    if (npts <= 0) return KEM_OK;
    KEM_CK(c, enter(c));
    if (counts) KEM_CK(c, hipMemsetAsync(z->counts, 0, 3 * sizeof(unsigned long long), c->stream));
    const kem_mesh* m = z->mesh;
    KEM_CK(c, kem::launch_zh_field(c->stream, z->dev, m->geom, m->nodes, m->recs, z->sigma, c->policy, x, y, zz,
                                   npts, phi, ex, ey, ez, counts ? z->counts : nullptr));
    if (counts) {
        unsigned long long h[3];
        KEM_CK(c, hipMemcpyAsync(h, z->counts, sizeof(h), hipMemcpyDeviceToHost, c->stream));
        KEM_CK(c, hipStreamSynchronize(c->stream));
        for (int k = 0; k < 3; k++) counts[k] = (int64_t)h[k];
    }
    return KEM_OK;
}

int kem_zh_field_eval(kem_zh* z, int64_t npts, const double* P, double* phi, double* E, int64_t* counts)
{
    if (!z || (npts > 0 && !P)) return KEM_E_ARG;
    kem_ctx* c = z->ctx;
    if (npts <= 0) {
        if (counts) counts[0] = counts[1] = counts[2] = 0;
        return KEM_OK;
    }
    KEM_CK(c, enter(c));
    std::vector<double> soa((size_t)npts * 3);
    for (int64_t i = 0; i < npts; i++)
        for (int k = 0; k < 3; k++) soa[(size_t)k * npts + i] = P[3 * i + k];
    double *dP = nullptr, *dO = nullptr;
    hipError_t e = dalloc(&dP, 3 * npts);
    if (e == hipSuccess) e = dalloc(&dO, 4 * npts);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dP, soa.data(), sizeof(double) * 3 * npts, hipMemcpyHostToDevice, c->stream);
    int rc = KEM_OK;
    if (e == hipSuccess) {
        rc = kem_zh_field_eval_device(z, npts, dP, dP + npts, dP + 2 * npts, dO, dO + npts, dO + 2 * npts,
                                      dO + 3 * npts, counts);
        if (rc == KEM_OK) {
            std::vector<double> o((size_t)npts * 4);
            e = hipMemcpyAsync(o.data(), dO, sizeof(double) * 4 * npts, hipMemcpyDeviceToHost, c->stream);
            if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
            if (e == hipSuccess)
                for (int64_t i = 0; i < npts; i++) {
                    if (phi) phi[i] = o[(size_t)i];
                    if (E)
                        for (int k = 0; k < 3; k++) E[3 * i + k] = o[(size_t)(k + 1) * npts + i];
                }
        }
    }
    dfree(dP);
    dfree(dO);
    if (rc != KEM_OK) return rc;
    if (e != hipSuccess) return fail(c, KEM_E_HIP, std::string("zonal harmonics: ") + hipGetErrorString(e));
    return KEM_OK;
}

}  // extern "C"
