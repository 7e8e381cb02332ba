// kem_kernels.hip -- HIP kernels of the electrostatic BEM path for gfx950 (MI355X).
//
// Field evaluation and matrix assembly map one field point / collocation target to each lane
// and stream the element records wave-uniformly (scalar loads), so a 64-lane wavefront
// evaluates one element against 64 points per step; lanes diverge only on the quadrature
// rule.  Dense linear algebra is HBM-bound streaming over column-major storage.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kem_device.hpp"
#include "kem_kernels.h"
#include "kem_policy.hpp"

namespace kem {

using dev::F4;

namespace {

constexpr int kBlock = 256;
constexpr int kDotParts = 256;  // block partials of a dot product

// ------------------------------------------------------------------ field evaluation
// Point block and element range of a far-pass block.  xcd_pb = 0: blockIdx.x and blockIdx.y.
// xcd_pb > 0 (the number of point blocks): a 1-D grid of 8 R element ranges per point block;
// block L takes point block (L / 8) % xcd_pb and element range L % 8 + 8 ((L / 8) / xcd_pb).
// The hardware hands consecutive blocks to the 8 XCDs in turn, so XCD x runs the ranges
// x, x + 8, x + 16, ... one after the other, each over all point blocks, and the range it
// works on stays in its L2.
struct FarTile {
    int64_t pb, split;
};
__device__ __forceinline__ FarTile far_tile(int64_t xcd_pb)
{
    if (xcd_pb) {
        const int64_t q = (int64_t)(blockIdx.x >> 3);
        return {q % xcd_pb, (int64_t)(blockIdx.x & 7) + 8 * (q / xcd_pb)};
    }
    return {(int64_t)blockIdx.x, (int64_t)blockIdx.y};
}

// One lane per field point.  Elements [j0, j1) are summed in container order with Kahan
// compensation per component (KElectrostaticIntegratingFieldSolver.hh:303-339).  The far pass
// evaluates the cubature / Gauss-Legendre pairs; the near pass then adds the closed-form pairs
// to the same sums.  Keeping the closed forms out of this kernel keeps its register use low
// enough for full occupancy.  With nsplit > 1 each blockIdx.y handles one contiguous element
// range and writes its partial sums; k_field_combine adds the partials in range order.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_field_far(const double* __restrict__ geom, const double* __restrict__ nodes,
                                                      const double* __restrict__ sigma,
                                                      const int32_t* __restrict__ pset, int64_t nsets, int64_t nelem,
                                                      int64_t per_split, const double* __restrict__ px,
                                                      const double* __restrict__ py, const double* __restrict__ pz,
                                                      int64_t npts, double* __restrict__ out0, double* __restrict__ out1,
                                                      double* __restrict__ out2, double* __restrict__ out3,
                                                      int64_t out_stride, int64_t xcd_pb)
{
    const FarTile t = far_tile(xcd_pb);
    int64_t i = t.pb * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double* __restrict__ sg = MULTI ? sigma + pset[i] : sigma;  // element-major sets
    const int64_t j0 = t.split * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > nelem) j1 = nelem;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    double c0 = 0., c1 = 0., c2 = 0., c3 = 0.;
    for (int64_t j = j0; j < j1; j++) {
        const double* g = geom + j * KG_STRIDE;
        const double r2 = dev::dist2(g, x, y, z);
        if (!dev::is_near2(g, r2, 0)) {
            const F4 v = dev::element_far(g, nodes + j * KEM_NODE_STRIDE, (int)g[KG_TYPE], r2, x, y, z);
            const double sj = sg[MULTI ? j * nsets : j];
            dev::kahan_add(v.x, sj, s0, c0);
            dev::kahan_add(v.y, sj, s1, c1);
            dev::kahan_add(v.z, sj, s2, c2);
            dev::kahan_add(v.p, sj, s3, c3);
        }
    }
    if (live) {
        const int64_t o = t.split * out_stride + i;
        out0[o] = s0;
        out1[o] = s1;
        out2[o] = s2;
        out3[o] = s3;
    }
}

// Mixed-precision far pass: the same pairs as k_field_far.  Pairs beyond the distance ratio
// fp64_ratio run with FP32 node sums (dev::element_far32, accuracy-targeted rule ladder); the
// band between the closed-form limit and fp64_ratio (empty by default) keeps the FP64 rules.
// All sums over elements are FP64.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_field_far32(const double* __restrict__ geom,
                                                        const double* __restrict__ nodes,
                                                        const float* __restrict__ nodes32, double fp64_ratio,
                                                        const double* __restrict__ sigma,
                                                        const int32_t* __restrict__ pset, int64_t nsets, int64_t nelem,
                                                        int64_t per_split, const double* __restrict__ px,
                                                        const double* __restrict__ py, const double* __restrict__ pz,
                                                        int64_t npts, double* __restrict__ out0,
                                                        double* __restrict__ out1, double* __restrict__ out2,
                                                        double* __restrict__ out3, int64_t out_stride, int64_t xcd_pb)
{
    const FarTile t = far_tile(xcd_pb);
    int64_t i = t.pb * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double* __restrict__ sg = MULTI ? sigma + pset[i] : sigma;  // element-major sets
    const int64_t j0 = t.split * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > nelem) j1 = nelem;
    const double band2 = fp64_ratio * fp64_ratio;
    double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
    for (int64_t j = j0; j < j1; j++) {
        const double* g = geom + j * KG_STRIDE;
        const double r2 = dev::dist2(g, x, y, z);
        if (dev::is_near2(g, r2, 0)) continue;
        const int type = (int)g[KG_TYPE];
        if (band2 == 0. || r2 * (g[KG_INVSIZE] * g[KG_INVSIZE]) > band2) {
            const float dx = (float)(x - g[KG_CX]), dy = (float)(y - g[KG_CY]), dz = (float)(z - g[KG_CZ]);
            float s0, s1, s2, s3;
            dev::element_far32(g, nodes32 + j * KEM_NODE32_STRIDE, type, r2, dx, dy, dz, s0, s1, s2, s3);
            const double wj = g[KG_PREF] * sg[MULTI ? j * nsets : j];
            a0 += (double)s0 * wj;
            a1 += (double)s1 * wj;
            a2 += (double)s2 * wj;
            a3 += (double)s3 * wj;
        } else {
            const F4 v = dev::element_far(g, nodes + j * KEM_NODE_STRIDE, type, r2, x, y, z);
            const double sj = sg[MULTI ? j * nsets : j];
            a0 += v.x * sj;
            a1 += v.y * sj;
            a2 += v.z * sj;
            a3 += v.p * sj;
        }
    }
    if (live) {
        const int64_t o = t.split * out_stride + i;
        out0[o] = a0;
        out1[o] = a1;
        out2[o] = a2;
        out3[o] = a3;
    }
}

// Near pass over all elements of a range: closed-form pairs added to the far pass's sums
// (numeric policy, accumulate = 1), or every pair from zero (rwg policy; PC = 1: the analytic
// and reference policies, from the raw shape records; PC = 2: meshes with conic sections or
// rings, any policy).
template <int PC, bool MULTI>
__global__ __launch_bounds__(kBlock) void k_field_near(const double* __restrict__ geom,
                                                       const double* __restrict__ nodes,
                                                       const double* __restrict__ recs,
                                                       const double* __restrict__ sigma,
                                                       const int32_t* __restrict__ pset, int64_t nsets, int64_t nelem,
                                                       int64_t per_split, const double* __restrict__ px,
                                                       const double* __restrict__ py, const double* __restrict__ pz,
                                                       int64_t npts, int policy, int accumulate,
                                                       double* __restrict__ out0, double* __restrict__ out1,
                                                       double* __restrict__ out2, double* __restrict__ out3,
                                                       int64_t out_stride)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double* __restrict__ sg = MULTI ? sigma + pset[i] : sigma;  // element-major sets
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > nelem) j1 = nelem;
    const int64_t o = (int64_t)blockIdx.y * out_stride + i;
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    if (accumulate) {
        s0 = out0[o];
        s1 = out1[o];
        s2 = out2[o];
        s3 = out3[o];
    }
    double c0 = 0., c1 = 0., c2 = 0., c3 = 0.;
    for (int64_t j = j0; j < j1; j++) {
        const double* g = geom + j * KG_STRIDE;
        if (PC == 1 || dev::is_near2(g, dev::dist2(g, x, y, z), policy)) {
            const F4 v = PC == 1   ? dev::element_policy(g, recs + j * KEM_REC, policy, x, y, z)
                         : PC == 2 ? dev::element_any<true>(g, nodes + j * KEM_NODE_STRIDE, recs + j * KEM_REC,
                                                            policy, x, y, z)
                                   : dev::element_near(g, (int)g[KG_TYPE], x, y, z);
            const double sj = sg[MULTI ? j * nsets : j];
            dev::kahan_add(v.x, sj, s0, c0);
            dev::kahan_add(v.y, sj, s1, c1);
            dev::kahan_add(v.z, sj, s2, c2);
            dev::kahan_add(v.p, sj, s3, c3);
        }
    }
    if (live) {
        out0[o] = s0;
        out1[o] = s1;
        out2[o] = s2;
        out3[o] = s3;
    }
}

__device__ __forceinline__ double wave_min(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

// Near pass with cluster culling (numeric policy).  Each wavefront bounds its 64 points by the
// sphere around their bounding box; a cluster of KEM_CLUSTER elements (Morton order) whose
// sphere of near distances cannot reach it is skipped as a whole.  The pairs evaluated are
// exactly the near set (the same dist2 test as the far pass); culling only removes clusters
// proven to contain none.  Writes its own sums (one slice of the partial-sum layout that
// k_field_combine adds), so it may run beside the far pass.  Blocks stride over the point
// blocks, so a small grid keeps the pass's footprint to a few wavefronts per SIMD beside the
// far kernel.  Capped at 168 VGPRs (3 wavefronts per SIMD, 16 bytes of spills): uncapped it
// takes 173 and runs 2 per SIMD (4 M points in uneven batches 253 -> 247 ms with the cap).
template <bool MULTI, bool AXI>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(3))) void k_field_near_culled(const double* __restrict__ geom,
                                                              const double* __restrict__ sigma,
                                                              const int32_t* __restrict__ pset, int64_t nsets,
                                                              const double* __restrict__ clusters, int64_t ncl,
                                                              const int32_t* __restrict__ order, int64_t nelem,
                                                              const double* __restrict__ px,
                                                              const double* __restrict__ py,
                                                              const double* __restrict__ pz, int64_t npts,
                                                              double* __restrict__ out0, double* __restrict__ out1,
                                                              double* __restrict__ out2, double* __restrict__ out3)
{
  for (int64_t pb = blockIdx.x; pb * kBlock < npts; pb += gridDim.x) {
    int64_t i = pb * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double* __restrict__ sg = MULTI ? sigma + pset[i] : sigma;  // element-major sets
    // bounding sphere of the wavefront's points
    const double lx = wave_min(x), ly = wave_min(y), lz = wave_min(z);
    const double hx = wave_max(x), hy = wave_max(y), hz = wave_max(z);
    const double wx = 0.5 * (lx + hx), wy = 0.5 * (ly + hy), wz = 0.5 * (lz + hz);
    const double wr = 0.5 * sqrt((hx - lx) * (hx - lx) + (hy - ly) * (hy - ly) + (hz - lz) * (hz - lz));
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    double c0 = 0., c1 = 0., c2 = 0., c3 = 0.;
    for (int64_t k = 0; k < ncl; k++) {
        const double* cl = clusters + 4 * k;
        const double dx = cl[0] - wx, dy = cl[1] - wy, dz = cl[2] - wz;
        const double reach = (cl[3] + wr) * (1. + 1e-9) + 1e-12;
        if (dx * dx + dy * dy + dz * dz > reach * reach) continue;
        int64_t e1 = (k + 1) * KEM_CLUSTER;
        if (e1 > nelem) e1 = nelem;
        for (int64_t e = k * KEM_CLUSTER; e < e1; e++) {
            const int64_t j = order[e];
            const double* g = geom + j * KG_STRIDE;
            if (dev::is_near2(g, dev::dist2(g, x, y, z), 0)) {
                const F4 v = AXI ? dev::element_near_any(g, (int)g[KG_TYPE], x, y, z)
                                 : dev::element_near(g, (int)g[KG_TYPE], x, y, z);
                const double sj = sg[MULTI ? j * nsets : j];
                dev::kahan_add(v.x, sj, s0, c0);
                dev::kahan_add(v.y, sj, s1, c1);
                dev::kahan_add(v.z, sj, s2, c2);
                dev::kahan_add(v.p, sj, s3, c3);
            }
        }
    }
    if (live) {
        out0[i] = s0;
        out1[i] = s1;
        out2[i] = s2;
        out3[i] = s3;
    }
  }
}

// A value every lane holds alike, moved to scalar registers so branches on it are uniform.
__device__ __forceinline__ double uniform(double v)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(v));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(v));
    return __hiloint2double(hi, lo);
}

typedef float v16f __attribute__((ext_vector_type(16)));

// the double whose low and high words are the bits of lo and hi
__device__ __forceinline__ double f2d(float lo, float hi)
{
    return __hiloint2double(__float_as_int(hi), __float_as_int(lo));
}

// weights of the degree-4 6-point rule's two orbits as the FP32 node table stores them
constexpr float kW6a = (float)0.223381589678011, kW6b = (float)0.109951743655322;
constexpr float kW3 = 1.f / 3.f;  // the degree-2 3-point rule (interior nodes)

// one FP32 quadrature node at offset (ex, ey, ez) from the point, weight w (dev::cubature32)
__device__ __forceinline__ void node32(float ex, float ey, float ez, float w, float& s0, float& s1, float& s2,
                                       float& s3)
{
    const float ir = __builtin_amdgcn_rsqf(ex * ex + ey * ey + ez * ez);
    const float wi = w * ir;
    const float f = wi * ir * ir;
    s0 += ex * f;
    s1 += ey * f;
    s2 += ez * f;
    s3 += wi;
}

// Per-call element weights of the far pass, sets element-major (entry j*nsets + s for
// element j, set s; sigma is set-major, sigma[s*nelem + j]):
//   wk[k*nsets + s] = prefactor * sigma of the k-th element in cluster order (order != null),
//   wk[j*nsets + s] = sigma (order == null, nsets > 1: the transposed sets).
__global__ __launch_bounds__(kBlock) void k_far_weights(const double* __restrict__ geom,
                                                        const int32_t* __restrict__ order,
                                                        const double* __restrict__ sigma, int64_t nelem,
                                                        int64_t nsets, double* __restrict__ wk)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nsets * nelem) return;
    const int64_t k = t / nsets, s = t - k * nsets;
    if (order) {
        const int64_t j = order[k];
        wk[t] = geom[j * KG_STRIDE + KG_PREF] * sigma[s * nelem + j];
    } else {
        wk[t] = sigma[s * nelem + k];
    }
}

// Mixed-precision far pass over element clusters.  Each wavefront bounds its 64 points by a
// sphere; an element whose 6-point reach (KEM_FAR6_REACH) lies beyond that sphere is a 6-point
// pair for every lane -- the per-pair ladder of k_field_far32 would pick the same rule -- so
// the wavefront evaluates it from the element's far record without per-lane regime tests; a
// cluster beyond its own reach skips even the per-element test, and a cluster beyond its 3-point
// reach (KEM_FAR3_RATIO sizes of each element) takes the 3-point records.  Every other element skips the
// near lanes and evaluates the highest rule that k_field_far32's per-pair ladder picks for any
// far lane, once for all of them (dev::element_far32_wave).  A 6-point pair evaluates P - node as (P - cluster centre)
// rounded to FP32 once per cluster minus the node's FP32 offset from the cluster centre (the
// per-pair path rounds P - element centre per pair); the FP64 sums run over elements in
// cluster order.  Elements
// [c0*32, c1*32) of the cluster order per blockIdx.y; wk from k_far_weights.
template <bool MULTI>
__global__ __launch_bounds__(kBlock) void k_field_far32c(const double* __restrict__ geom,
                                                         const float* __restrict__ nodes32,
                                                         const float* __restrict__ far6,
                                                         const float* __restrict__ far3,
                                                         const double* __restrict__ far6_cl,
                                                         const int32_t* __restrict__ order,
                                                         const double* __restrict__ wk,
                                                         const int32_t* __restrict__ pset, int64_t nsets, int64_t nelem,
                                                         int64_t ncl, int64_t cl_per_split,
                                                         const double* __restrict__ px,
                                                         const double* __restrict__ py,
                                                         const double* __restrict__ pz, int64_t npts,
                                                         double* __restrict__ out0, double* __restrict__ out1,
                                                         double* __restrict__ out2, double* __restrict__ out3,
                                                         int64_t out_stride, int64_t xcd_pb)
{
    const FarTile t = far_tile(xcd_pb);
    int64_t i = t.pb * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double* __restrict__ wl = MULTI ? wk + pset[i] : wk;  // element-major sets
    // bounding sphere of the wavefront's points, radius padded against rounding
    const double lx = wave_min(x), ly = wave_min(y), lz = wave_min(z);
    const double hx = wave_max(x), hy = wave_max(y), hz = wave_max(z);
    const double wx = uniform(0.5 * (lx + hx)), wy = uniform(0.5 * (ly + hy)), wz = uniform(0.5 * (lz + hz));
    const double wr =
        uniform(0.5 * sqrt((hx - lx) * (hx - lx) + (hy - ly) * (hy - ly) + (hz - lz) * (hz - lz)) * (1. + 1e-9) +
                1e-12);
    const int64_t c0 = t.split * cl_per_split;
    int64_t c1 = c0 + cl_per_split;
    if (c1 > ncl) c1 = ncl;
    double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
    for (int64_t c = c0; c < c1; c++) {
        const double* cl = far6_cl + KEM_FAR_CL_STRIDE * c;
        const double qx = wx - cl[0], qy = wy - cl[1], qz = wz - cl[2];
        const double q2 = qx * qx + qy * qy + qz * qz;
        const double tc = cl[3] + wr, t3 = cl[4] + wr;
        const bool all6 = q2 > tc * tc;
        // P - cluster centre, rounded to FP32 once per cluster; the far records hold the nodes
        // relative to the cluster centre
        const float dx = (float)(x - cl[0]), dy = (float)(y - cl[1]), dz = (float)(z - cl[2]);
        int64_t k1 = (c + 1) * KEM_CLUSTER;
        if (k1 > nelem) k1 = nelem;
        if (!MULTI && q2 > t3 * t3) {
            // every element of the cluster lies beyond KEM_FAR3_RATIO sizes of every lane: the
            // degree-2 3-point rule (single-set passes: in the several-set kernel the branch
            // costs 3 % on the 4 M-point batches, which have no such clusters)
            for (int64_t k = c * KEM_CLUSTER; k < k1; k++) {
                const v16f R = *reinterpret_cast<const v16f*>(far3 + k * KEM_FAR3_STRIDE);
                const double wj = wl[MULTI ? k * nsets : k];
                float s0 = 0.f, s1 = 0.f, s2 = 0.f, s3 = 0.f;
                node32(dx - R[0], dy - R[1], dz - R[2], kW3, s0, s1, s2, s3);
                node32(dx - R[4], dy - R[5], dz - R[6], kW3, s0, s1, s2, s3);
                node32(dx - R[8], dy - R[9], dz - R[10], kW3, s0, s1, s2, s3);
                a0 += (double)s0 * wj;
                a1 += (double)s1 * wj;
                a2 += (double)s2 * wj;
                a3 += (double)s3 * wj;
            }
            continue;
        }
        for (int64_t k = c * KEM_CLUSTER; k < k1; k++) {
            // the whole far record in two 16-dword loads: nodes 0-3, then nodes 4-5, centre, reach
            const v16f* rv = reinterpret_cast<const v16f*>(far6 + k * KEM_FAR6_STRIDE);
            const v16f A = rv[0], B = rv[1];
            const double wj = wl[MULTI ? k * nsets : k];
            const double cx = f2d(B[8], B[9]), cy = f2d(B[10], B[11]), cz = f2d(B[12], B[13]);
            bool six = all6;
            if (!six) {
                const double ux = wx - cx, uy = wy - cy, uz = wz - cz;
                const double te = f2d(B[14], B[15]) + wr;
                six = ux * ux + uy * uy + uz * uz > te * te;
            }
            float s0, s1, s2, s3;
            if (six) {
                s0 = s1 = s2 = s3 = 0.f;
                node32(dx - A[0], dy - A[1], dz - A[2], kW6a, s0, s1, s2, s3);
                node32(dx - A[4], dy - A[5], dz - A[6], kW6a, s0, s1, s2, s3);
                node32(dx - A[8], dy - A[9], dz - A[10], kW6a, s0, s1, s2, s3);
                node32(dx - A[12], dy - A[13], dz - A[14], kW6b, s0, s1, s2, s3);
                node32(dx - B[0], dy - B[1], dz - B[2], kW6b, s0, s1, s2, s3);
                node32(dx - B[4], dy - B[5], dz - B[6], kW6b, s0, s1, s2, s3);
            } else {
                const int64_t j = order[k];
                const double* g = geom + j * KG_STRIDE;
                const double r2 = dev::dist2(g, x, y, z);
                if (dev::is_near2(g, r2, 0)) continue;
                const float ex = (float)(x - g[KG_CX]), ey = (float)(y - g[KG_CY]), ez = (float)(z - g[KG_CZ]);
                dev::element_far32_wave(g, nodes32 + j * KEM_NODE32_STRIDE, (int)g[KG_TYPE], r2, ex, ey, ez, s0,
                                        s1, s2, s3);
            }
            a0 += (double)s0 * wj;
            a1 += (double)s1 * wj;
            a2 += (double)s2 * wj;
            a3 += (double)s3 * wj;
        }
    }
    if (live) {
        const int64_t o = t.split * out_stride + i;
        out0[o] = a0;
        out1[o] = a1;
        out2[o] = a2;
        out3[o] = a3;
    }
}

// Path census of k_field_far32c over a single charge-density set: the kernel's control flow
// (wavefront bounding sphere, cluster and element reach tests, near skip, wave rule) for points
// in the order the pass visits them, counting per live lane instead of evaluating.
// counts[0] 3-point cluster pairs, [1] 6-point record pairs, [2..5] pairs of the per-pair path
// by the rule the lane's own ladder picks (6, 7, 12, 33 points), [6, 7] segment pairs with 4 /
// 16 nodes, [8] near pairs the per-pair path skips, [9] node evaluations the per-pair path
// runs (the wave rule's nodes for each far lane).  Element ranges are whole clusters, so one
// range over all clusters takes the same decisions as any split of them.
__global__ __launch_bounds__(kBlock) void k_census_far32c(const double* __restrict__ geom,
                                                          const float* __restrict__ far6,
                                                          const double* __restrict__ far6_cl,
                                                          const int32_t* __restrict__ order, int64_t nelem,
                                                          int64_t ncl, const double* __restrict__ px,
                                                          const double* __restrict__ py,
                                                          const double* __restrict__ pz, int64_t npts,
                                                          unsigned long long* __restrict__ counts)
{
    __shared__ unsigned long long sc[10];
    if (threadIdx.x < 10) sc[threadIdx.x] = 0;
    __syncthreads();
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < npts;
    if (!live) i = npts - 1;
    const double x = px[i], y = py[i], z = pz[i];
    const double lx = wave_min(x), ly = wave_min(y), lz = wave_min(z);
    const double hx = wave_max(x), hy = wave_max(y), hz = wave_max(z);
    const double wx = uniform(0.5 * (lx + hx)), wy = uniform(0.5 * (ly + hy)), wz = uniform(0.5 * (lz + hz));
    const double wr =
        uniform(0.5 * sqrt((hx - lx) * (hx - lx) + (hy - ly) * (hy - ly) + (hz - lz) * (hz - lz)) * (1. + 1e-9) +
                1e-12);
    unsigned long long c[10] = {0};
    for (int64_t cc = 0; cc < ncl; cc++) {
        const double* cl = far6_cl + KEM_FAR_CL_STRIDE * cc;
        const double qx = wx - cl[0], qy = wy - cl[1], qz = wz - cl[2];
        const double q2 = qx * qx + qy * qy + qz * qz;
        const double tc = cl[3] + wr, t3 = cl[4] + wr;
        const bool all6 = q2 > tc * tc;
        int64_t k1 = (cc + 1) * KEM_CLUSTER;
        if (k1 > nelem) k1 = nelem;
        const int64_t cnt = k1 - cc * KEM_CLUSTER;
        if (q2 > t3 * t3) {
            c[0] += cnt;
            continue;
        }
        for (int64_t k = cc * KEM_CLUSTER; k < k1; k++) {
            const v16f B = reinterpret_cast<const v16f*>(far6 + k * KEM_FAR6_STRIDE)[1];
            bool six = all6;
            if (!six) {
                const double cx = f2d(B[8], B[9]), cy = f2d(B[10], B[11]), cz = f2d(B[12], B[13]);
                const double ux = wx - cx, uy = wy - cy, uz = wz - cz;
                const double te = f2d(B[14], B[15]) + wr;
                six = ux * ux + uy * uy + uz * uz > te * te;
            }
            if (six) {
                c[1]++;
                continue;
            }
            const int64_t j = order[k];
            const double* g = geom + j * KG_STRIDE;
            const double r2 = dev::dist2(g, x, y, z);
            const bool near = dev::is_near2(g, r2, 0);
            const int type = (int)g[KG_TYPE];
            // the wave rule (dev::element_far32_wave) over the far lanes; as in the pass, lanes
            // past the last point stand in for it
            const bool fl = !near;
            int wn;
            if (type == KEM_LINESEGMENT) {
                wn = __builtin_amdgcn_ballot_w64(fl && !(r2 > g[KG_MIX7_2])) == 0 ? 4 : 16;
            } else {
                wn = __builtin_amdgcn_ballot_w64(fl && !(r2 > g[KG_MIX12_2])) != 0   ? 33
                     : __builtin_amdgcn_ballot_w64(fl && !(r2 > g[KG_MIX7_2])) != 0 ? 12
                     : __builtin_amdgcn_ballot_w64(fl && !(r2 > g[KG_MIX6_2])) != 0 ? 7
                                                                                      : 6;
            }
            if (near) {
                c[8]++;
                continue;
            }
            c[9] += wn;
            if (type == KEM_LINESEGMENT)
                c[r2 > g[KG_MIX7_2] ? 6 : 7]++;
            else
                c[r2 > g[KG_MIX6_2] ? 2 : r2 > g[KG_MIX7_2] ? 3 : r2 > g[KG_MIX12_2] ? 4 : 5]++;
        }
    }
    for (int k = 0; k < 10; k++)
        if (live && c[k]) atomicAdd(&sc[k], c[k]);
    __syncthreads();
    if (threadIdx.x < 10 && sc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], sc[threadIdx.x]);
}

// Kahan-sum the per-range partials (ws[c][s][i], c = Ex, Ey, Ez, phi) in range order.
__global__ __launch_bounds__(kBlock) void k_field_combine(const double* __restrict__ ws, int nsplit, int64_t npts,
                                                          double* __restrict__ phi, double* __restrict__ ex,
                                                          double* __restrict__ ey, double* __restrict__ ez)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npts) return;
    double* outs[4] = {ex, ey, ez, phi};
    for (int c = 0; c < 4; c++) {
        if (!outs[c]) continue;
        const double* p = ws + (int64_t)c * nsplit * npts;
        double s = 0., comp = 0.;
        for (int k = 0; k < nsplit; k++) {
            const double yk = p[(int64_t)k * npts + i] - comp;
            const double t = s + yk;
            comp = (t - s) - yk;
            s = t;
        }
        outs[c][i] = s;
    }
}

// ------------------------------------------------------------------ element pairs
template <int PC>
__global__ __launch_bounds__(kBlock) void k_pairs(const double* __restrict__ geom, const double* __restrict__ nodes,
                                                  const double* __restrict__ recs, int policy, int64_t npairs,
                                                  const int64_t* __restrict__ elem, const double* __restrict__ P,
                                                  double* __restrict__ out, const int64_t* __restrict__ gstart)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= npairs) return;
    // a symmetry group sums its members in order (KElectrostaticElementIntegrator.hh:52-79)
    const int64_t m0 = gstart ? gstart[elem[k]] : elem[k], m1 = gstart ? gstart[elem[k] + 1] : elem[k] + 1;
    F4 s{0., 0., 0., 0.};
    for (int64_t j = m0; j < m1; j++) {
        const F4 v = PC == 1   ? dev::element_policy(geom + j * KG_STRIDE, recs + j * KEM_REC, policy, P[3 * k],
                                                     P[3 * k + 1], P[3 * k + 2])
                     : PC == 2 ? dev::element_any<true>(geom + j * KG_STRIDE, nodes + j * KEM_NODE_STRIDE,
                                                        recs + j * KEM_REC, policy, P[3 * k], P[3 * k + 1],
                                                        P[3 * k + 2])
                               : dev::element_eval<true>(geom + j * KG_STRIDE, nodes + j * KEM_NODE_STRIDE, policy,
                                                         P[3 * k], P[3 * k + 1], P[3 * k + 2]);
        s.x += v.x;
        s.y += v.y;
        s.z += v.z;
        s.p += v.p;
    }
    out[4 * k + 0] = s.x;
    out[4 * k + 1] = s.y;
    out[4 * k + 2] = s.z;
    out[4 * k + 3] = s.p;
}

// charge densities per member of symmetry groups: out[s*nm + i] = sigma[s*n + gof[i]]
__global__ __launch_bounds__(kBlock) void k_expand_groups(const int32_t* __restrict__ gof, int64_t nm, int64_t n,
                                                          int64_t nsets, const double* __restrict__ sigma,
                                                          double* __restrict__ out)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= nm * nsets) return;
    const int64_t s = t / nm, i = t - s * nm;
    out[t] = sigma[s * n + gof[i]];
}

// Step-level sum over fields: out_c[i] += t[c*npts + i] (c = phi, Ex, Ey, Ez), one field at a time
// in the order the fields were added (KSRootElectricField.cxx:104-121 adds per sample point)
__global__ __launch_bounds__(kBlock) void k_accumulate4(const double* __restrict__ t, int64_t npts,
                                                        double* __restrict__ phi, double* __restrict__ ex,
                                                        double* __restrict__ ey, double* __restrict__ ez)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= npts) return;
    phi[i] += t[i];
    ex[i] += t[npts + i];
    ey[i] += t[2 * npts + i];
    ez[i] += t[3 * npts + i];
}

// ------------------------------------------------------------------ assembly
// Lane = collocation target i (rows), wave-uniform source j (columns).  Column-major storage:
// A(i, j) at A[j*n + i], so the 64 lanes of a wavefront write 64 consecutive entries.
// KElectrostaticBoundaryIntegrator::ComputeBoundaryIntegral (Core/include/...hh:260-280).
template <bool FIELD, int PC, bool GROUPED>
__global__ __launch_bounds__(kBlock) void k_assemble(const double* __restrict__ geom,
                                                     const double* __restrict__ nodes,
                                                     const double* __restrict__ recs,
                                                     const double* __restrict__ targets, int policy, int64_t n,
                                                     int64_t per_split, double* __restrict__ A64,
                                                     float* __restrict__ A32, const int64_t* __restrict__ gstart)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < n;
    if (!live) i = n - 1;
    const double* t = targets + i * KT_STRIDE;
    const double cx = t[KT_CX], cy = t[KT_CY], cz = t[KT_CZ];
    const double nx = t[KT_NX], ny = t[KT_NY], nz = t[KT_NZ];
    const bool dirichlet = t[KT_BC] == 0.;
    const double f = t[KT_FLUX];
    const double diag = ((1. + f) / (1. - f)) / (2. * dev::kEps0);
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > n) j1 = n;
    for (int64_t j = j0; j < j1; j++) {
        // column j: element j, or the members of symmetry group j summed in order
        // (KElectrostaticElementIntegrator.hh:52-79)
        const int64_t m0 = GROUPED ? gstart[j] : j, m1 = GROUPED ? gstart[j + 1] : j + 1;
        F4 v{0., 0., 0., 0.};
        for (int64_t q = m0; q < m1; q++) {
            const F4 u = PC == 1   ? dev::element_policy(geom + q * KG_STRIDE, recs + q * KEM_REC, policy, cx, cy, cz)
                         : PC == 2 ? dev::element_any<FIELD>(geom + q * KG_STRIDE, nodes + q * KEM_NODE_STRIDE,
                                                             recs + q * KEM_REC, policy, cx, cy, cz)
                                   : dev::element_eval<FIELD>(geom + q * KG_STRIDE, nodes + q * KEM_NODE_STRIDE,
                                                              policy, cx, cy, cz);
            if (!GROUPED) {
                v = u;
                break;
            }
            v.x += u.x;
            v.y += u.y;
            v.z += u.z;
            v.p += u.p;
        }
        double a;
        if (!FIELD || dirichlet) {
            a = v.p;
        } else {
            const double* tj = targets + j * KT_STRIDE;
            const double dx = tj[KT_CX] - cx, dy = tj[KT_CY] - cy, dz = tj[KT_CZ] - cz;
            a = (sqrt(dx * dx + dy * dy + dz * dz) >= 1.e-12) ? (v.x * nx + v.y * ny + v.z * nz) : diag;
        }
        if (live) {
            if (A64) A64[j * n + i] = a;
            if (A32) A32[j * n + i] = (float)a;
        }
    }
}

// ------------------------------------------------------------------ assembly, split by regime
// The numeric-policy assembly of a mesh without symmetry groups or axisymmetric shapes runs in
// two kernels, as the field pass does: k_assemble_far evaluates the cubature / Gauss-Legendre
// entries and skips the closed-form ones, k_assemble_near evaluates only the closed-form
// entries for the targets whose wavefront can reach an element cluster.  Both classify a pair
// by the field passes' squared-distance tests (dist2 against the element's (dr_cut size)^2,
// no square root per pair), so every entry is written by exactly one of them; the rule a pair
// gets differs from the reference's dr > cut only within an ulp of a cut-off, where both rules
// meet the reference's accuracy.

// Far entries: lane = collocation target i, wave-uniform source column j, as k_assemble, but
// the closed-form entries are left to k_assemble_near (no closed form inlined here: the kernel
// keeps the far field kernels' register budget), and the far lanes of a wavefront share one
// cubature rule per column (dev::element_far_wave: 112.9 -> 108 ms on the 100,000-triangle
// assembly).  Dirichlet rows take k_assemble_far_pot instead.  KElectrostaticBoundaryIntegrator::
// ComputeBoundaryIntegral (Core/include/KElectrostaticBoundaryIntegrator.hh:260-280).
template <bool FIELD>
__global__ __launch_bounds__(kBlock) void k_assemble_far(const double* __restrict__ geom,
                                                         const double* __restrict__ nodes,
                                                         const double* __restrict__ targets, int64_t n,
                                                         int64_t per_split, double* __restrict__ A64,
                                                         float* __restrict__ A32)
{
    int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = i < n;
    if (!live) i = n - 1;
    const double* t = targets + i * KT_STRIDE;
    const double cx = t[KT_CX], cy = t[KT_CY], cz = t[KT_CZ];
    const double nx = t[KT_NX], ny = t[KT_NY], nz = t[KT_NZ];
    const bool dirichlet = t[KT_BC] == 0.;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > n) j1 = n;
    for (int64_t j = j0; j < j1; j++) {
        const double* g = geom + j * KG_STRIDE;
        const double r2 = dev::dist2(g, cx, cy, cz);
        if (dev::is_near2(g, r2, 0)) continue;
        const F4 v = dev::element_far_wave<FIELD>(g, nodes + j * KEM_NODE_STRIDE, (int)g[KG_TYPE], r2, cx, cy, cz);
        // far entries lie beyond 2 element sizes: never the Neumann self term
        const double a = (!FIELD || dirichlet) ? v.p : (v.x * nx + v.y * ny + v.z * nz);
        if (live) {
            if (A64) A64[j * n + i] = a;
            if (A32) A32[j * n + i] = (float)a;
        }
    }
}

// Far entries of Dirichlet rows with T targets per lane: wavefront w of a block takes rows
// 64 T w .. 64 T (w + 1) - 1 of the block's 256 T, lane l rows l + 64 t, so each wave-uniform
// node load of a column serves T evaluations; the nodes come from the centre-relative table an
// (k_asm_nodes: r^2 in 4 operations per node instead of 6, for 5 scalar values per node instead
// of 4).  100,000-triangle assembly, both kernels: 108 ms at one target per lane, 100-102 ms at
// 2 with the plain nodes, 96-98 ms with the table; the table at one target per lane was slower
// (120 ms at an unroll of 4, 149 ms at 16, SGPR spills fully unrolled).
template <int T>
__global__ __launch_bounds__(kBlock) void k_assemble_far_pot(const double* __restrict__ geom,
                                                             const double* __restrict__ nodes,
                                                             const double* __restrict__ an,
                                                             const double* __restrict__ targets, int64_t n,
                                                             int64_t per_split, double* __restrict__ A64,
                                                             float* __restrict__ A32)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = (int64_t)blockIdx.x * (T * kBlock) + wave * (64 * T) + lane;
    int64_t row[T];
    bool live[T];
    double px[T], py[T], pz[T];
#pragma unroll
    for (int t = 0; t < T; t++) {
        row[t] = i0 + 64 * t;
        live[t] = row[t] < n;
        if (!live[t]) row[t] = n - 1;
        px[t] = targets[row[t] * KT_STRIDE + KT_CX];
        py[t] = targets[row[t] * KT_STRIDE + KT_CY];
        pz[t] = targets[row[t] * KT_STRIDE + KT_CZ];
    }
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > n) j1 = n;
    for (int64_t j = j0; j < j1; j++) {
        const double* g = geom + j * KG_STRIDE;
        double r2[T], v[T], dx[T], dy[T], dz[T];
        bool far[T], any = false;
#pragma unroll
        for (int t = 0; t < T; t++) {
            dx[t] = px[t] - g[KG_CX];
            dy[t] = py[t] - g[KG_CY];
            dz[t] = pz[t] - g[KG_CZ];
            r2[t] = __builtin_fma(dz[t], dz[t], __builtin_fma(dy[t], dy[t], dx[t] * dx[t]));  // dev::dist2
            far[t] = !dev::is_near2(g, r2[t], 0);
            any = any || far[t];
        }
        if (__builtin_amdgcn_ballot_w64(any) == 0) continue;
        dev::element_far_wave_pot_n<T>(g, nodes + j * KEM_NODE_STRIDE, an + j * KEM_ANODE_STRIDE, (int)g[KG_TYPE], r2,
                                       far, dx, dy, dz, px, py, pz, v);
#pragma unroll
        for (int t = 0; t < T; t++)
            if (live[t] && far[t]) {
                if (A64) A64[j * n + row[t]] = v[t];
                if (A32) A32[j * n + row[t]] = (float)v[t];
            }
    }
}

// The assembly's potential node table (KEM_ANODE_STRIDE layout) from the cubature nodes: one
// thread per (element, node slot); unused slots hold values never read.
__global__ __launch_bounds__(kBlock) void k_asm_nodes(const double* __restrict__ geom,
                                                      const double* __restrict__ nodes, int64_t n,
                                                      double* __restrict__ an)
{
    const int64_t t = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (t >= n * KEM_NODE_SLOTS) return;
    const int64_t j = t / KEM_NODE_SLOTS;
    const int slot = (int)(t - j * KEM_NODE_SLOTS);
    const double* g = geom + j * KG_STRIDE;
    const double* q = nodes + j * KEM_NODE_STRIDE + 4 * slot;
    const double x = q[0] - g[KG_CX], y = q[1] - g[KG_CY], z = q[2] - g[KG_CZ];
    double* o = an + j * KEM_ANODE_STRIDE + 5 * slot;
    o[0] = -2. * x;
    o[1] = -2. * y;
    o[2] = -2. * z;
    o[3] = __builtin_fma(z, z, __builtin_fma(y, y, x * x));
    o[4] = q[3];
}

// Closed-form entries: targets in the Morton order of the element centres (order[k]), so a
// wavefront's targets are compact; each wavefront bounds its targets by a sphere and skips the
// element clusters that cannot hold a near pair of any of them (the culling of
// k_field_near_culled); the entries it finds are written in place (a few per target).
template <bool FIELD>
__global__ __launch_bounds__(kBlock) void k_assemble_near(const double* __restrict__ geom,
                                                          const double* __restrict__ targets,
                                                          const double* __restrict__ clusters, int64_t ncl,
                                                          const int32_t* __restrict__ order, int64_t n,
                                                          double* __restrict__ A64, float* __restrict__ A32)
{
    int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = k < n;
    if (!live) k = n - 1;
    const int64_t i = order[k];
    const double* t = targets + i * KT_STRIDE;
    const double cx = t[KT_CX], cy = t[KT_CY], cz = t[KT_CZ];
    const double nx = t[KT_NX], ny = t[KT_NY], nz = t[KT_NZ];
    const bool dirichlet = t[KT_BC] == 0.;
    const double f = t[KT_FLUX];
    const double diag = ((1. + f) / (1. - f)) / (2. * dev::kEps0);
    const double lx = wave_min(cx), ly = wave_min(cy), lz = wave_min(cz);
    const double hx = wave_max(cx), hy = wave_max(cy), hz = wave_max(cz);
    const double wx = 0.5 * (lx + hx), wy = 0.5 * (ly + hy), wz = 0.5 * (lz + hz);
    const double wr = 0.5 * sqrt((hx - lx) * (hx - lx) + (hy - ly) * (hy - ly) + (hz - lz) * (hz - lz));
    for (int64_t c = 0; c < ncl; c++) {
        const double* cl = clusters + 4 * c;
        const double dx = cl[0] - wx, dy = cl[1] - wy, dz = cl[2] - wz;
        const double reach = (cl[3] + wr) * (1. + 1e-9) + 1e-12;
        if (dx * dx + dy * dy + dz * dz > reach * reach) continue;
        int64_t e1 = (c + 1) * KEM_CLUSTER;
        if (e1 > n) e1 = n;
        for (int64_t e = c * KEM_CLUSTER; e < e1; e++) {
            const int64_t j = order[e];
            const double* g = geom + j * KG_STRIDE;
            if (!dev::is_near2(g, dev::dist2(g, cx, cy, cz), 0)) continue;
            const int type = (int)g[KG_TYPE];
            const F4 v = type == KEM_TRIANGLE    ? dev::closed_form<3, FIELD>(g, cx, cy, cz)
                         : type == KEM_RECTANGLE ? dev::closed_form<4, FIELD>(g, cx, cy, cz)
                                                 : dev::line_analytic(g, cx, cy, cz);
            double a;
            if (!FIELD || dirichlet) {
                a = v.p;
            } else {
                const double* tj = targets + j * KT_STRIDE;
                const double ex = tj[KT_CX] - cx, ey = tj[KT_CY] - cy, ez = tj[KT_CZ] - cz;
                a = (sqrt(ex * ex + ey * ey + ez * ez) >= 1.e-12) ? (v.x * nx + v.y * ny + v.z * nz) : diag;
            }
            if (live) {
                if (A64) A64[j * n + i] = a;
                if (A32) A32[j * n + i] = (float)a;
            }
        }
    }
}

// ------------------------------------------------------------------ dense linear algebra
// y_part[s][i] = sum_{j in split s} A[j*lda + i] * x[j]   (column-major, rows i)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_part(const T* __restrict__ A, int64_t lda, int64_t nrows,
                                                      int64_t ncols, int64_t per_split, const double* __restrict__ x,
                                                      double* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > ncols) j1 = ncols;
    if (i >= nrows) return;
    double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
    int64_t j = j0;
    for (; j + 4 <= j1; j += 4) {
        a0 += (double)A[(j + 0) * lda + i] * x[j + 0];
        a1 += (double)A[(j + 1) * lda + i] * x[j + 1];
        a2 += (double)A[(j + 2) * lda + i] * x[j + 2];
        a3 += (double)A[(j + 3) * lda + i] * x[j + 3];
    }
    for (; j < j1; j++) a0 += (double)A[j * lda + i] * x[j];
    part[(int64_t)blockIdx.y * nrows + i] = (a0 + a1) + (a2 + a3);
}

// y[i] = alpha * sum_s part[s][i] + beta * y[i]
// Same as k_gemv_part with R consecutive rows per thread read as one vector per column (used for
// FP32 matrices, R = 2: one 8-byte vector; needs lda and A aligned in R-row units and
// nrows % R == 0).  Measured on the 20,000-unknown FP32 matrix: 5.4 TB/s with R = 2 against
// 4.9-5.1 TB/s with R = 4 (16-byte vectors); FP64 matrices stay with k_gemv_part (5.9 TB/s).
template <typename T, int R>
__global__ __launch_bounds__(kBlock) void k_gemv_part_vec(const T* __restrict__ A, int64_t lda, int64_t nrows,
                                                          int64_t ncols, int64_t per_split,
                                                          const double* __restrict__ x, double* __restrict__ part,
                                                          const int* __restrict__ skip)
{
    typedef T vec __attribute__((ext_vector_type(R)));
    if (skip && *skip) return;
    const int64_t i0 = ((int64_t)blockIdx.x * kBlock + threadIdx.x) * R;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > ncols) j1 = ncols;
    if (i0 >= nrows) return;
    double a[R], b[R];
    for (int r = 0; r < R; r++) a[r] = b[r] = 0.;
    double c[R], d[R];
    for (int r = 0; r < R; r++) c[r] = d[r] = 0.;
    int64_t j = j0;
    for (; j + 4 <= j1; j += 4) {
        const vec u = *reinterpret_cast<const vec*>(A + j * lda + i0);
        const vec v = *reinterpret_cast<const vec*>(A + (j + 1) * lda + i0);
        const vec w = *reinterpret_cast<const vec*>(A + (j + 2) * lda + i0);
        const vec t = *reinterpret_cast<const vec*>(A + (j + 3) * lda + i0);
        const double x0 = x[j], x1 = x[j + 1], x2 = x[j + 2], x3 = x[j + 3];
        for (int r = 0; r < R; r++) {
            a[r] += (double)u[r] * x0;
            b[r] += (double)v[r] * x1;
            c[r] += (double)w[r] * x2;
            d[r] += (double)t[r] * x3;
        }
    }
    for (; j < j1; j++) {
        const vec u = *reinterpret_cast<const vec*>(A + j * lda + i0);
        const double xj = x[j];
        for (int r = 0; r < R; r++) a[r] += (double)u[r] * xj;
    }
    for (int r = 0; r < R; r++) part[(int64_t)blockIdx.y * nrows + i0 + r] = (a[r] + b[r]) + (c[r] + d[r]);
}

__global__ __launch_bounds__(kBlock) void k_gemv_reduce(const double* __restrict__ part, int nsplit, int64_t nrows,
                                                        double alpha, double beta, double* __restrict__ y,
                                                        const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nrows) return;
    double s = 0.;
    for (int k = 0; k < nsplit; k++) s += part[(int64_t)k * nrows + i];
    y[i] = (beta == 0. ? 0. : beta * y[i]) + alpha * s;
}

__device__ __forceinline__ double block_sum(double v, double* sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double r = 0.;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) r += sh[k];
    __syncthreads();
    return r;
}

// y[j] = sum_i A[j*lda + i] * x[i]  (one block per column; y = A^T x)
template <typename T>
__global__ __launch_bounds__(kBlock) void k_gemv_t(const T* __restrict__ A, int64_t lda, int64_t nrows,
                                                   const double* __restrict__ x, double* __restrict__ y,
                                                   const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sh[kBlock / 64];
    const int64_t j = blockIdx.x;
    const T* col = A + j * lda;
    double s = 0.;
    for (int64_t i = threadIdx.x; i < nrows; i += kBlock) s += (double)col[i] * x[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) y[j] = s;
}

// partial dot products: part[b] = sum over the block's grid-stride slice
__global__ __launch_bounds__(kBlock) void k_dot_part(const double* __restrict__ a, const double* __restrict__ b,
                                                     int64_t n, double* __restrict__ part)
{
    __shared__ double sh[kBlock / 64];
    double s = 0.;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        s += a[i] * b[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// out = sum part (sqrt if want_sqrt), single block
__global__ __launch_bounds__(kBlock) void k_dot_final(const double* __restrict__ part, int nparts, int want_sqrt,
                                                      double* __restrict__ out)
{
    __shared__ double sh[kBlock / 64];
    double s = 0.;
    for (int k = threadIdx.x; k < nparts; k += kBlock) s += part[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = want_sqrt ? sqrt(s) : s;
}

// y += sign * (*alpha) * x
__global__ __launch_bounds__(kBlock) void k_axpy_dev(const double* __restrict__ alpha, double sign,
                                                     const double* __restrict__ x, double* __restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] -= (-sign * (*alpha)) * x[i];
}

// y = x / (*d)
__global__ __launch_bounds__(kBlock) void k_dot_part_skip(const double* __restrict__ a, const double* __restrict__ b,
                                                          int64_t n, double* __restrict__ part,
                                                          const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ double sh[kBlock / 64];
    double s = 0.;
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock)
        s += a[i] * b[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ __launch_bounds__(kBlock) void k_dot_final_skip(const double* __restrict__ part, int nparts,
                                                           double* __restrict__ out, const int* __restrict__ skip)
{
    if (*skip) return;
    __shared__ double sh[kBlock / 64];
    double s = 0.;
    for (int k = threadIdx.x; k < nparts; k += kBlock) s += part[k];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) *out = sqrt(s);
}
__global__ __launch_bounds__(kBlock) void k_div_dev(const double* __restrict__ d, const double* __restrict__ x,
                                                    double* __restrict__ y, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) y[i] = x[i] * (1.0 / (*d));
}

// z = a*x + b*y
// ---- GMRES Arnoldi step on device state (ints: [0] j, [1] done, [2] jlimit, [3] converged)
// h[q] = V[:, q] . w for q <= j (one block per column of the basis's capacity)
// hp[q*S + s] = <V q, w> over rows [s per, (s+1) per) for q <= j (S = gridDim.y row ranges:
// a block per (q, range), so the latency chain of a column's dot product is 1/S as long; one
// block per column took 27 us per pass at 20,000 rows and 102 us at 100,000)
__global__ __launch_bounds__(kBlock) void k_arn_proj(const double* __restrict__ V, int64_t n,
                                                     const double* __restrict__ w, double* __restrict__ hp,
                                                     int64_t per, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int64_t q = blockIdx.x;
    if (q > ints[0]) return;
    const int S = gridDim.y, r = blockIdx.y;
    __shared__ double sh[kBlock / 64];
    const double* col = V + q * n;
    const int64_t i1 = std::min<int64_t>(n, (r + 1) * per);
    double s = 0.;
    for (int64_t i = r * per + threadIdx.x; i < i1; i += kBlock) s += col[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) hp[q * S + r] = s;
}

// h[q] = sum_s hp[q*S + s] in order, q <= j
__global__ __launch_bounds__(kBlock) void k_arn_hsum(const double* __restrict__ hp, int S, double* __restrict__ h,
                                                     const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int j = ints[0];
    for (int q = threadIdx.x; q <= j; q += kBlock) {
        double s = 0.;
        for (int r = 0; r < S; r++) s += hp[(int64_t)q * S + r];
        h[q] = s;
    }
}
// part[k][i] = sum over the k-th column range (of columns 0..j) of V[i, q] h[q]
__global__ __launch_bounds__(kBlock) void k_arn_update_part(const double* __restrict__ V, int64_t n,
                                                            int64_t per_split, const double* __restrict__ h,
                                                            double* __restrict__ part, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t ncols = (int64_t)ints[0] + 1;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    int64_t j1 = j0 + per_split;
    if (j1 > ncols) j1 = ncols;
    if (i >= n) return;
    double a0 = 0., a1 = 0., a2 = 0., a3 = 0.;
    int64_t q = j0;
    for (; q + 4 <= j1; q += 4) {
        a0 += V[(q + 0) * n + i] * h[q + 0];
        a1 += V[(q + 1) * n + i] * h[q + 1];
        a2 += V[(q + 2) * n + i] * h[q + 2];
        a3 += V[(q + 3) * n + i] * h[q + 3];
    }
    for (; q < j1; q++) a0 += V[q * n + i] * h[q];
    part[(int64_t)blockIdx.y * n + i] = (a0 + a1) + (a2 + a3);
}
// V[:, j+1] = vcur = w / nrm
__global__ __launch_bounds__(kBlock) void k_arn_normalize(const double* __restrict__ w, const double* __restrict__ nrm,
                                                          double* __restrict__ V, double* __restrict__ vcur,
                                                          int64_t n, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const double v = w[i] * (1.0 / (*nrm));
    V[((int64_t)ints[0] + 1) * n + i] = v;
    vcur[i] = v;
}
// Column j of the Hessenberg matrix from the two projection passes and the norm, rotated by
// the previous Givens rotations; the new rotation zeroes H[j+1, j] and rotates p
// (KGeneralizedMinimalResidual.hh:178-271).  Rounded as the host sequence it replaces: no
// contraction.
__global__ void k_arn_givens(const double* __restrict__ h1, const double* __restrict__ h2,
                             const double* __restrict__ nrm, double* __restrict__ H, int ld, double* __restrict__ cs,
                             double* __restrict__ sn, double* __restrict__ p, int* __restrict__ ints,
                             double* __restrict__ scal)
{
    _Pragma("clang fp contract(off)");
    if (threadIdx.x != 0 || ints[1]) return;
    const int j = ints[0];
    double* h = H + (int64_t)j * ld;
    for (int r = 0; r <= j; r++) h[r] = h1[r] + h2[r];
    const double beta = *nrm;
    h[j + 1] = beta;
    for (int r = 0; r < j; r++) {
        const double a = h[r], bb = h[r + 1];
        h[r] = cs[r] * a + sn[r] * bb;
        h[r + 1] = -1.0 * sn[r] * a + cs[r] * bb;
    }
    const double gam = sqrt(h[j] * h[j] + beta * beta);
    cs[j] = h[j] / gam;
    sn[j] = beta / gam;
    h[j] = gam;
    p[j + 1] = -1.0 * sn[j] * p[j];
    p[j] = cs[j] * p[j];
    const double rn = fabs(p[j + 1]);
    scal[1] = rn;
    ints[0] = j + 1;
    const bool conv = rn <= scal[0];
    ints[3] = conv ? 1 : 0;
    if (conv || j + 1 >= ints[2] || !(beta > 0.)) ints[1] = 1;
}

__global__ __launch_bounds__(kBlock) void k_axpby(double a, const double* __restrict__ x, double b,
                                                  const double* __restrict__ y, double* __restrict__ z, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) z[i] = a * x[i] + b * y[i];
}

// ------------------------------------------------------------------ Robin Hood
// `steps` Southwell sub-iterations of KRobinHood (KRobinHood.hh:76-97 with
// KRobinHood_SingleThread.hh:78-108) in one workgroup: r = b - b_iter, i = first argmax |r_i|,
// x_i += r_i / A_ii, b_iter += A(:, i) r_i / A_ii.  At the last step ||r||_inf (before its
// update, as the reference checks it) is written to *rnorm.
constexpr int kRHBlock = 1024;
__global__ __launch_bounds__(kRHBlock) void k_robin_hood(const double* __restrict__ A, const double* __restrict__ b,
                                                         double* __restrict__ x, double* __restrict__ bi, int64_t n,
                                                         int64_t steps, double* __restrict__ rnorm)
{
    __shared__ double sv[kRHBlock / 64];
    __shared__ int64_t si[kRHBlock / 64];
    __shared__ double s_corr;
    __shared__ int64_t s_idx;
    for (int64_t s = 0; s < steps; s++) {
        double best = -1.;
        int64_t bidx = n;
        for (int64_t i = threadIdx.x; i < n; i += kRHBlock) {
            const double r = fabs(b[i] - bi[i]);
            if (r > best) {
                best = r;
                bidx = i;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_down(best, o, 64);
            const int64_t oi = __shfl_down(bidx, o, 64);
            if (ov > best || (ov == best && oi < bidx)) {
                best = ov;
                bidx = oi;
            }
        }
        const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
        if (lane == 0) {
            sv[w] = best;
            si[w] = bidx;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double bv = sv[0];
            int64_t bx = si[0];
            for (int k = 1; k < kRHBlock / 64; k++)
                if (sv[k] > bv || (sv[k] == bv && si[k] < bx)) {
                    bv = sv[k];
                    bx = si[k];
                }
            // no lane found a residual > -1 (NaN residuals): the reference starts from index 0
            if (bx >= n) bx = 0;
            if (s == steps - 1) *rnorm = bv;
            const double corr = (b[bx] - bi[bx]) / A[bx * n + bx];
            x[bx] += corr;
            s_corr = corr;
            s_idx = bx;
        }
        __syncthreads();
        const double corr = s_corr;
        const double* col = A + s_idx * n;
        for (int64_t i = threadIdx.x; i < n; i += kRHBlock) bi[i] += col[i] * corr;
        __syncthreads();
    }
}

// The same sub-iterations over G co-resident workgroups (cooperative launch), for matrices too
// large for one workgroup to scan each step.  Workgroup g owns rows [g*per, g*per + per) and
// keeps their b and b_iter in LDS; per step it publishes its first argmax of |b_i - b_iter_i|
// (value, index, signed residual) in a slot, all workgroups meet at a grid barrier, each
// reduces the G slots to the same global first argmax (ties to the smaller index, as the scan
// of k_robin_hood), divides the owner's residual by A_kk, workgroup 0 adds it to x_k, and
// every workgroup adds the column's share to its rows.  The arithmetic of every value is
// k_robin_hood's, so both take the same corrections.  Slots alternate between two buffers by
// step parity; the barrier is a monotonic arrival counter (zero on entry) read with
// device-scope acquire loads, its wait bounded (a workgroup that waits past the bound sets
// *fault and the grid stops).
constexpr int kRHGBlock = 256;
constexpr int kRHRowsMax = 4096;  // rows per workgroup: b and b_iter in 64 KB of LDS
// from this many unknowns on, the sub-iterations run over several workgroups
constexpr int64_t kRobinHoodGridMin = 8192;

__device__ __forceinline__ void rh_store(double* p, double v)
{
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double rh_load(const double* p)
{
    return __longlong_as_double((long long)__hip_atomic_load(reinterpret_cast<const unsigned long long*>(p),
                                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
}

__global__ __launch_bounds__(kRHGBlock) void k_robin_hood_grid(const double* __restrict__ A,
                                                               const double* __restrict__ b, double* __restrict__ x,
                                                               double* __restrict__ bi, int64_t n, int64_t per,
                                                               int64_t steps, double* __restrict__ rnorm,
                                                               double* slots, unsigned int* bar,
                                                               unsigned int* fault)
{
    __shared__ double lb[kRHRowsMax], lbi[kRHRowsMax];
    __shared__ double sv[kRHGBlock / 64], sr[kRHGBlock / 64];
    __shared__ int64_t si[kRHGBlock / 64];
    __shared__ double s_corr;
    __shared__ int64_t s_idx;
    __shared__ int s_stop;
    const int G = gridDim.x;
    const int64_t r0 = (int64_t)blockIdx.x * per;
    const int64_t r1 = r0 + per < n ? r0 + per : n;
    const int64_t nr = r1 > r0 ? r1 - r0 : 0;
    for (int64_t i = threadIdx.x; i < nr; i += kRHGBlock) {
        lb[i] = b[r0 + i];
        lbi[i] = bi[r0 + i];
    }
    if (threadIdx.x == 0) s_stop = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int64_t s = 0; s < steps; s++) {
        // this workgroup's first argmax, with the signed residual there
        double best = -1., bres = 0.;
        int64_t bidx = n;
        for (int64_t i = threadIdx.x; i < nr; i += kRHGBlock) {
            const double d = lb[i] - lbi[i];
            const double r = fabs(d);
            if (r > best) {
                best = r;
                bidx = r0 + i;
                bres = d;
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_down(best, o, 64);
            const int64_t oi = __shfl_down(bidx, o, 64);
            const double orr = __shfl_down(bres, o, 64);
            if (ov > best || (ov == best && oi < bidx)) {
                best = ov;
                bidx = oi;
                bres = orr;
            }
        }
        if (lane == 0) {
            sv[w] = best;
            si[w] = bidx;
            sr[w] = bres;
        }
        __syncthreads();
        double* slot = slots + ((s & 1) * (int64_t)G + blockIdx.x) * 3;
        if (threadIdx.x == 0) {
            double bv = sv[0], br = sr[0];
            int64_t bx = si[0];
            for (int k = 1; k < kRHGBlock / 64; k++)
                if (sv[k] > bv || (sv[k] == bv && si[k] < bx)) {
                    bv = sv[k];
                    bx = si[k];
                    br = sr[k];
                }
            rh_store(slot, bv);
            rh_store(slot + 1, (double)bx);
            rh_store(slot + 2, br);
            // arrive, then wait for every workgroup's arrival at this step
            __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            const unsigned int want = (unsigned int)(G * (s + 1));
            unsigned int spins = 0;
            while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < want) {
                if (++spins > (1u << 26) || __hip_atomic_load(fault, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                    __hip_atomic_store(fault, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    s_stop = 1;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        __syncthreads();
        if (s_stop) break;
        if (w == 0) {
            // the global first argmax over the slots of this step: lane k reads slots k, k + 64,
            // ... (loads in flight together), then the wavefront reduces with first-index ties
            const double* sl = slots + (s & 1) * (int64_t)G * 3;
            double gv = -1., gr = 0.;
            int64_t gx = n;
            for (int k = lane; k < G; k += 64) {
                const double v = rh_load(sl + 3 * k);
                const int64_t ix = (int64_t)rh_load(sl + 3 * k + 1);
                const double rr = rh_load(sl + 3 * k + 2);
                if (v > gv || (v == gv && ix < gx)) {
                    gv = v;
                    gx = ix;
                    gr = rr;
                }
            }
            for (int o = 32; o > 0; o >>= 1) {
                const double ov = __shfl_down(gv, o, 64);
                const int64_t oi = __shfl_down(gx, o, 64);
                const double orr = __shfl_down(gr, o, 64);
                if (ov > gv || (ov == gv && oi < gx)) {
                    gv = ov;
                    gx = oi;
                    gr = orr;
                }
            }
            if (lane == 0) {
                // no workgroup found a residual > -1: every residual is NaN, the reference
                // corrects index 0 by its (NaN) residual
                if (gx >= n) {
                    gx = 0;
                    gr = __builtin_nan("");
                }
                if (s == steps - 1 && blockIdx.x == 0) *rnorm = gv;
                const double corr = gr / A[gx * n + gx];
                if (blockIdx.x == 0) x[gx] += corr;
                s_corr = corr;
                s_idx = gx;
            }
        }
        __syncthreads();
        const double corr = s_corr;
        const double* col = A + s_idx * n + r0;
        for (int64_t i = threadIdx.x; i < nr; i += kRHGBlock) lbi[i] += col[i] * corr;
        __syncthreads();
    }
    for (int64_t i = threadIdx.x; i < nr; i += kRHGBlock) bi[r0 + i] = lbi[i];
}

// ------------------------------------------------------------------ Gauss-Jordan
// KGaussianElimination::Solve (LinearAlgebra/Solvers/include/KGaussianElimination.hh:24-58):
// no pivoting; for each column col every row != col of the augmented matrix [A | b] becomes
// a[row][:] - (a[row][col] / a[col][col]) * a[col][:] over all n+1 columns; then
// x[row] = a[row][n] / a[row][row].  Within one column step the rows are independent (the
// pivot row is not modified), so the step is one factor kernel and one update kernel over the
// whole matrix.  Column-major M (entry (row, j) at j*n + row), lanes along rows.  Built
// without FMA contraction so each entry sees the reference's roundings.
__global__ __launch_bounds__(kBlock) void k_gj_factors(const double* __restrict__ M, int64_t n, int64_t col,
                                                       double* __restrict__ f)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    f[i] = (i == col) ? 0. : M[col * n + i] / M[col * n + col];
}

__global__ __launch_bounds__(kBlock) void k_gj_update(double* __restrict__ M, int64_t n, int64_t col,
                                                      const double* __restrict__ f)
{
    _Pragma("clang fp contract(off)");
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n || i == col) return;
    const double fi = f[i];
    for (int64_t j = blockIdx.y; j <= n; j += gridDim.y) M[j * n + i] = M[j * n + i] - fi * M[j * n + col];
}

__global__ __launch_bounds__(kBlock) void k_gj_solution(const double* __restrict__ M, int64_t n,
                                                        double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n) x[i] = M[n * n + i] / M[i * n + i];
}

// ------------------------------------------------------------------ set-index validation
// *bad = 1 if any pset[i] lies outside [0, nsets) (benign race: every writer stores 1)
__global__ __launch_bounds__(kBlock) void k_check_sets(const int32_t* __restrict__ pset, int64_t n, int64_t nsets,
                                                       int* __restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < n && (pset[i] < 0 || pset[i] >= nsets)) *bad = 1;
}

// ------------------------------------------------------------------ regime census
// Counts the (point, element) pairs of a field evaluation by quadrature regime with the field
// kernels' own classification (numeric policy; mixed = 1: the FP32 ladder beyond distance
// ratio fp64_ratio): counts[0..3] triangle 7/12/33-point/closed form, [4..7] rectangle,
// [8..10] segment 4-node/16-node/analytic, [11] triangle 6-point (FP32 ladder only).  Used to
// price a launch in FLOPs.
__global__ __launch_bounds__(kBlock) void k_census(const double* __restrict__ geom, int64_t nelem, int64_t per_split,
                                                   const double* __restrict__ px, const double* __restrict__ py,
                                                   const double* __restrict__ pz, int64_t npts, int mixed,
                                                   double fp64_ratio, unsigned long long* __restrict__ counts)
{
    __shared__ unsigned long long sc[12];
    if (threadIdx.x < 12) sc[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    unsigned long long c[12] = {0};
    if (i < npts) {
        const double x = px[i], y = py[i], z = pz[i];
        const int64_t j0 = (int64_t)blockIdx.y * per_split;
        int64_t j1 = j0 + per_split;
        if (j1 > nelem) j1 = nelem;
        for (int64_t j = j0; j < j1; j++) {
            const double* g = geom + j * KG_STRIDE;
            const int type = (int)g[KG_TYPE];
            const double r2 = dev::dist2(g, x, y, z);
            const int base = type == KEM_TRIANGLE ? 0 : type == KEM_RECTANGLE ? 4 : 8;
            int k;
            if (dev::is_near2(g, r2, 0)) {
                k = type == KEM_LINESEGMENT ? 2 : 3;
            } else {
                const bool f32 = mixed && r2 * (g[KG_INVSIZE] * g[KG_INVSIZE]) > fp64_ratio * fp64_ratio;
                const double t7 = f32 ? g[KG_MIX7_2] : g[KG_REF7_2];
                const double t12 = f32 ? g[KG_MIX12_2] : g[KG_REF12_2];
                if (type == KEM_LINESEGMENT)
                    k = r2 > t7 ? 0 : 1;
                else if (f32 && r2 > g[KG_MIX6_2])
                    k = 11 - base;
                else
                    k = r2 > t7 ? 0 : r2 > t12 ? 1 : 2;
            }
            c[base + k]++;
        }
    }
    for (int k = 0; k < 12; k++)
        if (c[k]) atomicAdd(&sc[k], c[k]);
    __syncthreads();
    if (threadIdx.x < 12 && sc[threadIdx.x]) atomicAdd(&counts[threadIdx.x], sc[threadIdx.x]);
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

}  // namespace

// ====================================================================== launchers
namespace {
thread_local LaunchTuning g_tuning;
constexpr int64_t kPartialDoubles = 4 * (int64_t)(2048 + 256) * 256;
// From this many point blocks on, the numeric-policy far pass of one charge-density set splits
// the elements into ranges of at most range_elems elements, a multiple of 8 of them (far_tile),
// over chunks of points whose 8 R partial sums per point fit kXcdPartialDoubles.  (Measured on
// 1 M points x 20,000 triangles, 8 ranges: FP64 399 -> 361 ms, mixed 68 -> 60 ms.  Passes over
// several sets keep one element range per block: with the per-lane set weights 8 ranges
// measured 23 % slower.)
constexpr int64_t kXcdMinBlocks = 1024;
constexpr int64_t kXcdPartialDoubles = (int64_t)128 << 20;
constexpr int kXcds = 8;

bool xcd_split(int policy, const FieldMesh& m, int64_t npts, int64_t range_elems)
{
    return policy == 0 && m.clusters && m.order && range_elems > 0 && (int64_t)nblocks(npts) >= kXcdMinBlocks;
}

// the culled near pass writes one more slice of partial sums beside the far pass's
bool culled_near(int policy, const FieldMesh& m) { return policy == 0 && m.clusters && m.order; }

// Culled near pass over points [p0, p0 + np) into slice `slice` of the partial layout q (ns
// slices of np doubles per component).  With a side stream it runs there, forked from st
// after the work queued so far; join_side() makes st wait for it.
template <bool MULTI>
void near_slice(hipStream_t st, const SideStream* side, const FieldMesh& m, const double* sg, const int32_t* pset,
                int64_t nsets, const double* px, const double* py, const double* pz, int64_t np, double* q,
                int64_t ns, int64_t slice, bool mixed)
{
    hipStream_t s = st;
    if (side && side->st) {
        (void)hipEventRecord(side->fork, st);
        (void)hipStreamWaitEvent(side->st, side->fork, 0);
        s = side->st;
    }
    const int64_t ncl = (m.n + KEM_CLUSTER - 1) / KEM_CLUSTER;
    double* o = q + slice * np;
    // beside the mixed far pass a grid of g_tuning.near_blocks blocks (0: one per point block):
    // a few wavefronts per SIMD, which the far kernel's blocks do not crowd out, so the pass ends
    // with the far kernel instead of after it (measured on 1 M points x 20,000 triangles: 67.5 ms
    // with one block per point block, 64.9 ms with 64 to 512 blocks, 69.3 ms with 768).  The FP64
    // far pass runs long enough that a small grid only takes its occupancy (362 -> 367..402 ms),
    // and so does the near pass over several sets (4 M points in uneven batches: 253 -> 261 ms).
    unsigned gn = nblocks(np);
    if (s != st && mixed && !MULTI && g_tuning.near_blocks > 0 && (int64_t)gn > g_tuning.near_blocks)
        gn = (unsigned)g_tuning.near_blocks;
    if (m.axisym)
        k_field_near_culled<MULTI, true><<<gn, kBlock, 0, s>>>(m.geom, sg, pset, nsets, m.clusters, ncl,
                                                                        m.order, m.n, px, py, pz, np, o,
                                                                        o + ns * np, o + 2 * ns * np, o + 3 * ns * np);
    else
        k_field_near_culled<MULTI, false><<<gn, kBlock, 0, s>>>(m.geom, sg, pset, nsets, m.clusters, ncl,
                                                                         m.order, m.n, px, py, pz, np, o,
                                                                         o + ns * np, o + 2 * ns * np, o + 3 * ns * np);
    if (s != st) (void)hipEventRecord(side->join, s);
}

void join_side(hipStream_t st, const SideStream* side)
{
    if (side && side->st) (void)hipStreamWaitEvent(st, side->join, 0);
}

template <bool MULTI>
void far_kernel(hipStream_t st, dim3 grid, int kind, const FieldMesh& m, double fp64_ratio, const double* wk,
                const double* sg, const int32_t* pset, int64_t nsets, int64_t per, const double* px,
                const double* py, const double* pz, int64_t npts, double* o0, double* o1, double* o2, double* o3,
                int64_t stride, int64_t xcd_pb)
{
    const int64_t nelem = m.n;
    if (kind == 2) {
        const int64_t ncl = (nelem + KEM_CLUSTER - 1) / KEM_CLUSTER;
        k_field_far32c<MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes32, m.far6, m.far3, m.far6_cl, m.order, wk, pset, nsets,
                                                       nelem, ncl, per / KEM_CLUSTER, px, py, pz, npts, o0, o1, o2,
                                                       o3, stride, xcd_pb);
    } else if (kind == 1) {
        k_field_far32<MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes, m.nodes32, fp64_ratio, sg, pset, nsets, nelem,
                                                      per, px, py, pz, npts, o0, o1, o2, o3, stride, xcd_pb);
    } else {
        k_field_far<MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes, sg, pset, nsets, nelem, per, px, py, pz, npts, o0,
                                                    o1, o2, o3, stride, xcd_pb);
    }
}

template <bool MULTI>
hipError_t field_passes(hipStream_t st, int policy, const FieldMesh& m, bool mixed, double fp64_ratio,
                        const double* sigma, const int32_t* pset, int64_t nsets, const double* px, const double* py,
                        const double* pz, int64_t npts, double* phi, double* ex, double* ey, double* ez, double* ws,
                        int64_t ws_doubles, int64_t range_elems, hipEvent_t far_begin, hipEvent_t far_end,
                        const SideStream* side)
{
    const int64_t nelem = m.n;
    const FieldPlan p = plan_field(policy, m, mixed, fp64_ratio, MULTI, nsets, npts, range_elems, ws_doubles);
    // every partial sum the passes write lies inside ws
    if (p.weights + p.partials > ws_doubles) return hipErrorInvalidValue;
    const bool clustered = p.kind == 2;
    // ws: the cluster pass's element weights, then (several sets) the sets transposed to
    // element-major order, then the partial sums of split element ranges
    const int64_t nsig = (nsets * nelem + 31) / 32 * 32;
    double* wk = ws;
    if (clustered) ws += nsig;
    const double* sg = sigma;
    if (MULTI) {
        k_far_weights<<<nblocks(nsets * nelem), kBlock, 0, st>>>(m.geom, nullptr, sigma, nelem, nsets, ws);
        sg = ws;
        ws += nsig;
    }
    const bool culled = culled_near(policy, m);
    const int64_t nr = p.ranges, ns = p.slices;
    if (p.xcd) {
        // 8 R element ranges of at most range_elems elements, R per XCD; partial sums per
        // range (and of the near pass) over chunks of points
        if (far_begin) (void)hipEventRecord(far_begin, st);
        if (clustered)
            k_far_weights<<<nblocks(nsets * nelem), kBlock, 0, st>>>(m.geom, m.order, sigma, nelem, nsets, wk);
        for (int64_t p0 = 0; p0 < npts; p0 += p.chunk) {
            const int64_t np = std::min(p.chunk, npts - p0);
            const unsigned gc = nblocks(np);
            double* q = ws;
            if (culled)
                near_slice<MULTI>(st, side, m, sg, MULTI ? pset + p0 : nullptr, nsets, px + p0, py + p0, pz + p0, np,
                                  q, ns, nr, p.kind != 0);
            far_kernel<MULTI>(st, dim3((unsigned)(gc * nr)), p.kind, m, fp64_ratio, wk, sg,
                              MULTI ? pset + p0 : nullptr, nsets, p.per, px + p0, py + p0, pz + p0, np, q,
                              q + ns * np, q + 2 * ns * np, q + 3 * ns * np, np, (int64_t)gc);
            // far_end right after the last far kernel: the events time the far pass (with the
            // joins and combines of earlier chunks), not the wait for the near pass
            if (far_end && p0 + np >= npts) (void)hipEventRecord(far_end, st);
            if (culled) join_side(st, side);
            k_field_combine<<<gc, kBlock, 0, st>>>(ws, (int)ns, np, phi + p0, ex + p0, ey + p0, ez + p0);
        }
    } else {
        const unsigned gx = nblocks(npts);
        dim3 grid(gx, (unsigned)nr);
        double *o0 = ex, *o1 = ey, *o2 = ez, *o3 = phi;
        int64_t stride = 0;
        if (p.partials > 0) {
            o0 = ws;
            o1 = ws + ns * npts;
            o2 = ws + 2 * ns * npts;
            o3 = ws + 3 * ns * npts;
            stride = npts;
        }
        if (culled) near_slice<MULTI>(st, side, m, sg, pset, nsets, px, py, pz, npts, ws, ns, nr, p.kind != 0);
        if (far_begin) (void)hipEventRecord(far_begin, st);
        if (clustered)
            k_far_weights<<<nblocks(nsets * nelem), kBlock, 0, st>>>(m.geom, m.order, sigma, nelem, nsets, wk);
        if (policy == 0) far_kernel<MULTI>(st, grid, p.kind, m, fp64_ratio, wk, sg, pset, nsets, p.per, px, py, pz,
                                           npts, o0, o1, o2, o3, stride, 0);
        if (far_end) (void)hipEventRecord(far_end, st);
        if (!culled) {
            if (m.axisym)
                k_field_near<2, MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes, m.recs, sg, pset, nsets, nelem, p.per,
                                                                px, py, pz, npts, policy, policy == 0 ? 1 : 0, o0, o1,
                                                                o2, o3, stride);
            else if (policy >= 2)
                k_field_near<1, MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes, m.recs, sg, pset, nsets, nelem, p.per,
                                                                px, py, pz, npts, policy, 0, o0, o1, o2, o3, stride);
            else
                k_field_near<0, MULTI><<<grid, kBlock, 0, st>>>(m.geom, m.nodes, m.recs, sg, pset, nsets, nelem, p.per,
                                                                px, py, pz, npts, policy, policy == 0 ? 1 : 0, o0, o1,
                                                                o2, o3, stride);
        }
        if (culled) join_side(st, side);
        if (p.partials > 0) k_field_combine<<<gx, kBlock, 0, st>>>(ws, (int)ns, npts, phi, ex, ey, ez);
    }
    return hipSuccess;
}
}  // namespace

FieldPlan plan_field(int policy, const FieldMesh& m, bool mixed, double fp64_ratio, bool multi, int64_t nsets,
                     int64_t npts, int64_t range_elems, int64_t ws_doubles)
{
    FieldPlan p;
    const int64_t nelem = m.n;
    if (npts <= 0 || nelem <= 0 || nsets < 1) return p;
    const bool clustered = policy == 0 && mixed && m.nodes32 && fp64_ratio == 0. && m.far6 && m.far3 &&
                           m.far6_cl && m.order;
    p.kind = clustered ? 2 : (policy == 0 && mixed && m.nodes32) ? 1 : 0;
    const int64_t nsig = (nsets * nelem + 31) / 32 * 32;
    p.weights = (clustered ? nsig : 0) + (multi ? nsig : 0);
    if (ws_doubles < 0)
        ws_doubles = field_workspace_doubles(policy, m, mixed, fp64_ratio, multi, nsets, npts, range_elems);
    const int64_t avail = std::max<int64_t>(0, ws_doubles - p.weights);
    const int64_t extra = culled_near(policy, m) ? 1 : 0;  // the near pass's slice, after the far pass's
    p.xcd = !multi && xcd_split(policy, m, npts, range_elems) && avail >= kXcdPartialDoubles;
    if (p.xcd) {
        const int64_t r = std::min<int64_t>(kXcdPartialDoubles / (4 * kXcds * kBlock) - 1,
                                            std::max<int64_t>(1, ((nelem + kXcds - 1) / kXcds + range_elems - 1) /
                                                                     range_elems));
        p.ranges = kXcds * r;
        p.slices = p.ranges + extra;
        p.per = (nelem + p.ranges - 1) / p.ranges;
        if (clustered) p.per = (p.per + KEM_CLUSTER - 1) / KEM_CLUSTER * KEM_CLUSTER;
        p.chunk = std::max<int64_t>(kBlock, kXcdPartialDoubles / (4 * p.slices) / kBlock * kBlock);
        p.partials = 4 * p.slices * std::min(p.chunk, npts);
        return p;
    }
    // enough blocks to fill 256 CUs several times over; each range >= 32 elements (one cluster);
    // fewer ranges when their partial sums would not fit the workspace
    const int64_t gx = nblocks(npts);
    int64_t nsplit = (2048 + gx - 1) / gx;
    if (nsplit > (nelem + 31) / 32) nsplit = (nelem + 31) / 32;
    if (nsplit < 1) nsplit = 1;
    while (nsplit > 1 && 4 * (nsplit + extra) * npts > avail) nsplit--;
    p.per = (nelem + nsplit - 1) / nsplit;
    if (clustered) p.per = (p.per + KEM_CLUSTER - 1) / KEM_CLUSTER * KEM_CLUSTER;
    p.ranges = (nelem + p.per - 1) / p.per;
    p.slices = p.ranges + extra;
    p.chunk = npts;
    // one range and no near slice: the passes write the outputs directly
    p.partials = p.slices > 1 ? 4 * p.slices * npts : 0;
    return p;
}

int64_t field_workspace_doubles(int policy, const FieldMesh& m, bool mixed, double fp64_ratio, bool multi,
                                int64_t nsets, int64_t npts, int64_t range_elems)
{
    const int64_t nsig = (nsets * m.n + 31) / 32 * 32;
    int64_t partial = std::max(kPartialDoubles, 8 * npts);
    if (!multi && xcd_split(policy, m, npts, range_elems)) partial = std::max(partial, kXcdPartialDoubles);
    return partial + 2 * nsig;
}

hipError_t launch_field(hipStream_t st, int policy, const FieldMesh& m, bool mixed, double fp64_ratio,
                        const double* sigma, const int32_t* pset, int64_t nsets, const double* px, const double* py,
                        const double* pz, int64_t npts, double* phi, double* ex, double* ey, double* ez, double* ws,
                        int64_t ws_doubles, int64_t range_elems, hipEvent_t far_begin, hipEvent_t far_end,
                        const SideStream* side)
{
    if (npts <= 0) return hipSuccess;
    if (nsets < 1 || !phi || !ex || !ey || !ez) return hipErrorInvalidValue;
    if (side && side->st && (!side->fork || !side->join)) return hipErrorInvalidValue;
    hipError_t e;
    if (pset)
        e = field_passes<true>(st, policy, m, mixed, fp64_ratio, sigma, pset, nsets, px, py, pz, npts, phi, ex, ey,
                               ez, ws, ws_doubles, range_elems, far_begin, far_end, side);
    else
        e = field_passes<false>(st, policy, m, mixed, fp64_ratio, sigma, nullptr, 1, px, py, pz, npts, phi, ex, ey,
                                ez, ws, ws_doubles, range_elems, far_begin, far_end, side);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

hipError_t launch_check_sets(hipStream_t st, const int32_t* pset, int64_t n, int64_t nsets, int* bad)
{
    if (n <= 0) return hipSuccess;
    k_check_sets<<<nblocks(n), kBlock, 0, st>>>(pset, n, nsets, bad);
    return hipGetLastError();
}

hipError_t launch_census(hipStream_t st, const double* geom, int64_t nelem, const double* px, const double* py,
                         const double* pz, int64_t npts, int mixed, double fp64_ratio, unsigned long long* counts)
{
    if (npts <= 0 || nelem <= 0) return hipSuccess;
    const int64_t per = 4096;
    dim3 grid(nblocks(npts), (unsigned)((nelem + per - 1) / per));
    k_census<<<grid, kBlock, 0, st>>>(geom, nelem, per, px, py, pz, npts, mixed, fp64_ratio, counts);
    return hipGetLastError();
}

hipError_t launch_census_paths(hipStream_t st, const FieldMesh& m, const double* px, const double* py,
                               const double* pz, int64_t npts, unsigned long long* counts)
{
    if (npts <= 0 || m.n <= 0) return hipSuccess;
    if (!m.far6 || !m.far6_cl || !m.order) return hipErrorInvalidValue;
    const int64_t ncl = (m.n + KEM_CLUSTER - 1) / KEM_CLUSTER;
    k_census_far32c<<<nblocks(npts), kBlock, 0, st>>>(m.geom, m.far6, m.far6_cl, m.order, m.n, ncl, px, py, pz, npts,
                                                      counts);
    return hipGetLastError();
}

hipError_t launch_pairs(hipStream_t st, int policy, const double* geom, const double* nodes, const double* recs,
                        int64_t npairs, const int64_t* elem, const double* P, double* out, bool axisym,
                        const int64_t* gstart)
{
    if (npairs <= 0) return hipSuccess;
    if (axisym)
        k_pairs<2><<<nblocks(npairs), kBlock, 0, st>>>(geom, nodes, recs, policy, npairs, elem, P, out, gstart);
    else if (policy >= 2)
        k_pairs<1><<<nblocks(npairs), kBlock, 0, st>>>(geom, nodes, recs, policy, npairs, elem, P, out, gstart);
    else
        k_pairs<0><<<nblocks(npairs), kBlock, 0, st>>>(geom, nodes, recs, policy, npairs, elem, P, out, gstart);
    return hipGetLastError();
}

namespace {
// targets per lane of the Dirichlet cubature kernel (100,000-triangle assembly, both kernels,
// plain node table: 108 ms at one target per lane, 100-102 ms at 2, 106 ms at 3, 108 ms at 4)
constexpr int kAsmTargets = 2;

template <bool G>
hipError_t assemble_kernels(hipStream_t st, int policy, bool field, const double* geom, const double* nodes,
                            const double* recs, const double* targets, int64_t n, double* A64, float* A32,
                            bool axisym, const int64_t* gstart, const double* clusters, const int32_t* order,
                            double* anodes)
{
    if (n <= 0) return hipSuccess;
    const unsigned gx = nblocks(n);
    int64_t nsplit = (4096 + gx - 1) / gx;
    if (nsplit > n) nsplit = n;
    const int64_t per = (n + nsplit - 1) / nsplit;
    nsplit = (n + per - 1) / per;
    dim3 grid(gx, (unsigned)nsplit);
    if (!G && !axisym && policy == 0 && clusters && order) {
        // cubature entries and closed-form entries in separate kernels (measured on the
        // 20,000-triangle coax: 12.8 ms in one kernel at 129 VGPRs).  The cubature kernel's
        // grid: enough column splits for ~16 rounds of wavefronts over 256 CUs x 32 slots, so
        // the last round's partial occupancy costs little (4,096 blocks were 2.1 rounds, a
        // third of the last one idle)
        int64_t sp = (32768 + gx - 1) / gx;
        if (sp > n) sp = n;
        const int64_t per2 = (n + sp - 1) / sp;
        sp = (n + per2 - 1) / per2;
        grid = dim3(gx, (unsigned)sp);
        const int64_t ncl = (n + KEM_CLUSTER - 1) / KEM_CLUSTER;
        if (field) {
            k_assemble_far<true><<<grid, kBlock, 0, st>>>(geom, nodes, targets, n, per2, A64, A32);
            k_assemble_near<true><<<gx, kBlock, 0, st>>>(geom, targets, clusters, ncl, order, n, A64, A32);
        } else {
            // kAsmTargets targets per lane: fewer row blocks, more column splits
            constexpr int TT = kAsmTargets;
            const unsigned gx2 = (unsigned)((n + TT * kBlock - 1) / (TT * kBlock));
            int64_t sp2 = (32768 + gx2 - 1) / gx2;
            if (sp2 > n) sp2 = n;
            const int64_t per3 = (n + sp2 - 1) / sp2;
            sp2 = (n + per3 - 1) / per3;
            if (!anodes) return hipErrorInvalidValue;
            k_asm_nodes<<<nblocks(n * KEM_NODE_SLOTS), kBlock, 0, st>>>(geom, nodes, n, anodes);
            k_assemble_far_pot<TT><<<dim3(gx2, (unsigned)sp2), kBlock, 0, st>>>(geom, nodes, anodes, targets, n, per3,
                                                                              A64, A32);
            k_assemble_near<false><<<gx, kBlock, 0, st>>>(geom, targets, clusters, ncl, order, n, A64, A32);
        }
        return hipGetLastError();
    }
    if (axisym) {
        if (field)
            k_assemble<true, 2, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
        else
            k_assemble<false, 2, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
    } else if (policy >= 2) {
        if (field)
            k_assemble<true, 1, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
        else
            k_assemble<false, 1, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
    } else if (field) {
        k_assemble<true, 0, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
    } else {
        k_assemble<false, 0, G><<<grid, kBlock, 0, st>>>(geom, nodes, recs, targets, policy, n, per, A64, A32, gstart);
    }
    return hipGetLastError();
}
}  // namespace

int64_t assemble_nodes_doubles(int64_t n) { return n * KEM_ANODE_STRIDE; }

hipError_t launch_assemble(hipStream_t st, int policy, bool field, const double* geom, const double* nodes,
                           const double* recs, const double* targets, int64_t n, double* A64, float* A32,
                           bool axisym, const int64_t* gstart, const double* clusters, const int32_t* order,
                           double* anodes)
{
    if (gstart)
        return assemble_kernels<true>(st, policy, field, geom, nodes, recs, targets, n, A64, A32, axisym, gstart,
                                      nullptr, nullptr, nullptr);
    return assemble_kernels<false>(st, policy, field, geom, nodes, recs, targets, n, A64, A32, axisym, nullptr,
                                   clusters, order, anodes);
}

hipError_t launch_accumulate4(hipStream_t st, const double* t, int64_t npts, double* phi, double* ex, double* ey,
                              double* ez)
{
    if (npts <= 0) return hipSuccess;
    k_accumulate4<<<nblocks(npts), kBlock, 0, st>>>(t, npts, phi, ex, ey, ez);
    return hipGetLastError();
}

hipError_t launch_expand_groups(hipStream_t st, const int32_t* gof, int64_t nm, int64_t n, int64_t nsets,
                                const double* sigma, double* out)
{
    if (nm <= 0 || nsets <= 0) return hipSuccess;
    k_expand_groups<<<nblocks(nm * nsets), kBlock, 0, st>>>(gof, nm, n, nsets, sigma, out);
    return hipGetLastError();
}

void set_launch_tuning(const LaunchTuning& t) { g_tuning = t; }
const LaunchTuning& launch_tuning() { return g_tuning; }

int gemv_splits(int64_t nrows, int64_t ncols)
{
    const int64_t gx = (nrows + kBlock - 1) / kBlock;
    int64_t s = (g_tuning.gemv_blocks + gx - 1) / gx;
    if (s > (ncols + 15) / 16) s = (ncols + 15) / 16;
    if (s < 1) s = 1;
    if (s > 512) s = 512;
    return (int)s;
}

template <typename T>
hipError_t launch_gemv(hipStream_t st, const T* A, int64_t lda, int64_t nrows, int64_t ncols, const double* x,
                       double alpha, double beta, double* y, double* part, const int* skip)
{
    int nsplit = gemv_splits(nrows, ncols);
    const int64_t per = (ncols + nsplit - 1) / nsplit;
    nsplit = (int)((ncols + per - 1) / per);
    constexpr int R = 8 / (int)sizeof(T);  // 8-byte lanes: 68 % of HBM against 62 % at 16 bytes
    if (sizeof(T) == 4 && nrows % R == 0 && lda % R == 0 && ((uintptr_t)A % 16) == 0) {
        dim3 grid(nblocks(nrows / R), (unsigned)nsplit);
        k_gemv_part_vec<T, R><<<grid, kBlock, 0, st>>>(A, lda, nrows, ncols, per, x, part, skip);
    } else {
        dim3 grid(nblocks(nrows), (unsigned)nsplit);
        k_gemv_part<T><<<grid, kBlock, 0, st>>>(A, lda, nrows, ncols, per, x, part, skip);
    }
    k_gemv_reduce<<<nblocks(nrows), kBlock, 0, st>>>(part, nsplit, nrows, alpha, beta, y, skip);
    return hipGetLastError();
}
template hipError_t launch_gemv<double>(hipStream_t, const double*, int64_t, int64_t, int64_t, const double*, double,
                                        double, double*, double*, const int*);
template hipError_t launch_gemv<float>(hipStream_t, const float*, int64_t, int64_t, int64_t, const double*, double,
                                       double, double*, double*, const int*);

template <typename T>
hipError_t launch_gemv_t(hipStream_t st, const T* A, int64_t lda, int64_t nrows, int64_t ncols, const double* x,
                         double* y, const int* skip)
{
    k_gemv_t<T><<<(unsigned)ncols, kBlock, 0, st>>>(A, lda, nrows, x, y, skip);
    return hipGetLastError();
}
template hipError_t launch_gemv_t<double>(hipStream_t, const double*, int64_t, int64_t, int64_t, const double*,
                                          double*, const int*);
template hipError_t launch_gemv_t<float>(hipStream_t, const float*, int64_t, int64_t, int64_t, const double*,
                                         double*, const int*);

int arn_proj_splits(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>(n / 4096, kArnProjSplits)); }

hipError_t launch_arnoldi_step(hipStream_t st, const ArnoldiState& a)
{
    const int64_t n = a.n, ncap = (int64_t)a.kmax + 1;
    int nsplit = gemv_splits(n, ncap);
    const int64_t per = (ncap + nsplit - 1) / nsplit;
    nsplit = (int)((ncap + per - 1) / per);
    const int* ints = a.ints;
    double* gpart = a.part + 256;
    const int ps = arn_proj_splits(n);
    const int64_t pper = (n + ps - 1) / ps;
    // CGS2: h1 = V^T w, w -= V h1; h2 = V^T w, w -= V h2 (columns 0..j)
    for (double* h : {a.h1, a.h2}) {
        k_arn_proj<<<dim3((unsigned)ncap, (unsigned)ps), kBlock, 0, st>>>(a.V, n, a.w, a.hp, pper, ints);
        k_arn_hsum<<<1, kBlock, 0, st>>>(a.hp, ps, h, ints);
        k_arn_update_part<<<dim3(nblocks(n), (unsigned)nsplit), kBlock, 0, st>>>(a.V, n, per, h, gpart, ints);
        k_gemv_reduce<<<nblocks(n), kBlock, 0, st>>>(gpart, nsplit, n, -1.0, 1.0, a.w, ints + 1);
    }
    const int nb = (int)std::max<int64_t>(1, std::min<int64_t>(kDotParts, nblocks(n)));
    k_dot_part_skip<<<nb, kBlock, 0, st>>>(a.w, a.w, n, a.part, ints + 1);
    k_dot_final_skip<<<1, kBlock, 0, st>>>(a.part, nb, a.nrm, ints + 1);
    k_arn_normalize<<<nblocks(n), kBlock, 0, st>>>(a.w, a.nrm, a.V, a.vcur, n, ints);
    k_arn_givens<<<1, 64, 0, st>>>(a.h1, a.h2, a.nrm, a.H, a.kmax + 1, a.cs, a.sn, a.p, a.ints, a.scal);
    return hipGetLastError();
}


hipError_t launch_dot(hipStream_t st, const double* a, const double* b, int64_t n, double* part, double* out,
                      bool want_sqrt)
{
    int nb = (int)((n + kBlock - 1) / kBlock);
    if (nb > kDotParts) nb = kDotParts;
    if (nb < 1) nb = 1;
    k_dot_part<<<nb, kBlock, 0, st>>>(a, b, n, part);
    k_dot_final<<<1, kBlock, 0, st>>>(part, nb, want_sqrt ? 1 : 0, out);
    return hipGetLastError();
}

hipError_t launch_axpy_dev(hipStream_t st, const double* alpha, double sign, const double* x, double* y, int64_t n)
{
    k_axpy_dev<<<nblocks(n), kBlock, 0, st>>>(alpha, sign, x, y, n);
    return hipGetLastError();
}

hipError_t launch_div_dev(hipStream_t st, const double* d, const double* x, double* y, int64_t n)
{
    k_div_dev<<<nblocks(n), kBlock, 0, st>>>(d, x, y, n);
    return hipGetLastError();
}

hipError_t launch_axpby(hipStream_t st, double a, const double* x, double b, const double* y, double* z, int64_t n)
{
    k_axpby<<<nblocks(n), kBlock, 0, st>>>(a, x, b, y, z, n);
    return hipGetLastError();
}

hipError_t launch_gauss_jordan(hipStream_t st, double* M, int64_t n, double* f, double* x)
{
    if (n <= 0) return hipSuccess;
    const unsigned gx = nblocks(n);
    const unsigned gy = (unsigned)std::min<int64_t>(n + 1, 1024);
    for (int64_t col = 0; col < n; col++) {
        k_gj_factors<<<gx, kBlock, 0, st>>>(M, n, col, f);
        k_gj_update<<<dim3(gx, gy), kBlock, 0, st>>>(M, n, col, f);
    }
    k_gj_solution<<<gx, kBlock, 0, st>>>(M, n, x);
    return hipGetLastError();
}

int robin_hood_groups(int64_t n)
{
    if (n <= kRobinHoodGridMin || g_tuning.rh_groups == 1) return 1;
    int64_t g = (n + 1023) / 1024;
    g = std::max<int64_t>(g, (n + kRHRowsMax - 1) / kRHRowsMax);
    return (int)std::min<int64_t>(g, 256);
}

hipError_t launch_robin_hood(hipStream_t st, const double* A, const double* b, double* x, double* bi, int64_t n,
                             int64_t steps, double* rnorm, double* slots, unsigned int* flags)
{
    const int G = robin_hood_groups(n);
    if (G == 1) {
        k_robin_hood<<<1, kRHBlock, 0, st>>>(A, b, x, bi, n, steps, rnorm);
        return hipGetLastError();
    }
    int64_t per = (n + G - 1) / G;
    if (per > kRHRowsMax || !slots || !flags) return hipErrorInvalidValue;
    // arrival counter and fault flag start at zero
    hipError_t e = hipMemsetAsync(flags, 0, 2 * sizeof(unsigned int), st);
    if (e != hipSuccess) return e;
    unsigned int* bar = flags;
    unsigned int* fault = flags + 1;
    void* args[] = {(void*)&A, (void*)&b,     (void*)&x,     (void*)&bi,  (void*)&n,    (void*)&per,
                    (void*)&steps, (void*)&rnorm, (void*)&slots, (void*)&bar, (void*)&fault};
    e = hipLaunchCooperativeKernel((const void*)k_robin_hood_grid, dim3((unsigned)G), dim3(kRHGBlock), args, 0, st);
    if (e != hipSuccess) return e;
    return hipGetLastError();
}

}  // namespace kem
