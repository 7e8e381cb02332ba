// kem_sort.hip -- spatial ordering of field points for the field kernels (gfx950).
//
// A wavefront evaluates one element against 64 points at a time and executes every
// quadrature branch that any of its lanes needs.  Points in the caller's order are spatially
// random, so a wavefront typically runs both the 12- and the 33-point loops (and the closed
// form) for one element.  Sorting the points along a Morton (Z-order) curve of their bounding
// box makes the 64 points of a wavefront neighbours, so they almost always share one regime.
// The results are gathered back to the caller's order; each point's value does not depend on
// the ordering.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstdint>

#include "kem_kernels.h"

namespace kem {

namespace {

constexpr int kBlock = 256;
constexpr int kBoxParts = 256;

// per-block min/max of the three coordinates: part[b*6 + {0,1,2}] = min, [3,4,5] = max
__global__ __launch_bounds__(kBlock) void k_bbox_part(const double* __restrict__ x, const double* __restrict__ y,
                                                      const double* __restrict__ z, int64_t n, double* __restrict__ part)
{
    __shared__ double sh[6][kBlock];
    double mn[3] = {1e300, 1e300, 1e300}, mx[3] = {-1e300, -1e300, -1e300};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
        const double v[3] = {x[i], y[i], z[i]};
        for (int k = 0; k < 3; k++) {
            mn[k] = fmin(mn[k], v[k]);
            mx[k] = fmax(mx[k], v[k]);
        }
    }
    for (int k = 0; k < 3; k++) {
        sh[k][threadIdx.x] = mn[k];
        sh[3 + k][threadIdx.x] = mx[k];
    }
    __syncthreads();
    for (int s = kBlock / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int k = 0; k < 3; k++) {
                sh[k][threadIdx.x] = fmin(sh[k][threadIdx.x], sh[k][threadIdx.x + s]);
                sh[3 + k][threadIdx.x] = fmax(sh[3 + k][threadIdx.x], sh[3 + k][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sh[threadIdx.x][0];
}

// spread the low 21 bits of a so that bit k lands at bit 3k
__device__ __forceinline__ uint64_t spread3(uint64_t a)
{
    uint64_t v = a & 0x1fffffull;
    v = (v | (v << 32)) & 0x1f00000000ffffull;
    v = (v | (v << 16)) & 0x1f0000ff0000ffull;
    v = (v | (v << 8)) & 0x100f00f00f00f00full;
    v = (v | (v << 4)) & 0x10c30c30c30c30c3ull;
    v = (v | (v << 2)) & 0x1249249249249249ull;
    return v;
}

// 63-bit Morton key of each point in the box spanned by the block partials (points of several
// charge-density sets sort together: the kernels read the sets element-major, so a wavefront
// may mix sets)
__global__ __launch_bounds__(kBlock) void k_morton(const double* __restrict__ x, const double* __restrict__ y,
                                                   const double* __restrict__ z, int64_t n, const double* __restrict__ part, int nparts,
                                                   uint64_t* __restrict__ key, int32_t* __restrict__ idx)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int b = 0; b < nparts; b++)
        for (int k = 0; k < 3; k++) {
            lo[k] = fmin(lo[k], part[b * 6 + k]);
            hi[k] = fmax(hi[k], part[b * 6 + 3 + k]);
        }
    const double v[3] = {x[i], y[i], z[i]};
    uint64_t code = 0;
    for (int k = 0; k < 3; k++) {
        const double span = hi[k] - lo[k];
        double q = span > 0. ? (v[k] - lo[k]) / span * 2097151.0 : 0.;
        q = q < 0. ? 0. : (q > 2097151.0 ? 2097151.0 : q);
        code |= spread3((uint64_t)q) << k;
    }
    key[i] = code;
    idx[i] = (int32_t)i;
}

__global__ __launch_bounds__(kBlock) void k_gather3(const int32_t* __restrict__ idx, const double* __restrict__ x,
                                                    const double* __restrict__ y, const double* __restrict__ z,
                                                    const int32_t* __restrict__ pset, int64_t n,
                                                    double* __restrict__ sx, double* __restrict__ sy,
                                                    double* __restrict__ sz, int32_t* __restrict__ sset)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int32_t i = idx[k];
    sx[k] = x[i];
    sy[k] = y[i];
    sz[k] = z[i];
    if (pset) sset[k] = pset[i];
}

__global__ __launch_bounds__(kBlock) void k_scatter4(const int32_t* __restrict__ idx, const double* __restrict__ s0,
                                                     const double* __restrict__ s1, const double* __restrict__ s2,
                                                     const double* __restrict__ s3, int64_t n, double* __restrict__ d0,
                                                     double* __restrict__ d1, double* __restrict__ d2,
                                                     double* __restrict__ d3)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= n) return;
    const int32_t i = idx[k];
    if (d0) d0[i] = s0[k];
    if (d1) d1[i] = s1[k];
    if (d2) d2[i] = s2[k];
    if (d3) d3[i] = s3[k];
}

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

size_t cub_temp_bytes(int64_t n)
{
    size_t bytes = 0;
    (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const int32_t*)nullptr, (int32_t*)nullptr, (int)n, 0, 64);
    return bytes;
}

}  // namespace

size_t sort_workspace_bytes(int64_t n)
{
    return align256(8 * n) * 2 + align256(4 * n) * 3 + align256(8 * n) * 7 + align256(8 * 6 * kBoxParts) +
           align256(cub_temp_bytes(n));
}

hipError_t launch_field_sorted(hipStream_t st, int policy, const FieldMesh& m, bool mixed, double fp64_ratio,
                               const double* sigma, const int32_t* pset, int64_t nsets, const double* px,
                               const double* py, const double* pz, int64_t npts, double* phi, double* ex, double* ey,
                               double* ez, void* sort_ws, size_t sort_ws_bytes, double* ws, int64_t ws_doubles,
                               int64_t range_elems, hipEvent_t far_begin, hipEvent_t far_end,
                               const SideStream* side)
{
    if (npts <= 0) return hipSuccess;
    if (npts > INT32_MAX || sort_ws_bytes < sort_workspace_bytes(npts)) return hipErrorInvalidValue;
    if (!phi || !ex || !ey || !ez) return hipErrorInvalidValue;
    char* p = (char*)sort_ws;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align256(bytes);
        return r;
    };
    uint64_t* k0 = (uint64_t*)take(8 * npts);
    uint64_t* k1 = (uint64_t*)take(8 * npts);
    int32_t* i0 = (int32_t*)take(4 * npts);
    int32_t* i1 = (int32_t*)take(4 * npts);
    int32_t* ss = (int32_t*)take(4 * npts);
    double* sx = (double*)take(8 * npts);
    double* sy = (double*)take(8 * npts);
    double* sz = (double*)take(8 * npts);
    double* o[4];
    for (int c = 0; c < 4; c++) o[c] = (double*)take(8 * npts);
    double* box = (double*)take(8 * 6 * kBoxParts);
    size_t tb = cub_temp_bytes(npts);
    void* tmp = take(tb);

    int nb = (int)((npts + kBlock - 1) / kBlock);
    if (nb > kBoxParts) nb = kBoxParts;
    k_bbox_part<<<nb, kBlock, 0, st>>>(px, py, pz, npts, box);
    k_morton<<<nblocks(npts), kBlock, 0, st>>>(px, py, pz, npts, box, nb, k0, i0);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, i1, (int)npts, 0, 63, st);
    if (e != hipSuccess) return e;
    k_gather3<<<nblocks(npts), kBlock, 0, st>>>(i1, px, py, pz, pset, npts, sx, sy, sz, ss);
    e = launch_field(st, policy, m, mixed, fp64_ratio, sigma, pset ? ss : nullptr, nsets, sx, sy, sz, npts, o[0], o[1],
                     o[2], o[3], ws, ws_doubles, range_elems, far_begin, far_end, side);
    if (e != hipSuccess) return e;
    k_scatter4<<<nblocks(npts), kBlock, 0, st>>>(i1, o[0], o[1], o[2], o[3], npts, phi, ex, ey, ez);
    return hipGetLastError();
}

hipError_t launch_census_paths_sorted(hipStream_t st, const FieldMesh& m, const double* px, const double* py,
                                      const double* pz, int64_t npts, void* sort_ws, size_t sort_ws_bytes,
                                      unsigned long long* counts)
{
    if (npts <= 0) return hipSuccess;
    if (npts > INT32_MAX || sort_ws_bytes < sort_workspace_bytes(npts)) return hipErrorInvalidValue;
    // the same carving and ordering as launch_field_sorted
    char* p = (char*)sort_ws;
    auto take = [&](size_t bytes) {
        char* r = p;
        p += align256(bytes);
        return r;
    };
    uint64_t* k0 = (uint64_t*)take(8 * npts);
    uint64_t* k1 = (uint64_t*)take(8 * npts);
    int32_t* i0 = (int32_t*)take(4 * npts);
    int32_t* i1 = (int32_t*)take(4 * npts);
    int32_t* ss = (int32_t*)take(4 * npts);
    double* sx = (double*)take(8 * npts);
    double* sy = (double*)take(8 * npts);
    double* sz = (double*)take(8 * npts);
    for (int c = 0; c < 4; c++) (void)take(8 * npts);
    double* box = (double*)take(8 * 6 * kBoxParts);
    size_t tb = cub_temp_bytes(npts);
    void* tmp = take(tb);

    int nb = (int)((npts + kBlock - 1) / kBlock);
    if (nb > kBoxParts) nb = kBoxParts;
    k_bbox_part<<<nb, kBlock, 0, st>>>(px, py, pz, npts, box);
    k_morton<<<nblocks(npts), kBlock, 0, st>>>(px, py, pz, npts, box, nb, k0, i0);
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, i0, i1, (int)npts, 0, 63, st);
    if (e != hipSuccess) return e;
    k_gather3<<<nblocks(npts), kBlock, 0, st>>>(i1, px, py, pz, nullptr, npts, sx, sy, sz, ss);
    return launch_census_paths(st, m, sx, sy, sz, npts, counts);
}

}  // namespace kem
