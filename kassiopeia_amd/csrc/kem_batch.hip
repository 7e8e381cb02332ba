// kem_batch.hip -- multi-right-hand-side kernels: block matrix products on MFMA (FP32 matrix)
// and on the FP64 vector ALU, and the batched Krylov-basis operations of lockstep GMRES.
//
// Layouts.  The matrix is column-major, A(i, j) at A[j*n + i].  Block operands fed to the
// products are row-major "transposed" panels Vt[j*KP + c] (row j of the n x KP panel holds one
// entry per right-hand side c, zero-padded to KP columns).  Per-right-hand-side vectors are
// column-major: W[c*ldw + i].  The Krylov bases are V[(r*(kmax+1) + q)*n + i] (basis vector q of
// right-hand side r).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kem_kernels.h"

namespace kem {
namespace {

constexpr int kBlock = 256;
typedef float f16v __attribute__((ext_vector_type(16)));

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ FP32 MFMA block product
// part[s][i][c] = sum_{j in split s} A32(i, j) * Vt[j][c]   (c < KP = 32*NT)
// One wavefront computes a 64-row x KP tile with v_mfma_f32_32x32x2f32: operand A is 32 rows x
// 2 columns of the matrix (lane l holds row l%32 of column l/32: two 128-byte runs per load),
// operand B is 2 rows x 32 columns of the panel (lane l holds panel row l/32, column l%32), and
// accumulator register v of lane l holds tile row 8*(v/4) + 4*(l/32) + v%4, column l%32 (lane
// layout measured on gfx950).  Four wavefronts per workgroup take consecutive 64-row tiles of the
// same element range; blockIdx.y splits the columns of A so that the grid fills the chip.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_bmm32(const float* __restrict__ A, int64_t n,
                                                  const float* __restrict__ Vt, int64_t per_split,
                                                  float* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int KP = 32 * NT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * 64;
    if (i0 >= n) return;  // whole wavefront beyond the last row
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    const int64_t j1 = std::min<int64_t>(n, j0 + per_split);
    const int lr = lane & 31, lk = lane >> 5;
    const int64_t ra = std::min<int64_t>(i0 + lr, n - 1), rb = std::min<int64_t>(i0 + 32 + lr, n - 1);
    f16v acc0[NT], acc1[NT];
    for (int t = 0; t < NT; t++)
        for (int v = 0; v < 16; v++) {
            acc0[t][v] = 0.f;
            acc1[t][v] = 0.f;
        }
    // KB column pairs are loaded before their MFMAs so that the loads of a batch are in flight
    // together (the kernel is otherwise bound by one memory latency per column pair)
    constexpr int KB = 8;
    int64_t j = j0;
    for (; j + 2 * KB <= j1; j += 2 * KB) {
        float a0[KB], a1[KB], b[KB][NT];
        for (int q = 0; q < KB; q++) {
            const int64_t js = j + 2 * q + lk;
            a0[q] = A[js * n + ra];
            a1[q] = A[js * n + rb];
            for (int t = 0; t < NT; t++) b[q][t] = Vt[js * KP + 32 * t + lr];
        }
        for (int q = 0; q < KB; q++)
            for (int t = 0; t < NT; t++) {
                acc0[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[q], b[q][t], acc0[t], 0, 0, 0);
                acc1[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[q], b[q][t], acc1[t], 0, 0, 0);
            }
    }
    for (; j < j1; j += 2) {
        const int64_t jj = j + lk;
        const bool ok = jj < j1;
        const int64_t js = ok ? jj : j;  // j < j1: a valid column
        float a0 = A[js * n + ra], a1 = A[js * n + rb];
        float b[NT];
        for (int t = 0; t < NT; t++) b[t] = Vt[js * KP + 32 * t + lr];
        if (!ok) {
            a0 = a1 = 0.f;
            for (int t = 0; t < NT; t++) b[t] = 0.f;
        }
        for (int t = 0; t < NT; t++) {
            acc0[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b[t], acc0[t], 0, 0, 0);
            acc1[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b[t], acc1[t], 0, 0, 0);
        }
    }
    float* out = part + (int64_t)blockIdx.y * n * KP;
    for (int t = 0; t < NT; t++)
        for (int v = 0; v < 16; v++) {
            const int64_t r = 8 * (v >> 2) + 4 * lk + (v & 3);
            if (i0 + r < n) out[(i0 + r) * KP + 32 * t + lr] = acc0[t][v];
            if (i0 + 32 + r < n) out[(i0 + 32 + r) * KP + 32 * t + lr] = acc1[t][v];
        }
}

// W[c*ldw + i] = sum_s part[s][i][c] for c < ncols: 64-row tiles summed in FP64 (splits in
// order) and transposed through LDS so that both the reads and the writes are coalesced.
template <int KP>
__global__ __launch_bounds__(kBlock) void k_bmm32_reduce(const float* __restrict__ part, int nsplit, int64_t n,
                                                         int ncols, double* __restrict__ W, int64_t ldw,
                                                         const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sh[KP][65];
    const int64_t i0 = (int64_t)blockIdx.x * 64;
    for (int e = threadIdx.x; e < 64 * KP; e += kBlock) {
        const int ii = e / KP, c = e % KP;
        const int64_t i = i0 + ii;
        double s = 0.;
        if (i < n)
            for (int k = 0; k < nsplit; k++) s += (double)part[((int64_t)k * n + i) * KP + c];
        sh[c][ii] = s;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * KP; e += kBlock) {
        const int c = e / 64, ii = e % 64;
        const int64_t i = i0 + ii;
        if (c < ncols && i < n) W[c * ldw + i] = sh[c][ii];
    }
}

// ------------------------------------------------------------------ FP64 block product (VALU)
// part[s][c][i] = sum_{j in split s} A64(i, j) * Xt[j][c], KB right-hand sides per lane; the
// panel row Xt[j][:] is wave-uniform (scalar loads).
template <int KB>
__global__ __launch_bounds__(kBlock) void k_bmm64(const double* __restrict__ A, int64_t n,
                                                  const double* __restrict__ Xt, int64_t per_split,
                                                  double* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t ic = std::min<int64_t>(i, n - 1);
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    const int64_t j1 = std::min<int64_t>(n, j0 + per_split);
    double acc[KB];
    for (int c = 0; c < KB; c++) acc[c] = 0.;
    for (int64_t j = j0; j < j1; j++) {
        const double a = A[j * n + ic];
        const double* x = Xt + j * KB;
        for (int c = 0; c < KB; c++) acc[c] += a * x[c];
    }
    if (i < n) {
        double* out = part + (int64_t)blockIdx.y * KB * n;
        for (int c = 0; c < KB; c++) out[c * n + i] = acc[c];
    }
}

// W[c*ldw + i] = sum_s part[s][c][i] for c < ncols
__global__ __launch_bounds__(kBlock) void k_bmm64_reduce(const double* __restrict__ part, int nsplit, int kb,
                                                         int64_t n, int ncols, double* __restrict__ W, int64_t ldw,
                                                         const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int c = blockIdx.y;
    if (i >= n || c >= ncols) return;
    double s = 0.;
    for (int k = 0; k < nsplit; k++) s += part[((int64_t)k * kb + c) * n + i];
    W[c * ldw + i] = s;
}

// Vt[i][c] = X[c*ldx + i] (c < ncols), 0 for the padding columns; T = float or double
template <typename T>
__global__ __launch_bounds__(kBlock) void k_panel(const double* __restrict__ X, int64_t ldx, int64_t n, int ncols,
                                                  int kp, T* __restrict__ Vt)
{
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e >= n * kp) return;
    const int64_t i = e / kp;
    const int c = (int)(e % kp);
    Vt[e] = c < ncols ? (T)X[c * ldx + i] : (T)0;
}

// ------------------------------------------------------------------ batched Krylov operations
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

// h[r*ldh + q] = <V_r q, W_r> for q <= j (block per (q, r))
__global__ __launch_bounds__(kBlock) void k_bproj(const double* __restrict__ V, int64_t vstride,
                                                  const double* __restrict__ W, int64_t n, double* __restrict__ h,
                                                  int ldh)
{
    __shared__ double sh[kBlock / 64];
    const int q = blockIdx.x, r = blockIdx.y;
    const double* v = V + r * vstride + (int64_t)q * n;
    const double* w = W + (int64_t)r * n;
    double s = 0.;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) s += v[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) h[r * ldh + q] = s;
}

// Y_r += sign * sum_{q < cnt_r} V_r q * h[r*ldh + q]   (grid: rows x r)
__global__ __launch_bounds__(kBlock) void k_bupdate(const double* __restrict__ V, int64_t vstride,
                                                    const double* __restrict__ h, int ldh,
                                                    const int* __restrict__ cnt, double sign, int64_t n,
                                                    double* __restrict__ Y, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double* v = V + r * vstride + i;
    const double* hr = h + r * ldh;
    const int m = cnt[r];
    double s = 0.;
    for (int q = 0; q < m; q++) s += v[(int64_t)q * n] * hr[q];
    Y[(int64_t)r * n + i] += sign * s;
}

// nrm[r] = ||W_r|| (block per r)
__global__ __launch_bounds__(kBlock) void k_bnorm(const double* __restrict__ W, int64_t n, double* __restrict__ nrm,
                                                  int stride, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sh[kBlock / 64];
    const int r = blockIdx.x;
    const double* w = W + (int64_t)r * n;
    double s = 0.;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) s += w[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) nrm[r * stride] = sqrt(s);
}

// V_r q = W_r * (1 / nrm_r); panel column r = V_r q for active r, 0 otherwise
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bnormalize(const double* __restrict__ W, const double* __restrict__ nrm,
                                                       int stride, const int* __restrict__ active, int64_t n,
                                                       double* __restrict__ V, int64_t vstride, int q, int kp,
                                                       T* __restrict__ Vt)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double d = nrm[r * stride];
    const double v = active[r] ? W[(int64_t)r * n + i] * (1.0 / d) : 0.;
    V[r * vstride + (int64_t)q * n + i] = v;
    Vt[i * kp + r] = (T)v;
}

// Z_r = a X_r + b Y_r over all r (column-major, ld n), elementwise
__global__ __launch_bounds__(kBlock) void k_baxpby(double a, const double* __restrict__ X, double b,
                                                   const double* __restrict__ Y, double* __restrict__ Z, int64_t len)
{
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < len) Z[e] = a * X[e] + b * Y[e];
}

// ---- one lockstep GMRES iteration on device state (ints: [0] j, [1] stop, [2] jlimit, [3] steps)
// h[r*ldh + q] = V_r q . W_r for q <= j
__global__ __launch_bounds__(kBlock) void k_bproj_idx(const double* __restrict__ V, int64_t vstride,
                                                      const double* __restrict__ W, int64_t n, double* __restrict__ h,
                                                      int ldh, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int q = blockIdx.x, r = blockIdx.y;
    if (q > ints[0]) return;
    __shared__ double sh[kBlock / 64];
    const double* v = V + r * vstride + (int64_t)q * n;
    const double* w = W + (int64_t)r * n;
    double s = 0.;
    for (int64_t i = threadIdx.x; i < n; i += kBlock) s += v[i] * w[i];
    s = block_sum(s, sh);
    if (threadIdx.x == 0) h[r * ldh + q] = s;
}

// V_r (j+1) = W_r / nrm_r; panel column r = that for active r, 0 otherwise
template <typename T>
__global__ __launch_bounds__(kBlock) void k_bnormalize_idx(const double* __restrict__ W,
                                                           const double* __restrict__ nrm,
                                                           const int* __restrict__ active, int64_t n,
                                                           double* __restrict__ V, int64_t vstride, int kp,
                                                           T* __restrict__ Vt, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double v = active[r] ? W[(int64_t)r * n + i] * (1.0 / nrm[r]) : 0.;
    V[r * vstride + (int64_t)(ints[0] + 1) * n + i] = v;
    Vt[i * kp + r] = (T)v;
}

// Column j of each active right-hand side's Hessenberg matrix, its Givens rotations and the
// stop test (the single-RHS algorithm per column, KGeneralizedMinimalResidual.hh:178-271);
// j += 1, cnt = j + 1 for the next step; stop when no column is active or j reaches jlimit.
// One block of 64 threads, thread r = right-hand side r.  No contraction: the roundings of
// the host sequence it replaces.
__global__ __launch_bounds__(64) void k_bgivens(const double* __restrict__ h1, const double* __restrict__ h2,
                                                const double* __restrict__ nrm, double* __restrict__ H, int ldh,
                                                double* __restrict__ cs, double* __restrict__ sn,
                                                double* __restrict__ p, const double* __restrict__ bnorm,
                                                const double* __restrict__ tol, double* __restrict__ resid,
                                                int* __restrict__ active, int* __restrict__ done,
                                                int* __restrict__ jr, int* __restrict__ cnt,
                                                int* __restrict__ ints, int nrhs)
{
    _Pragma("clang fp contract(off)");
    if (ints[1]) return;
    const int j = ints[0];
    const int r = threadIdx.x;
    __shared__ int nact;
    if (r == 0) nact = 0;
    __syncthreads();
    if (r < nrhs && active[r]) {
        double* h = H + ((int64_t)r * ldh + j) * ldh;
        const double* a1 = h1 + (int64_t)r * ldh;
        const double* a2 = h2 + (int64_t)r * ldh;
        for (int q = 0; q <= j; q++) h[q] = a1[q] + a2[q];
        const double beta = nrm[r];
        h[j + 1] = beta;
        double* csr = cs + (int64_t)r * ldh;
        double* snr = sn + (int64_t)r * ldh;
        double* pr = p + (int64_t)r * ldh;
        for (int q = 0; q < j; q++) {
            const double a = h[q], bb = h[q + 1];
            h[q] = csr[q] * a + snr[q] * bb;
            h[q + 1] = -1.0 * snr[q] * a + csr[q] * bb;
        }
        const double gam = sqrt(h[j] * h[j] + beta * beta);
        csr[j] = h[j] / gam;
        snr[j] = beta / gam;
        h[j] = gam;
        pr[j + 1] = -1.0 * snr[j] * pr[j];
        pr[j] = csr[j] * pr[j];
        jr[r] = j + 1;
        const double rr = fabs(pr[j + 1]) / bnorm[r];
        resid[r] = rr;
        if (rr <= tol[r] || !(beta > 0.)) {
            active[r] = 0;
            if (rr <= tol[r]) done[r] = 1;
        } else {
            atomicAdd(&nact, 1);
        }
    }
    __syncthreads();
    if (r < nrhs) cnt[r] = j + 2;
    if (r == 0) {
        ints[0] = j + 1;
        ints[3] += 1;
        if (nact == 0 || j + 1 >= ints[2]) ints[1] = 1;
    }
}

// y_r = H_r^{-1} p_r over the jr[r] leading columns (back substitution, host order);
// cnt[r] = jr[r].  One thread per right-hand side.
__global__ __launch_bounds__(64) void k_btrisolve(const double* __restrict__ H, int ldh, const double* __restrict__ p,
                                                  const int* __restrict__ jr, double* __restrict__ Y,
                                                  int* __restrict__ cnt, int nrhs)
{
    _Pragma("clang fp contract(off)");
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= nrhs) return;
    const int m = jr[r];
    const double* Hr = H + (int64_t)r * ldh * ldh;
    const double* pr = p + (int64_t)r * ldh;
    double* yr = Y + (int64_t)r * ldh;
    for (int q = m - 1; q >= 0; q--) {
        double s = pr[q];
        for (int u = q + 1; u < m; u++) s -= Hr[(int64_t)u * ldh + q] * yr[u];
        yr[q] = s / Hr[(int64_t)q * ldh + q];
    }
    cnt[r] = m;
}

int bmm_splits(int64_t n, int64_t tiles)
{
    int64_t s = (launch_tuning().bmm_blocks + tiles - 1) / std::max<int64_t>(tiles, 1);
    s = std::min<int64_t>(s, std::max<int64_t>(1, n / 256));
    return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}

}  // namespace

size_t bmm_workspace_bytes(int64_t n, int nrhs, bool fp32)
{
    const int kp = nrhs <= 32 ? 32 : 64;
    if (fp32) {
        const int64_t tiles = (n + 63) / 64;
        const int s = bmm_splits(n, tiles);
        return sizeof(float) * ((size_t)s * n * kp + (size_t)n * kp) + 256;
    }
    const int kb = nrhs <= 16 ? 16 : 64;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int s = bmm_splits(n, blocks * 4);
    return sizeof(double) * ((size_t)s * n * kb + (size_t)n * kb) + 256;
}

hipError_t launch_bmm(hipStream_t st, const float* A32, const double* A64, int64_t n, const double* X, int64_t ldx,
                      int nrhs, double* Y, int64_t ldy, void* ws, size_t ws_bytes, bool x_is_panel, const int* skip)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    if (nrhs > 64 || ws_bytes < bmm_workspace_bytes(n, nrhs, A32 != nullptr)) return hipErrorInvalidValue;
    if (A32) {
        const int nt = nrhs <= 32 ? 1 : 2, kp = 32 * nt;
        const int64_t tiles = (n + 63) / 64;
        const int s = bmm_splits(n, tiles);
        int64_t per = (n + s - 1) / s;
        per += per & 1;
        const int ns = (int)((n + per - 1) / per);
        float* part = (float*)ws;
        float* Vt = part + (size_t)s * n * kp;
        const float* P = x_is_panel ? (const float*)X : Vt;
        if (!x_is_panel)
            k_panel<float><<<nblocks(n * kp), kBlock, 0, st>>>(X, ldx, n, nrhs, kp, Vt);
        dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)ns);
        if (nt == 1)
            k_bmm32<1><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
        else
            k_bmm32<2><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
        const unsigned gr = (unsigned)((n + 63) / 64);
        if (kp == 32)
            k_bmm32_reduce<32><<<gr, kBlock, 0, st>>>(part, ns, n, nrhs, Y, ldy, skip);
        else
            k_bmm32_reduce<64><<<gr, kBlock, 0, st>>>(part, ns, n, nrhs, Y, ldy, skip);
        return hipGetLastError();
    }
    const int kb = nrhs <= 16 ? 16 : 64;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int s = bmm_splits(n, blocks * 4);
    const int64_t per = (n + s - 1) / s;
    const int ns = (int)((n + per - 1) / per);
    double* part = (double*)ws;
    double* Xt = part + (size_t)s * n * kb;
    const double* P = x_is_panel ? X : Xt;
    if (!x_is_panel) k_panel<double><<<nblocks(n * kb), kBlock, 0, st>>>(X, ldx, n, nrhs, kb, Xt);
    dim3 grid((unsigned)blocks, (unsigned)ns);
    if (kb == 16)
        k_bmm64<16><<<grid, kBlock, 0, st>>>(A64, n, P, per, part, skip);
    else
        k_bmm64<64><<<grid, kBlock, 0, st>>>(A64, n, P, per, part, skip);
    k_bmm64_reduce<<<dim3((unsigned)blocks, (unsigned)nrhs), kBlock, 0, st>>>(part, ns, kb, n, nrhs, Y, ldy, skip);
    return hipGetLastError();
}

hipError_t launch_bproj(hipStream_t st, const double* V, int64_t vstride, const double* W, int64_t n, int nq,
                        int nrhs, double* h, int ldh)
{
    if (nq <= 0 || nrhs <= 0) return hipSuccess;
    k_bproj<<<dim3((unsigned)nq, (unsigned)nrhs), kBlock, 0, st>>>(V, vstride, W, n, h, ldh);
    return hipGetLastError();
}

hipError_t launch_bupdate(hipStream_t st, const double* V, int64_t vstride, const double* h, int ldh, const int* cnt,
                          double sign, int64_t n, int nrhs, double* Y)
{
    if (nrhs <= 0) return hipSuccess;
    k_bupdate<<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(V, vstride, h, ldh, cnt, sign, n, Y, nullptr);
    return hipGetLastError();
}

hipError_t launch_bnorm(hipStream_t st, const double* W, int64_t n, int nrhs, double* nrm, int stride)
{
    if (nrhs <= 0) return hipSuccess;
    k_bnorm<<<(unsigned)nrhs, kBlock, 0, st>>>(W, n, nrm, stride, nullptr);
    return hipGetLastError();
}

hipError_t launch_bnormalize(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                             int64_t n, int nrhs, double* V, int64_t vstride, int q, int kp, float* Vt32,
                             double* Vt64)
{
    if (nrhs <= 0) return hipSuccess;
    dim3 grid(nblocks(n), (unsigned)nrhs);
    if (Vt32)
        k_bnormalize<float><<<grid, kBlock, 0, st>>>(W, nrm, stride, active, n, V, vstride, q, kp, Vt32);
    else
        k_bnormalize<double><<<grid, kBlock, 0, st>>>(W, nrm, stride, active, n, V, vstride, q, kp, Vt64);
    return hipGetLastError();
}

hipError_t launch_batch_arnoldi_step(hipStream_t st, const BatchArnoldi& a)
{
    const int64_t n = a.n;
    const int nrhs = a.nrhs;
    int* stop = a.ints + 1;
    hipError_t e = launch_bmm(st, a.use32 ? a.A32 : nullptr, a.use32 ? nullptr : a.A64, n,
                              a.use32 ? (const double*)a.P32 : a.P64, n, nrhs, a.W, n, a.ws, a.ws_bytes, true, stop);
    if (e != hipSuccess) return e;
    for (double* h : {a.h1, a.h2}) {
        k_bproj_idx<<<dim3((unsigned)(a.kmax + 1), (unsigned)nrhs), kBlock, 0, st>>>(a.V, a.vstride, a.W, n, h, a.ldh,
                                                                                   a.ints);
        k_bupdate<<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(a.V, a.vstride, h, a.ldh, a.cnt, -1.0, n, a.W,
                                                                      stop);
    }
    k_bnorm<<<(unsigned)nrhs, kBlock, 0, st>>>(a.W, n, a.nrm, 1, stop);
    if (a.use32)
        k_bnormalize_idx<float><<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(a.W, a.nrm, a.active, n, a.V,
                                                                                    a.vstride, a.kp, a.P32, a.ints);
    else
        k_bnormalize_idx<double><<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(a.W, a.nrm, a.active, n, a.V,
                                                                                     a.vstride, a.kp, a.P64, a.ints);
    k_bgivens<<<1, 64, 0, st>>>(a.h1, a.h2, a.nrm, a.H, a.ldh, a.cs, a.sn, a.p, a.bnorm, a.tol, a.resid, a.active,
                                a.done, a.jr, a.cnt, a.ints, nrhs);
    return hipGetLastError();
}

hipError_t launch_batch_trisolve(hipStream_t st, const BatchArnoldi& a, double* Y)
{
    k_btrisolve<<<(unsigned)((a.nrhs + 63) / 64), 64, 0, st>>>(a.H, a.ldh, a.p, a.jr, Y, a.cnt, a.nrhs);
    return hipGetLastError();
}

hipError_t launch_baxpby(hipStream_t st, double a, const double* X, double b, const double* Y, double* Z, int64_t len)
{
    if (len <= 0) return hipSuccess;
    k_baxpby<<<nblocks(len), kBlock, 0, st>>>(a, X, b, Y, Z, len);
    return hipGetLastError();
}

}  // namespace kem
