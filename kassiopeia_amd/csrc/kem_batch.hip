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
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// FP32 -> BF16 bit pattern, round to nearest even (finite values)
__device__ __forceinline__ uint16_t bf16_bits(float f)
{
    uint32_t u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

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
    // rows i0 + 4 lk + rr, rr = 8 (v / 4) + v % 4 (+ 32 for the second tile): one base address
    // and constant offsets (per-store 64-bit addresses took 135 VGPRs and 2 wavefronts per SIMD)
    float* out = part + (int64_t)blockIdx.y * n * KP + (i0 + 4 * lk) * KP + lr;
    const int64_t lim = n - i0 - 4 * lk;
    for (int t = 0; t < NT; t++)
        for (int v = 0; v < 16; v++) {
            const int rr = 8 * (v >> 2) + (v & 3);
            if (rr < lim) out[rr * KP + 32 * t] = acc0[t][v];
            if (rr + 32 < lim) out[(rr + 32) * KP + 32 * t] = acc1[t][v];
        }
}

// The same product with 16-byte row loads (n % 4 == 0): lane l loads rows i0 + 4 (l % 32) .. + 3
// of column j + l / 32 as one float4 and feeds its four values to four MFMAs, row subtile m
// taking rows i0 + 4 r + m (r = the MFMA row index), so a wavefront covers 128 rows, 32 lanes read
// 512 contiguous bytes per column, and each panel load serves four MFMAs instead of two.
template <int NT>
__global__ __launch_bounds__(kBlock) void k_bmm32q(const float* __restrict__ A, int64_t n,
                                                   const float* __restrict__ Vt, int64_t per_split,
                                                   float* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int KP = 32 * NT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t i0 = ((int64_t)blockIdx.x * 4 + wave) * 128;
    if (i0 >= n) return;  // whole wavefront beyond the last row
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    const int64_t j1 = std::min<int64_t>(n, j0 + per_split);
    const int lr = lane & 31, lk = lane >> 5;
    // rows past n read the last four rows (n % 4 == 0 keeps the float4 aligned); never stored
    const int64_t ra = std::min<int64_t>(i0 + 4 * lr, n - 4);
    const float4* __restrict__ A4 = reinterpret_cast<const float4*>(A);
    f16v acc[4][NT];
    for (int m = 0; m < 4; m++)
        for (int t = 0; t < NT; t++)
            for (int v = 0; v < 16; v++) acc[m][t][v] = 0.f;
    constexpr int KB = 4;
    int64_t j = j0;
    for (; j + 2 * KB <= j1; j += 2 * KB) {
        float4 a[KB];
        float b[KB][NT];
        for (int q = 0; q < KB; q++) {
            const int64_t js = j + 2 * q + lk;
            a[q] = A4[(js * n + ra) >> 2];
            for (int t = 0; t < NT; t++) b[q][t] = Vt[js * KP + 32 * t + lr];
        }
        for (int q = 0; q < KB; q++)
            for (int t = 0; t < NT; t++) {
                acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].x, b[q][t], acc[0][t], 0, 0, 0);
                acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].y, b[q][t], acc[1][t], 0, 0, 0);
                acc[2][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].z, b[q][t], acc[2][t], 0, 0, 0);
                acc[3][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[q].w, b[q][t], acc[3][t], 0, 0, 0);
            }
    }
    for (; j < j1; j += 2) {
        const int64_t jj = j + lk;
        const bool ok = jj < j1;
        const int64_t js = ok ? jj : j;  // j < j1: a valid column
        float4 a = A4[(js * n + ra) >> 2];
        float b[NT];
        for (int t = 0; t < NT; t++) b[t] = ok ? Vt[js * KP + 32 * t + lr] : 0.f;
        for (int t = 0; t < NT; t++) {
            acc[0][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.x, b[t], acc[0][t], 0, 0, 0);
            acc[1][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.y, b[t], acc[1][t], 0, 0, 0);
            acc[2][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.z, b[t], acc[2][t], 0, 0, 0);
            acc[3][t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a.w, b[t], acc[3][t], 0, 0, 0);
        }
    }
    // accumulator v of subtile m: MFMA row 8 (v / 4) + 4 lk + v % 4, matrix row i0 + 4 (that) + m;
    // stored in groups of 4 v (32 rows apart) from a base that advances by 32 rows
    float* out = part + (int64_t)blockIdx.y * n * KP + (i0 + 16 * lk) * KP + lr;
    const int64_t lim = n - i0 - 16 * lk;
    for (int v4 = 0; v4 < 4; v4++) {
        for (int m = 0; m < 4; m++)
            for (int t = 0; t < NT; t++)
                for (int v = 0; v < 4; v++) {
                    const int rr = 4 * v + m;  // row offset within this group of 16 rows
                    if (32 * v4 + rr < lim) out[rr * KP + 32 * t] = acc[m][t][4 * v4 + v];
                }
        out += 32 * KP;
    }
}

constexpr int kReduceRows = 16;  // rows per block of k_bmm32_reduce

// W[c*ldw + i] = sum_s part[s][i][c] for c < ncols, splits summed in order in FP64.  A block
// takes kReduceRows rows: entry e = threadIdx.x + 256 m (row e / KP, column e % KP) is contiguous
// in e for each split, so every split's read is coalesced; a thread loads up to eight splits of
// its kReduceRows KP / 256 entries before adding them, so a reduce is about one round trip to
// memory (one split per round trip: 9 splits of a 20,000 x 64 block in 41 us; 32 rows and
// four splits per round trip: 20 us for 6 splits); the sums are transposed through LDS so that
// the writes run along rows.
template <int KP, typename P = float>
__global__ __launch_bounds__(kBlock) void k_bmm32_reduce(const P* __restrict__ part, int nsplit, int64_t n,
                                                         int ncols, double* __restrict__ W, int64_t ldw,
                                                         const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int R = kReduceRows, E = R * KP / kBlock;
    static_assert(E >= 1, "at least one entry per thread");
    __shared__ double sh[KP][R + 1];
    const int64_t i0 = (int64_t)blockIdx.x * R;
    const int rows = (int)std::min<int64_t>(R, n - i0);
    const int lim = rows * KP;
    const P* base = part + i0 * KP;
    double s[E];
    for (int m = 0; m < E; m++) s[m] = 0.;
    constexpr int KS = 8;
    for (int k0 = 0; k0 < nsplit; k0 += KS) {
        P v[KS][E];
        for (int kk = 0; kk < KS; kk++) {
            const P* p = base + (int64_t)(k0 + kk) * n * KP;
            for (int m = 0; m < E; m++) {
                const int e = threadIdx.x + kBlock * m;
                v[kk][m] = (k0 + kk < nsplit && e < lim) ? p[e] : (P)0;
            }
        }
        for (int kk = 0; kk < KS; kk++)
            if (k0 + kk < nsplit)
                for (int m = 0; m < E; m++) s[m] += (double)v[kk][m];
    }
    for (int m = 0; m < E; m++) {
        const int e = threadIdx.x + kBlock * m;
        sh[e % KP][e / KP] = s[m];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < R * KP; e += kBlock) {
        const int c = e / R, ii = e % R;
        if (c < ncols && ii < rows) W[c * ldw + i0 + ii] = sh[c][ii];
    }
}

// ------------------------------------------------------------------ BF16 MFMA block product
// BF16 matrix storage, packed for the MFMA operand: A16[(g*n + i)*8 + m] = A(i, 8g + m) (row
// i, column 8g + m; columns past n are zero), so lane i of a wavefront loads the 8 consecutive
// columns of its row in one 16-byte load and 32 lanes read 512 contiguous bytes.  Entries are
// rounded to FP32 and then to BF16, each to nearest even.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_pack_bf16(const T* __restrict__ A, int64_t n, int64_t ngroups,
                                                      uint4* __restrict__ A16)
{
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;  // e = g*n + i
    if (e >= ngroups * n) return;
    const int64_t g = e / n, i = e - g * n;
    uint16_t v[8];
    for (int m = 0; m < 8; m++) {
        const int64_t j = 8 * g + m;
        v[m] = j < n ? bf16_bits((float)A[j * n + i]) : (uint16_t)0;
    }
    uint4 w;
    w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
    w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
    w.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
    w.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
    A16[e] = w;
}

// BF16 panel of the right-hand sides: Vp[(g*kp + c)*8 + m] = X(8g + m, c) (zero past n rows and
// ncols columns).  A block takes 2 groups (16 rows) x kp columns through LDS, so the reads run
// down the columns (128 bytes) and the 16-byte writes along the panel.
__global__ __launch_bounds__(kBlock) void k_panel_bf16(const double* __restrict__ X, int64_t ldx, int64_t n,
                                                       int ncols, int kp, int64_t ngroups, uint4* __restrict__ Vp)
{
    __shared__ float sh[64][17];
    const int64_t g0 = (int64_t)blockIdx.x * 2, i0 = g0 * 8;
    for (int e = threadIdx.x; e < kp * 16; e += kBlock) {
        const int c = e / 16, r = e % 16;
        const int64_t i = i0 + r;
        sh[c][r] = (c < ncols && i < n) ? (float)X[c * ldx + i] : 0.f;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 2 * kp; e += kBlock) {
        const int g = e / kp, c = e % kp;
        if (g0 + g >= ngroups) continue;
        uint16_t v[8];
        for (int m = 0; m < 8; m++) v[m] = bf16_bits(sh[c][8 * g + m]);
        uint4 w;
        w.x = (uint32_t)v[0] | ((uint32_t)v[1] << 16);
        w.y = (uint32_t)v[2] | ((uint32_t)v[3] << 16);
        w.z = (uint32_t)v[4] | ((uint32_t)v[5] << 16);
        w.w = (uint32_t)v[6] | ((uint32_t)v[7] << 16);
        Vp[(g0 + g) * kp + c] = w;
    }
}


// ------------------------------------------------------------------ block products streamed through LDS
// 64 right-hand sides.  A workgroup takes 256 rows (wavefront w: rows 64 w .. 64 w + 63, the
// 64 x 64 tile of the kernels above, same accumulator layout) and a range of columns.  The
// matrix and the panel stream in chunks into three LDS buffers by direct-to-LDS loads
// (global_load_lds_dwordx4: lane l writes its 16 bytes at the buffer base + 16 l), two chunks
// ahead of the MFMAs.  Per chunk a wavefront waits for its own loads of this chunk (those of
// the next chunk stay in flight), meets the others at a bare barrier (a fence would also wait
// for the loads in flight), reads its operands into registers, issues the loads two chunks
// ahead into the buffer every wavefront finished reading one chunk ago, then runs the MFMAs.
// Loads into LDS hold no registers, so two chunks are in flight at any occupancy.  Measured on
// a 20,000-unknown matrix: BF16 0.148 ms at 3 column splits (the register-loaded k_bmm16: 0.150
// ms at 9 splits, three times the partial sums), FP32 0.569 ms at 6 splits (k_bmm32: 0.61-0.65).

// s_waitcnt with vmcnt = N and the other counters left open (gfx9 encoding)
template <int N>
__device__ __forceinline__ void wait_vmcnt()
{
    static_assert(N >= 0 && N < 64, "vmcnt");
    __builtin_amdgcn_s_waitcnt((N & 0xF) | ((N >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

constexpr int kStreamRows = 256;  // rows per workgroup
constexpr int kStreamKC = 16;     // FP32: matrix columns per chunk (3 buffers: 60 KB of LDS)
constexpr int kStreamG = 4;       // BF16: column groups of 8 per chunk (3 buffers: 60 KB)

// FP32 matrix (n % 4 == 0, A and Vt 16-byte aligned), KP = 64.  Chunk: KC columns x 256 rows
// of A (one 1 KB load per column: lane l takes rows 4l .. 4l + 3) and KC panel rows (4 per 1 KB
// load).  Splits start at multiples of KC columns.
__global__ __launch_bounds__(kBlock) void k_bmm32s(const float* __restrict__ A, int64_t n,
                                                   const float* __restrict__ Vt, int64_t per_split,
                                                   float* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int KP = 64, KC = kStreamKC, CH = KC * kStreamRows + KC * KP, NW = KC / 4 + KC / 16;
    static_assert(KC % 16 == 0, "one panel load per wavefront and chunk");
    __shared__ float lds[3 * CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lk = lane >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * kStreamRows;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    const int64_t j1 = std::min<int64_t>(n, j0 + per_split);
    if (j0 >= j1) return;
    const int64_t nch = (j1 - j0 + KC - 1) / KC;
    // this lane's 4 rows of a column (rows past n read the last 4 rows; never stored)
    const int64_t rr = std::min<int64_t>(r0 + 4 * lane, n - 4);
    auto issue = [&](int64_t k) {
        float* buf = lds + (k % 3) * CH;
        const int64_t jc = j0 + k * KC;
        for (int q = 0; q < KC / 4; q++) {
            const int c = wave + 4 * q;
            const int64_t col = std::min<int64_t>(jc + c, j1 - 1);  // past j1: a valid column, unused
            __builtin_amdgcn_global_load_lds(A + col * n + rr, buf + c * kStreamRows, 16, 0, 0);
        }
        for (int q = 0; q < KC / 16; q++) {
            // panel rows jc + 16 q + 4 wave .. + 3; jc and n are multiples of 4, so a group of
            // rows is either below n or wholly past it (then clamped, unused)
            const int rb = 16 * q + 4 * wave;
            const int64_t row = std::min<int64_t>(jc + rb, n - 4);
            __builtin_amdgcn_global_load_lds(Vt + row * KP + lane * 4, buf + KC * kStreamRows + rb * KP, 16, 0, 0);
        }
    };
    f16v acc0[2], acc1[2];
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < 16; v++) {
            acc0[t][v] = 0.f;
            acc1[t][v] = 0.f;
        }
    issue(0);
    if (nch > 1) issue(1);
    for (int64_t k = 0; k < nch; k++) {
        if (k + 1 < nch)
            wait_vmcnt<NW>();
        else
            wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        const float* buf = lds + (k % 3) * CH;
        const int kv = (int)std::min<int64_t>(KC, j1 - (j0 + k * KC));  // valid columns
        float a0[KC / 2], a1[KC / 2], b[KC / 2][2];
        for (int p = 0; p < KC / 2; p++) {
            const int col = 2 * p + lk;
            a0[p] = buf[col * kStreamRows + wave * 64 + lr];
            a1[p] = buf[col * kStreamRows + wave * 64 + 32 + lr];
            b[p][0] = buf[KC * kStreamRows + col * KP + lr];
            b[p][1] = buf[KC * kStreamRows + col * KP + 32 + lr];
            if (col >= kv) a0[p] = a1[p] = 0.f;
        }
        if (k + 2 < nch) issue(k + 2);
        for (int p = 0; p < KC / 2; p++)
            for (int t = 0; t < 2; t++) {
                acc0[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0[p], b[p][t], acc0[t], 0, 0, 0);
                acc1[t] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1[p], b[p][t], acc1[t], 0, 0, 0);
            }
    }
    const int64_t i0 = r0 + wave * 64;
    if (i0 >= n) return;
    float* out = part + (int64_t)blockIdx.y * n * KP + (i0 + 4 * lk) * KP + lr;
    const int64_t lim = n - i0 - 4 * lk;
    for (int t = 0; t < 2; t++)
        for (int v = 0; v < 16; v++) {
            const int rr2 = 8 * (v >> 2) + (v & 3);
            if (rr2 < lim) out[rr2 * KP + 32 * t] = acc0[t][v];
            if (rr2 + 32 < lim) out[(rr2 + 32) * KP + 32 * t] = acc1[t][v];
        }
}

// BF16 matrix, KP = 32 NT (NT = 1: 2-32 right-hand sides, 2: 33-64).  Chunk: G column groups x
// 256 rows of A16 (one 1 KB load per group and wavefront: lane l takes row 64 w + l) and the G
// groups of the panel (one load per group, its lanes past KP idle); the panel is read from LDS
// by all four wavefronts.
template <int NT>
__device__ __forceinline__ void bmm16s_issue(const bf8v* __restrict__ A16, int64_t n, const bf8v* __restrict__ Vp,
                                             bf8v* lds, int64_t row, int64_t g0, int64_t g1, int64_t k, int lane,
                                             int wave)
{
    constexpr int KP = 32 * NT, G = kStreamG, CH = G * kStreamRows + G * KP;
    bf8v* buf = lds + (k % 3) * CH;
    const int64_t gc = g0 + k * G;
    for (int q = 0; q < G; q++) {
        const int64_t g = std::min<int64_t>(gc + q, g1 - 1);  // past g1: a valid group, unused
        __builtin_amdgcn_global_load_lds(A16 + g * n + row, buf + q * kStreamRows + 64 * wave, 16, 0, 0);
    }
    for (int q = wave; q < G; q += 4) {
        const int64_t g = std::min<int64_t>(gc + q, g1 - 1);
        if (lane < KP) __builtin_amdgcn_global_load_lds(Vp + g * KP + lane, buf + G * kStreamRows + q * KP, 16, 0, 0);
    }
}

template <int NT>
__global__ __launch_bounds__(kBlock) void k_bmm16s(const bf8v* __restrict__ A16, int64_t n,
                                                   const bf8v* __restrict__ Vp, int64_t ngroups, int64_t gper,
                                                   float* __restrict__ part, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int KP = 32 * NT, G = kStreamG, CH = G * kStreamRows + G * KP, NW = G + G / 4;
    static_assert(G % 4 == 0, "panel loads spread over the four wavefronts");
    __shared__ bf8v lds[3 * CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 31, lk = lane >> 5;
    const int64_t r0 = (int64_t)blockIdx.x * kStreamRows;
    const int64_t g0 = (int64_t)blockIdx.y * gper;
    const int64_t g1 = std::min<int64_t>(ngroups, g0 + gper);
    if (g0 >= g1) return;
    const int64_t nch = (g1 - g0 + G - 1) / G;
    const int64_t row = std::min<int64_t>(r0 + 64 * wave + lane, n - 1);  // past n: unused
    f16v acc0[NT], acc1[NT];
    for (int t = 0; t < NT; t++)
        for (int v = 0; v < 16; v++) {
            acc0[t][v] = 0.f;
            acc1[t][v] = 0.f;
        }
    bmm16s_issue<NT>(A16, n, Vp, lds, row, g0, g1, 0, lane, wave);
    if (nch > 1) bmm16s_issue<NT>(A16, n, Vp, lds, row, g0, g1, 1, lane, wave);
    for (int64_t k = 0; k < nch; k++) {
        if (k + 1 < nch)
            wait_vmcnt<NW>();
        else
            wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        const bf8v* buf = lds + (k % 3) * CH;
        const int kv = (int)std::min<int64_t>(G, g1 - (g0 + k * G));  // valid groups
        bf8v a0[G / 2], a1[G / 2], b[G / 2][NT];
        for (int p = 0; p < G / 2; p++) {
            const int q = 2 * p + lk;
            a0[p] = buf[q * kStreamRows + wave * 64 + lr];
            a1[p] = buf[q * kStreamRows + wave * 64 + 32 + lr];
            for (int t = 0; t < NT; t++) b[p][t] = buf[G * kStreamRows + q * KP + 32 * t + lr];
            if (q >= kv)
                for (int m = 0; m < 8; m++) {
                    a0[p][m] = (__bf16)0.f;
                    a1[p][m] = (__bf16)0.f;
                }
        }
        if (k + 2 < nch) bmm16s_issue<NT>(A16, n, Vp, lds, row, g0, g1, k + 2, lane, wave);
        for (int p = 0; p < G / 2; p++)
            for (int t = 0; t < NT; t++) {
                acc0[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0[p], b[p][t], acc0[t], 0, 0, 0);
                acc1[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1[p], b[p][t], acc1[t], 0, 0, 0);
            }
    }
    const int64_t i0 = r0 + wave * 64;
    if (i0 >= n) return;
    float* out = part + (int64_t)blockIdx.y * n * KP + (i0 + 4 * lk) * KP + lr;
    const int64_t lim = n - i0 - 4 * lk;
    for (int t = 0; t < NT; t++)
        for (int v = 0; v < 16; v++) {
            const int rr2 = 8 * (v >> 2) + (v & 3);
            if (rr2 < lim) out[rr2 * KP + 32 * t] = acc0[t][v];
            if (rr2 + 32 < lim) out[(rr2 + 32) * KP + 32 * t] = acc1[t][v];
        }
}

// FP32 or FP64 matrix, KP = 16, 32 or 64 panel columns, on 16 x 16 MFMA tiles
// (v_mfma_f32_16x16x4f32 / v_mfma_f64_16x16x4f64: operand A is 16 rows x 4 columns of the
// matrix, lane l holding row l % 16 of column l / 16; operand B is 4 panel rows x 16 columns, lane
// l holding panel row l / 16, column l % 16; accumulator register v of lane l holds column l % 16
// of tile row 4 (l / 16) + v in FP32, of tile row l / 16 + 4 v in FP64: layouts measured on
// gfx950).  Sixteen columns need no padding to 32, so 16 right-hand sides
// cost half the MFMA work of the 32 x 32 tiles, and the same kernel serves the FP64 matrix.  A
// workgroup takes 256 rows (wavefront w: rows 64 w .. 64 w + 63, four row tiles x KP / 16 column
// tiles) and streams a range of columns through three LDS buffers as k_bmm32s does.  Chunk:
// 64 bytes of every row, i.e. KC = 16 FP32 or 8 FP64 columns (sixteen 1 KB loads, four per
// wavefront: FP32 lane l takes rows 4 l .. 4 l + 3 of a column, FP64 rows 2 l, 2 l + 1 of one
// half), and the KC panel rows (KC KP elements, a quarter per wavefront, in one load whose
// lanes past the quarter are idle).  Columns sit 64 bytes apart beyond their 256 rows in LDS so
// that the two column halves a wavefront's operand read touches fall in different banks.
// Requires n % (16 / sizeof(T)) == 0 and 16-byte aligned A and Vt.  part[s][i][c] in T.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    // tile row of accumulator register v in lanes 16 lk .. 16 lk + 15
    static constexpr __device__ int row(int lk, int v) { return 4 * lk + v; }
};
template <>
struct Mfma16<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ __forceinline__ v4 step(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static constexpr __device__ int row(int lk, int v) { return lk + 4 * v; }
};

// the loads of chunk k of k_bmmt into its LDS buffer (see there)
template <typename T, int KP>
__device__ __forceinline__ void bmmt_issue(const T* __restrict__ A, int64_t n, const T* __restrict__ Vt, T* lds,
                                           int64_t r0, int64_t j0, int64_t j1, int64_t k, int lane, int wave)
{
    constexpr int SZ = sizeof(T), KC = 64 / SZ, PER = 16 / SZ, LPC = kStreamRows * SZ / 1024;
    constexpr int CS = kStreamRows + KC, PC = KC * KP, PQ = PC / 4, CH = KC * CS + PC;
    T* buf = lds + (k % 3) * CH;
    const int64_t jc = j0 + k * KC;
    for (int q = 0; q < 4; q++) {
        const int u = wave + 4 * q;  // load u of the chunk: column u / LPC, part u % LPC
        const int c = u / LPC, h = u % LPC;
        const int64_t col = std::min<int64_t>(jc + c, j1 - 1);  // past j1: a valid column, unused
        // rows past n read the last rows (never stored)
        const int64_t row = std::min<int64_t>(r0 + 1024 / SZ * h + PER * lane, n - PER);
        __builtin_amdgcn_global_load_lds(A + col * n + row, buf + c * CS + 1024 / SZ * h, 16, 0, 0);
    }
    // this wavefront's quarter of the chunk's panel rows; rows past n read row n - 1 (their
    // matrix operands are zero)
    if (lane * PER < PQ) {
        const int e = wave * PQ + lane * PER;
        int64_t prow = jc + e / KP;
        if (prow > n - 1) prow = n - 1;
        __builtin_amdgcn_global_load_lds(Vt + prow * KP + e % KP, buf + KC * CS + wave * PQ, 16, 0, 0);
    }
}

template <typename T, int KP>
__global__ __launch_bounds__(kBlock, 2) void k_bmmt(const T* __restrict__ A, int64_t n, const T* __restrict__ Vt,
                                                 int64_t per_split, T* __restrict__ part,
                                                 const int* __restrict__ skip)
{
    if (skip && *skip) return;
    constexpr int SZ = sizeof(T), KC = 64 / SZ, PER = 16 / SZ;  // PER: elements per 16-byte lane load
    constexpr int LPC = kStreamRows * SZ / 1024;                 // 1 KB loads per column
    constexpr int CS = kStreamRows + KC;                          // LDS column stride (elements)
    constexpr int PC = KC * KP, PQ = PC / 4;                      // panel elements per chunk, per wavefront
    constexpr int CH = KC * CS + PC, NW = KC * LPC / 4 + 1;
    constexpr int CT = KP / 16;
    static_assert(KC * LPC == 16 && PQ * SZ <= 1024 && PQ % PER == 0 && KP % 16 == 0, "chunk shape");
    typedef typename Mfma16<T>::v4 acc_t;
    __shared__ T lds[3 * CH];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lk = lane >> 4;
    const int64_t r0 = (int64_t)blockIdx.x * kStreamRows;
    const int64_t j0 = (int64_t)blockIdx.y * per_split;
    const int64_t j1 = std::min<int64_t>(n, j0 + per_split);
    if (j0 >= j1) return;
    const int64_t nch = (j1 - j0 + KC - 1) / KC;
    acc_t acc[4][CT];
    for (int m = 0; m < 4; m++)
        for (int t = 0; t < CT; t++)
            for (int v = 0; v < 4; v++) acc[m][t][v] = (T)0;
    bmmt_issue<T, KP>(A, n, Vt, lds, r0, j0, j1, 0, lane, wave);
    if (nch > 1) bmmt_issue<T, KP>(A, n, Vt, lds, r0, j0, j1, 1, lane, wave);
    for (int64_t k = 0; k < nch; k++) {
        if (k + 1 < nch)
            wait_vmcnt<NW>();
        else
            wait_vmcnt<0>();
        __builtin_amdgcn_s_barrier();
        asm volatile("" ::: "memory");
        const T* buf = lds + (k % 3) * CH;
        const int kv = (int)std::min<int64_t>(KC, j1 - (j0 + k * KC));  // valid columns
        T a[KC / 4][4], b[KC / 4][CT];
        for (int p = 0; p < KC / 4; p++) {
            const int col = 4 * p + lk;
            for (int m = 0; m < 4; m++) a[p][m] = col < kv ? buf[col * CS + wave * 64 + 16 * m + lr] : (T)0;
            for (int t = 0; t < CT; t++) b[p][t] = buf[KC * CS + col * KP + 16 * t + lr];
        }
        if (k + 2 < nch) bmmt_issue<T, KP>(A, n, Vt, lds, r0, j0, j1, k + 2, lane, wave);
        for (int p = 0; p < KC / 4; p++)
            for (int m = 0; m < 4; m++)
                for (int t = 0; t < CT; t++) acc[m][t] = Mfma16<T>::step(a[p][m], b[p][t], acc[m][t]);
    }
    const int64_t i0 = r0 + wave * 64;
    if (i0 >= n) return;
    // tile m, register v: row i0 + 16 m + Mfma16<T>::row(lk, v), column 16 t + lr
    T* out = part + (int64_t)blockIdx.y * n * KP + i0 * KP + lr;
    const int64_t lim = n - i0;
    for (int m = 0; m < 4; m++)
        for (int v = 0; v < 4; v++) {
            const int rr = 16 * m + Mfma16<T>::row(lk, v);
            if (rr < lim)
                for (int t = 0; t < CT; t++) out[rr * KP + 16 * t] = acc[m][t][v];
        }
}

// One right-hand side on the BF16 matrix: part[s][i] = sum_{g in split s} sum_m A16(i, 8g + m) *
// x(8g + m) on the vector ALU (the block product would pad one column to 32 MFMA columns).
// Lane = row i: one 16-byte load per column group holds the row's 8 BF16 entries; the group's
// 8 panel values are wave-uniform (one 16-byte scalar load).  FP32 products and sums, as the
// MFMA product accumulates.  Groups [s*gper, (s+1)*gper) per blockIdx.y.
__device__ __forceinline__ float bf16_lo(uint32_t w) { return __uint_as_float(w << 16); }
__device__ __forceinline__ float bf16_hi(uint32_t w) { return __uint_as_float(w & 0xffff0000u); }

__global__ __launch_bounds__(kBlock) void k_gemv16_part(const uint4* __restrict__ A16, int64_t n,
                                                        const uint4* __restrict__ Vp, int kp, int64_t ngroups,
                                                        int64_t gper, float* __restrict__ part,
                                                        const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const int64_t g0 = (int64_t)blockIdx.y * gper;
    const int64_t g1 = g0 + gper < ngroups ? g0 + gper : ngroups;
    float a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f;
    int64_t g = g0;
    for (; g + 2 <= g1; g += 2) {
        const uint4 u = A16[g * n + i], v = A16[(g + 1) * n + i];
        const uint4 x = Vp[g * kp], y = Vp[(g + 1) * kp];  // column 0 of the panel: 8 values
        a0 += bf16_lo(u.x) * bf16_lo(x.x) + bf16_hi(u.x) * bf16_hi(x.x);
        a1 += bf16_lo(u.y) * bf16_lo(x.y) + bf16_hi(u.y) * bf16_hi(x.y);
        a2 += bf16_lo(u.z) * bf16_lo(x.z) + bf16_hi(u.z) * bf16_hi(x.z);
        a3 += bf16_lo(u.w) * bf16_lo(x.w) + bf16_hi(u.w) * bf16_hi(x.w);
        a0 += bf16_lo(v.x) * bf16_lo(y.x) + bf16_hi(v.x) * bf16_hi(y.x);
        a1 += bf16_lo(v.y) * bf16_lo(y.y) + bf16_hi(v.y) * bf16_hi(y.y);
        a2 += bf16_lo(v.z) * bf16_lo(y.z) + bf16_hi(v.z) * bf16_hi(y.z);
        a3 += bf16_lo(v.w) * bf16_lo(y.w) + bf16_hi(v.w) * bf16_hi(y.w);
    }
    for (; g < g1; g++) {
        const uint4 u = A16[g * n + i];
        const uint4 x = Vp[g * kp];
        a0 += bf16_lo(u.x) * bf16_lo(x.x) + bf16_hi(u.x) * bf16_hi(x.x);
        a1 += bf16_lo(u.y) * bf16_lo(x.y) + bf16_hi(u.y) * bf16_hi(x.y);
        a2 += bf16_lo(u.z) * bf16_lo(x.z) + bf16_hi(u.z) * bf16_hi(x.z);
        a3 += bf16_lo(u.w) * bf16_lo(x.w) + bf16_hi(u.w) * bf16_hi(x.w);
    }
    part[(int64_t)blockIdx.y * n + i] = (a0 + a1) + (a2 + a3);
}

// W[i] = sum_s part[s][i] (FP64)
__global__ __launch_bounds__(kBlock) void k_gemv16_reduce(const float* __restrict__ part, int nsplit, int64_t n,
                                                          double* __restrict__ W, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s = 0.;
    for (int k = 0; k < nsplit; k++) s += (double)part[(int64_t)k * n + i];
    W[i] = s;
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
// (a block transposes kPanelRows rows x kp <= 64 columns through LDS: 128-byte column reads,
// row writes contiguous)
constexpr int kPanelRows = 16;
template <typename T>
__global__ __launch_bounds__(kBlock) void k_panel(const double* __restrict__ X, int64_t ldx, int64_t n, int ncols,
                                                  int kp, T* __restrict__ Vt)
{
    __shared__ double sh[64][kPanelRows + 1];
    const int64_t i0 = (int64_t)blockIdx.x * kPanelRows;
    for (int e = threadIdx.x; e < kp * kPanelRows; e += kBlock) {
        const int c = e / kPanelRows, r = e % kPanelRows;
        const int64_t i = i0 + r;
        sh[c][r] = (c < ncols && i < n) ? X[c * ldx + i] : 0.;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < kPanelRows * kp; e += kBlock) {
        const int r = e / kp, c = e % kp;
        const int64_t i = i0 + r;
        if (i < n) Vt[i * kp + c] = (T)sh[c][r];
    }
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

// Y_r += sign * sum_{q < cnt_r} V_r q * h[r*ldh + q]   (grid: rows x r); the basis V in FP64,
// or in FP32 for the inner solves of mixed precision (VT = float), summed in FP64
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_bupdate(const VT* __restrict__ V, int64_t vstride,
                                                    const double* __restrict__ h, int ldh,
                                                    const int* __restrict__ cnt, double sign, int64_t n,
                                                    double* __restrict__ Y, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const VT* v = V + r * vstride + i;
    const double* hr = h + r * ldh;
    const int m = cnt[r];
    // four independent sums, four basis loads in flight per thread
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    int q = 0;
    for (; q + 3 < m; q += 4) {
        s0 += (double)v[(int64_t)q * n] * hr[q];
        s1 += (double)v[(int64_t)(q + 1) * n] * hr[q + 1];
        s2 += (double)v[(int64_t)(q + 2) * n] * hr[q + 2];
        s3 += (double)v[(int64_t)(q + 3) * n] * hr[q + 3];
    }
    for (; q < m; q++) s0 += (double)v[(int64_t)q * n] * hr[q];
    Y[(int64_t)r * n + i] += sign * ((s0 + s1) + (s2 + s3));
}

// nrm[r] = ||W_r|| (block per r)
__global__ __launch_bounds__(kBlock) void k_bnorm(const double* __restrict__ W, int64_t n, double* __restrict__ nrm,
                                                  int stride, const int* __restrict__ skip)
{
    if (skip && *skip) return;
    __shared__ double sh[kBlock / 64];
    const int r = blockIdx.x;
    const double* w = W + (int64_t)r * n;
    // four independent sums, so four loads per thread are in flight (one block per right-hand
    // side is otherwise bound by one load latency per 256 entries)
    double s0 = 0., s1 = 0., s2 = 0., s3 = 0.;
    int64_t i = threadIdx.x;
    for (; i + 3 * kBlock < n; i += 4 * kBlock) {
        const double a = w[i], b = w[i + kBlock], c = w[i + 2 * kBlock], d = w[i + 3 * kBlock];
        s0 += a * a;
        s1 += b * b;
        s2 += c * c;
        s3 += d * d;
    }
    for (; i < n; i += kBlock) s0 += w[i] * w[i];
    double s = block_sum((s0 + s1) + (s2 + s3), sh);
    if (threadIdx.x == 0) nrm[r * stride] = sqrt(s);
}

// Panel entry (row i, column r) of a block operand: FP32 / FP64 panels are row-major n x kp;
// BF16 panels (uint16_t bit patterns) are packed in groups of 8 rows, Vp[((i/8)*kp + r)*8 + i%8],
// so that one 16-byte load holds the 8 consecutive k-values an MFMA lane takes.
template <typename T>
__device__ __forceinline__ int64_t panel_index(int64_t i, int r, int kp)
{
    return i * kp + r;
}
template <>
__device__ __forceinline__ int64_t panel_index<uint16_t>(int64_t i, int r, int kp)
{
    return ((i >> 3) * kp + r) * 8 + (i & 7);
}
template <typename T>
__device__ __forceinline__ T panel_value(double v)
{
    return (T)v;
}
template <>
__device__ __forceinline__ uint16_t panel_value<uint16_t>(double v)
{
    return bf16_bits((float)v);
}

// V_r q = W_r * (1 / nrm_r); panel column r = V_r q for active r, 0 otherwise (the basis in VT)
template <typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_bnormalize(const double* __restrict__ W, const double* __restrict__ nrm,
                                                       int stride, const int* __restrict__ active, int64_t n,
                                                       VT* __restrict__ V, int64_t vstride, int q, int kp,
                                                       T* __restrict__ Vt)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double d = nrm[r * stride];
    const double v = active[r] ? W[(int64_t)r * n + i] * (1.0 / d) : 0.;
    V[r * vstride + (int64_t)q * n + i] = (VT)v;
    Vt[panel_index<T>(i, r, kp)] = panel_value<T>(v);
}

// Z_r = a X_r + b Y_r over all r (column-major, ld n), elementwise
__global__ __launch_bounds__(kBlock) void k_baxpby(double a, const double* __restrict__ X, double b,
                                                   const double* __restrict__ Y, double* __restrict__ Z, int64_t len)
{
    const int64_t e = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (e < len) Z[e] = a * X[e] + b * Y[e];
}

// ---- one lockstep GMRES iteration on device state (ints: [0] j, [1] stop, [2] jlimit, [3] steps)

// Projections of the lockstep Gram-Schmidt, h[r*ldh + q] = <V_r q, W_r> for q <= j = ints[0], in
// two passes.  k_bproj_part: a block per (chunk of kProjRows rows, r) holds W_r's chunk in
// registers (lane l: rows l + 64 k) and wavefront w takes q = w, w + 4, ..., one pass over V_r q
// and a shuffle sum each: part[(c*nrhs + r)*ldh + q].  k_bproj_sum adds the chunks in order.  (A
// block per (q, r), every q up to kmax launched: 42.6 us per pass for 64 right-hand sides of
// 20,000 rows, most of its blocks returning at once.)
constexpr int kProjRows = 1024;
template <typename VT>
__global__ __launch_bounds__(kBlock) void k_bproj_part(const VT* __restrict__ V, int64_t vstride,
                                                       const double* __restrict__ W, int64_t n,
                                                       double* __restrict__ part, int ldh,
                                                       const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int j = min(ints[0], ldh - 2);  // basis vectors 0 .. j, j <= kmax
    const int c = blockIdx.x, r = blockIdx.y, nrhs = gridDim.y;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    constexpr int K = kProjRows / 64;
    const int64_t i0 = (int64_t)c * kProjRows + lane;
    const double* wr = W + (int64_t)r * n;
    double w[K];
    for (int k = 0; k < K; k++) {
        const int64_t i = i0 + 64 * k;
        w[k] = i < n ? wr[i] : 0.;
    }
    const VT* vr = V + r * vstride;
    for (int q = wave; q <= j; q += 4) {
        const VT* v = vr + (int64_t)q * n;
        double s = 0.;
        for (int k = 0; k < K; k++) {
            const int64_t i = i0 + 64 * k;
            if (i < n) s += (double)v[i] * w[k];
        }
        for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
        if (lane == 0) part[((int64_t)c * nrhs + r) * ldh + q] = s;
    }
}

__global__ __launch_bounds__(64) void k_bproj_sum(const double* __restrict__ part, int nchunks, int ldh,
                                                  double* __restrict__ h, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int j = min(ints[0], ldh - 2);
    const int r = blockIdx.x, nrhs = gridDim.x;
    for (int q = threadIdx.x; q <= j; q += 64) {
        // the chunks in order, four loads in flight
        double s = 0.;
        int c = 0;
        for (; c + 3 < nchunks; c += 4) {
            const double a = part[((int64_t)c * nrhs + r) * ldh + q], b = part[((int64_t)(c + 1) * nrhs + r) * ldh + q],
                         d = part[((int64_t)(c + 2) * nrhs + r) * ldh + q], e = part[((int64_t)(c + 3) * nrhs + r) * ldh + q];
            s += a;
            s += b;
            s += d;
            s += e;
        }
        for (; c < nchunks; c++) s += part[((int64_t)c * nrhs + r) * ldh + q];
        h[r * ldh + q] = s;
    }
}

// V_r (j+1) = W_r / nrm_r; panel column r = that for active r, 0 otherwise
template <typename T, typename VT>
__global__ __launch_bounds__(kBlock) void k_bnormalize_idx(const double* __restrict__ W,
                                                           const double* __restrict__ nrm,
                                                           const int* __restrict__ active, int64_t n,
                                                           VT* __restrict__ V, int64_t vstride, int kp,
                                                           T* __restrict__ Vt, const int* __restrict__ ints)
{
    if (ints[1]) return;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= n) return;
    const double v = active[r] ? W[(int64_t)r * n + i] * (1.0 / nrm[r]) : 0.;
    V[r * vstride + (int64_t)(ints[0] + 1) * n + i] = (VT)v;
    Vt[panel_index<T>(i, r, kp)] = panel_value<T>(v);
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

// column splits of the LDS-streamed products: the workgroup count of 256 rows x splits nearest
// bmm_stream_blocks (the measured optima sit at 3 and 6 splits of a 79-workgroup matrix)
int bmm_stream_splits(int64_t n)
{
    const int64_t rb = (n + 255) / 256;
    int64_t s = (launch_tuning().bmm_stream_blocks + rb / 2) / rb;
    s = std::min<int64_t>(s, std::max<int64_t>(1, n / 256));
    return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}

// column splits of k_bmmt for a matrix of sz-byte entries and kp panel columns.  Two
// workgroups fit a compute unit (54-62 KB of LDS each), so the grid of 256-row blocks x splits
// runs in rounds of 2 x CUs workgroups of equal work; of the split counts whose grid stays
// within bmm_tile_blocks, the one of least modelled cost per row: the matrix row (n sz bytes)
// stretched by the idle part of the last round, plus 8 kp sz bytes per split for its partials
// (written, read back and summed; the weight fits the measured optima: on the 20,000-unknown
// coax, 79 row blocks, 6 splits, 0.277 ms for 16 FP32 right-hand sides against 0.291 ms at 19
// and 0.33 ms at 10 or 13; 9 splits on a 100,000-unknown matrix, 7.3 ms against 8.5 ms at 4).
int bmm_tile_splits(int64_t n, int kp, int sz)
{
    const int64_t rb = (n + 255) / 256, cap = 2 * std::max<int64_t>(1, launch_tuning().cus);
    const int64_t smax = std::max<int64_t>(
        1, std::min<int64_t>({launch_tuning().bmm_tile_blocks / rb, std::max<int64_t>(1, n / 256), 64}));
    int64_t best = 1;
    double cost = 0.;
    for (int64_t s = 1; s <= smax; s++) {
        const int64_t g = rb * s;
        const double c = (double)(cap * ((g + cap - 1) / cap)) / (double)g * (double)n * sz + 8.0 * s * kp * sz;
        if (s == 1 || c < cost) {
            cost = c;
            best = s;
        }
    }
    return (int)best;
}

int bmm_splits(int64_t n, int64_t tiles)
{
    int64_t s = (launch_tuning().bmm_blocks + tiles - 1) / std::max<int64_t>(tiles, 1);
    s = std::min<int64_t>(s, std::max<int64_t>(1, n / 256));
    return (int)std::max<int64_t>(1, std::min<int64_t>(s, 64));
}

}  // namespace

int bmm_panel_cols(int64_t n, int nrhs, int prec)
{
    if (prec == kBmmFp64) return nrhs <= 16 ? 16 : 64;
    if (prec == kBmmFp32 && nrhs <= 16 && n % 4 == 0) return 16;  // k_bmmt only
    return nrhs <= 32 ? 32 : 64;
}

size_t bmm_workspace_bytes(int64_t n, int nrhs, int prec)
{
    const int kp = bmm_panel_cols(n, nrhs, prec);
    if (prec == kBmmBf16) {
        const int64_t tiles = (n + 63) / 64, ngroups = (n + 7) / 8;
        const int s = std::max({bmm_splits(n, tiles), kp == 64 ? bmm_stream_splits(n) : 1,
                                kp == 32 ? bmm_tile_splits(n, 32, 2) : 1});
        return sizeof(float) * (size_t)s * n * kp + sizeof(uint16_t) * (size_t)ngroups * 8 * kp + 512;
    }
    if (prec == kBmmFp32) {
        const int64_t tiles = (n + 63) / 64;
        const int s = std::max({bmm_splits(n, tiles), kp == 64 ? bmm_stream_splits(n) : 1,
                                kp <= 32 && n % 4 == 0 ? bmm_tile_splits(n, kp, 4) : 1});
        return sizeof(float) * ((size_t)s * n * kp + (size_t)n * kp) + 256;
    }
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int s = std::max(bmm_splits(n, blocks * 4), n % 2 == 0 ? bmm_tile_splits(n, kp, 8) : 1);
    return sizeof(double) * ((size_t)s * n * kp + (size_t)n * kp) + 256;
}

hipError_t launch_bmm(hipStream_t st, const float* A32, const double* A64, int64_t n, const double* X, int64_t ldx,
                      int nrhs, double* Y, int64_t ldy, void* ws, size_t ws_bytes, bool x_is_panel, const int* skip)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    if (nrhs > 64 || ws_bytes < bmm_workspace_bytes(n, nrhs, A32 ? kBmmFp32 : kBmmFp64)) return hipErrorInvalidValue;
    const int kp = bmm_panel_cols(n, nrhs, A32 ? kBmmFp32 : kBmmFp64);
    if (A32) {
        const int nt = kp == 64 ? 2 : 1;
        const int64_t tiles = (n + 63) / 64;
        const int s = bmm_splits(n, tiles);
        int64_t per = (n + s - 1) / s;
        per += per & 1;
        const int ns = (int)((n + per - 1) / per);
        float* part = (float*)ws;
        // the panel after the partial sums of any of the kernels
        const int sw = std::max({s, nt == 2 ? bmm_stream_splits(n) : 1,
                                 kp <= 32 && n % 4 == 0 ? bmm_tile_splits(n, kp, 4) : 1});
        float* Vt = part + (size_t)sw * n * kp;
        const float* P = x_is_panel ? (const float*)X : Vt;
        if (!x_is_panel)
            k_panel<float><<<(unsigned)((n + kPanelRows - 1) / kPanelRows), kBlock, 0, st>>>(X, ldx, n, nrhs, kp, Vt);
        const bool a16 = n % 4 == 0 && (uintptr_t)A32 % 16 == 0 && (uintptr_t)P % 16 == 0;
        if (kp <= 32 && a16) {
            // 16 x 16 tiles streamed through LDS; splits start at multiples of the chunk width
            const int ss = bmm_tile_splits(n, kp, 4);
            int64_t pc = (n + ss - 1) / ss;
            pc = (pc + 15) / 16 * 16;
            const int nc = (int)((n + pc - 1) / pc);
            dim3 grid((unsigned)((n + kStreamRows - 1) / kStreamRows), (unsigned)nc);
            const unsigned gr = (unsigned)((n + kReduceRows - 1) / kReduceRows);
            if (kp == 16) {
                k_bmmt<float, 16><<<grid, kBlock, 0, st>>>(A32, n, P, pc, part, skip);
                k_bmm32_reduce<16><<<gr, kBlock, 0, st>>>(part, nc, n, nrhs, Y, ldy, skip);
            } else {
                k_bmmt<float, 32><<<grid, kBlock, 0, st>>>(A32, n, P, pc, part, skip);
                k_bmm32_reduce<32><<<gr, kBlock, 0, st>>>(part, nc, n, nrhs, Y, ldy, skip);
            }
            return hipGetLastError();
        }
        if (kp == 16) return hipErrorInvalidValue;  // a 16-column panel needs the 16 x 16 tiles
        if (nt == 2 && a16) {
            // streamed through LDS; splits start at multiples of the chunk width
            const int ss = bmm_stream_splits(n);
            int64_t pc = (n + ss - 1) / ss;
            pc = (pc + kStreamKC - 1) / kStreamKC * kStreamKC;
            const int nc = (int)((n + pc - 1) / pc);
            dim3 grid((unsigned)((n + kStreamRows - 1) / kStreamRows), (unsigned)nc);
            k_bmm32s<<<grid, kBlock, 0, st>>>(A32, n, P, pc, part, skip);
            k_bmm32_reduce<64><<<(unsigned)((n + kReduceRows - 1) / kReduceRows), kBlock, 0, st>>>(part, nc, n, nrhs, Y, ldy, skip);
            return hipGetLastError();
        }
        if (launch_tuning().bmm_rows == 128 && n % 4 == 0 && n >= 4) {
            // 128-row wavefront tiles, 512 rows per block
            dim3 grid((unsigned)((n + 511) / 512), (unsigned)ns);
            if (nt == 1)
                k_bmm32q<1><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
            else
                k_bmm32q<2><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
        } else {
            dim3 grid((unsigned)((tiles + 3) / 4), (unsigned)ns);
            if (nt == 1)
                k_bmm32<1><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
            else
                k_bmm32<2><<<grid, kBlock, 0, st>>>(A32, n, P, per, part, skip);
        }
        const unsigned gr = (unsigned)((n + kReduceRows - 1) / kReduceRows);
        if (kp == 32)
            k_bmm32_reduce<32><<<gr, kBlock, 0, st>>>(part, ns, n, nrhs, Y, ldy, skip);
        else
            k_bmm32_reduce<64><<<gr, kBlock, 0, st>>>(part, ns, n, nrhs, Y, ldy, skip);
        return hipGetLastError();
    }
    const int kb = kp;
    const int64_t blocks = (n + kBlock - 1) / kBlock;
    const int s = bmm_splits(n, blocks * 4);
    const int64_t per = (n + s - 1) / s;
    const int ns = (int)((n + per - 1) / per);
    double* part = (double*)ws;
    const int sw = std::max(s, n % 2 == 0 ? bmm_tile_splits(n, kb, 8) : 1);
    double* Xt = part + (size_t)sw * n * kb;
    const double* P = x_is_panel ? X : Xt;
    if (!x_is_panel)
        k_panel<double><<<(unsigned)((n + kPanelRows - 1) / kPanelRows), kBlock, 0, st>>>(X, ldx, n, nrhs, kb, Xt);
    if (n % 2 == 0 && (uintptr_t)A64 % 16 == 0 && (uintptr_t)P % 16 == 0) {
        // FP64 MFMA on 16 x 16 tiles streamed through LDS
        const int ss = bmm_tile_splits(n, kb, 8);
        int64_t pc = (n + ss - 1) / ss;
        pc = (pc + 7) / 8 * 8;
        const int nc = (int)((n + pc - 1) / pc);
        dim3 tgrid((unsigned)((n + kStreamRows - 1) / kStreamRows), (unsigned)nc);
        const unsigned gr = (unsigned)((n + kReduceRows - 1) / kReduceRows);
        if (kb == 16) {
            k_bmmt<double, 16><<<tgrid, kBlock, 0, st>>>(A64, n, P, pc, part, skip);
            k_bmm32_reduce<16, double><<<gr, kBlock, 0, st>>>(part, nc, n, nrhs, Y, ldy, skip);
        } else {
            k_bmmt<double, 64><<<tgrid, kBlock, 0, st>>>(A64, n, P, pc, part, skip);
            k_bmm32_reduce<64, double><<<gr, kBlock, 0, st>>>(part, nc, n, nrhs, Y, ldy, skip);
        }
        return hipGetLastError();
    }
    dim3 grid((unsigned)blocks, (unsigned)ns);
    if (kb == 16)
        k_bmm64<16><<<grid, kBlock, 0, st>>>(A64, n, P, per, part, skip);
    else
        k_bmm64<64><<<grid, kBlock, 0, st>>>(A64, n, P, per, part, skip);
    k_bmm64_reduce<<<dim3((unsigned)blocks, (unsigned)nrhs), kBlock, 0, st>>>(part, ns, kb, n, nrhs, Y, ldy, skip);
    return hipGetLastError();
}

hipError_t launch_pack_bf16(hipStream_t st, const double* A64, const float* A32, int64_t n, uint16_t* A16)
{
    if (n <= 0) return hipSuccess;
    const int64_t ngroups = (n + 7) / 8;
    if (A64)
        k_pack_bf16<double><<<nblocks(ngroups * n), kBlock, 0, st>>>(A64, n, ngroups, (uint4*)A16);
    else if (A32)
        k_pack_bf16<float><<<nblocks(ngroups * n), kBlock, 0, st>>>(A32, n, ngroups, (uint4*)A16);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

size_t bf16_matrix_elements(int64_t n) { return (size_t)((n + 7) / 8) * 8 * (size_t)n; }

hipError_t launch_bmm16(hipStream_t st, const uint16_t* A16, int64_t n, const void* X, int64_t ldx, int nrhs,
                        double* Y, int64_t ldy, void* ws, size_t ws_bytes, bool x_is_panel, const int* skip)
{
    if (n <= 0 || nrhs <= 0) return hipSuccess;
    if (nrhs > 64 || ws_bytes < bmm_workspace_bytes(n, nrhs, kBmmBf16)) return hipErrorInvalidValue;
    const int nt = nrhs <= 32 ? 1 : 2, kp = 32 * nt;
    const int64_t tiles = (n + 63) / 64, ngroups = (n + 7) / 8;
    const int s = bmm_splits(n, tiles);
    float* part = (float*)ws;
    // the panel after the partial sums of any of the kernels, 256-byte aligned
    const int sw = std::max({s, nt == 2 ? bmm_stream_splits(n) : 1, nt == 1 ? bmm_tile_splits(n, 32, 2) : 1});
    uintptr_t pv = (uintptr_t)(part + (size_t)sw * n * kp);
    pv = (pv + 255) & ~(uintptr_t)255;
    uint4* Vp = (uint4*)pv;
    const bf8v* P = x_is_panel ? (const bf8v*)X : (const bf8v*)Vp;
    if (!x_is_panel)
        k_panel_bf16<<<(unsigned)((ngroups + 1) / 2), kBlock, 0, st>>>((const double*)X, ldx, n, nrhs, kp, ngroups, Vp);
    if (nrhs == 1) {
        // one right-hand side: the vector-ALU product, split over column groups into about 768
        // blocks (3 per CU: 6.0 TB/s on a 20,000-unknown matrix, against 4.5 TB/s at 22,000
        // blocks of short columns); its partials fit the block product's (s16 <= s kp)
        const int64_t rb = (n + kBlock - 1) / kBlock;
        int64_t s16 = (768 + rb - 1) / rb;
        s16 = std::max<int64_t>(1, std::min<int64_t>({s16, ngroups, (int64_t)s * kp}));
        const int64_t g16 = (ngroups + s16 - 1) / s16;
        const int n16 = (int)((ngroups + g16 - 1) / g16);
        k_gemv16_part<<<dim3((unsigned)rb, (unsigned)n16), kBlock, 0, st>>>((const uint4*)A16, n, (const uint4*)P,
                                                                            kp, ngroups, g16, part, skip);
        k_gemv16_reduce<<<nblocks(n), kBlock, 0, st>>>(part, n16, n, Y, skip);
        return hipGetLastError();
    }
    if (nt == 2) {
        const int ss = bmm_stream_splits(n);
        const int64_t gs = (ngroups + ss - 1) / ss;
        const int nss = (int)((ngroups + gs - 1) / gs);
        dim3 grid((unsigned)((n + kStreamRows - 1) / kStreamRows), (unsigned)nss);
        k_bmm16s<2><<<grid, kBlock, 0, st>>>((const bf8v*)A16, n, P, ngroups, gs, part, skip);
        k_bmm32_reduce<64><<<(unsigned)((n + kReduceRows - 1) / kReduceRows), kBlock, 0, st>>>(part, nss, n, nrhs, Y, ldy, skip);
        return hipGetLastError();
    } else {
        // 2-32 right-hand sides: streamed through LDS on one 32-column tile, HBM-bound, split
        // like the 16 x 16-tile products
        const int ss = bmm_tile_splits(n, 32, 2);
        const int64_t gs = (ngroups + ss - 1) / ss;
        const int nss = (int)((ngroups + gs - 1) / gs);
        dim3 grid((unsigned)((n + kStreamRows - 1) / kStreamRows), (unsigned)nss);
        k_bmm16s<1><<<grid, kBlock, 0, st>>>((const bf8v*)A16, n, P, ngroups, gs, part, skip);
        k_bmm32_reduce<32><<<(unsigned)((n + kReduceRows - 1) / kReduceRows), kBlock, 0, st>>>(part, nss, n, nrhs, Y, ldy, skip);
        return hipGetLastError();
    }
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
    k_bupdate<double><<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(V, vstride, h, ldh, cnt, sign, n, Y, nullptr);
    return hipGetLastError();
}

hipError_t launch_bupdate(hipStream_t st, const float* V, int64_t vstride, const double* h, int ldh, const int* cnt,
                          double sign, int64_t n, int nrhs, double* Y)
{
    if (nrhs <= 0) return hipSuccess;
    k_bupdate<float><<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(V, vstride, h, ldh, cnt, sign, n, Y, nullptr);
    return hipGetLastError();
}

hipError_t launch_bnorm(hipStream_t st, const double* W, int64_t n, int nrhs, double* nrm, int stride)
{
    if (nrhs <= 0) return hipSuccess;
    k_bnorm<<<(unsigned)nrhs, kBlock, 0, st>>>(W, n, nrm, stride, nullptr);
    return hipGetLastError();
}

namespace {
template <typename VT>
hipError_t bnormalize_any(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                          int64_t n, int nrhs, VT* V, int64_t vstride, int q, int kp, float* Vt32, double* Vt64,
                          uint16_t* Vt16)
{
    if (nrhs <= 0) return hipSuccess;
    dim3 grid(nblocks(n), (unsigned)nrhs);
    if (Vt16)
        k_bnormalize<uint16_t, VT><<<grid, kBlock, 0, st>>>(W, nrm, stride, active, n, V, vstride, q, kp, Vt16);
    else if (Vt32)
        k_bnormalize<float, VT><<<grid, kBlock, 0, st>>>(W, nrm, stride, active, n, V, vstride, q, kp, Vt32);
    else
        k_bnormalize<double, VT><<<grid, kBlock, 0, st>>>(W, nrm, stride, active, n, V, vstride, q, kp, Vt64);
    return hipGetLastError();
}

// the Gram-Schmidt passes, norm and normalisation of one lockstep step on a basis of type VT
template <typename VT>
void batch_orthonormalize(hipStream_t st, const BatchArnoldi& a, VT* V)
{
    const int64_t n = a.n;
    const int nrhs = a.nrhs;
    const int* stop = a.ints + 1;
    const unsigned nchunks = (unsigned)((n + kProjRows - 1) / kProjRows);
    for (double* h : {a.h1, a.h2}) {
        k_bproj_part<VT><<<dim3(nchunks, (unsigned)nrhs), kBlock, 0, st>>>(V, a.vstride, a.W, n, a.proj, a.ldh,
                                                                            a.ints);
        k_bproj_sum<<<(unsigned)nrhs, 64, 0, st>>>(a.proj, (int)nchunks, a.ldh, h, a.ints);
        k_bupdate<VT><<<dim3(nblocks(n), (unsigned)nrhs), kBlock, 0, st>>>(V, a.vstride, h, a.ldh, a.cnt, -1.0, n,
                                                                          a.W, stop);
    }
    k_bnorm<<<(unsigned)nrhs, kBlock, 0, st>>>(a.W, n, a.nrm, 1, stop);
    const dim3 grid(nblocks(n), (unsigned)nrhs);
    if (a.A16)
        k_bnormalize_idx<uint16_t, VT><<<grid, kBlock, 0, st>>>(a.W, a.nrm, a.active, n, V, a.vstride, a.kp, a.P16,
                                                               a.ints);
    else if (a.use32)
        k_bnormalize_idx<float, VT><<<grid, kBlock, 0, st>>>(a.W, a.nrm, a.active, n, V, a.vstride, a.kp, a.P32,
                                                            a.ints);
    else
        k_bnormalize_idx<double, VT><<<grid, kBlock, 0, st>>>(a.W, a.nrm, a.active, n, V, a.vstride, a.kp, a.P64,
                                                             a.ints);
}
}  // namespace

hipError_t launch_bnormalize(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                             int64_t n, int nrhs, double* V, int64_t vstride, int q, int kp, float* Vt32,
                             double* Vt64, uint16_t* Vt16)
{
    return bnormalize_any(st, W, nrm, stride, active, n, nrhs, V, vstride, q, kp, Vt32, Vt64, Vt16);
}

hipError_t launch_bnormalize(hipStream_t st, const double* W, const double* nrm, int stride, const int* active,
                             int64_t n, int nrhs, float* V, int64_t vstride, int q, int kp, float* Vt32, double* Vt64,
                             uint16_t* Vt16)
{
    return bnormalize_any(st, W, nrm, stride, active, n, nrhs, V, vstride, q, kp, Vt32, Vt64, Vt16);
}

int64_t batch_proj_doubles(int64_t n, int nrhs, int ldh)
{
    return (n + kProjRows - 1) / kProjRows * (int64_t)nrhs * ldh;
}

hipError_t launch_batch_arnoldi_step(hipStream_t st, const BatchArnoldi& a)
{
    const int64_t n = a.n;
    const int nrhs = a.nrhs;
    int* stop = a.ints + 1;
    hipError_t e = a.A16 ? launch_bmm16(st, a.A16, n, a.P16, n, nrhs, a.W, n, a.ws, a.ws_bytes, true, stop)
                         : launch_bmm(st, a.use32 ? a.A32 : nullptr, a.use32 ? nullptr : a.A64, n,
                                      a.use32 ? (const double*)a.P32 : a.P64, n, nrhs, a.W, n, a.ws, a.ws_bytes, true,
                                      stop);
    if (e != hipSuccess) return e;
    if (a.V32)
        batch_orthonormalize(st, a, a.V32);
    else
        batch_orthonormalize(st, a, a.V);
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
