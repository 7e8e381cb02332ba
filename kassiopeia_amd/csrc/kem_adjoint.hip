// kem_adjoint.hip -- derivatives of the BEM path with respect to the triangle vertices (no
// reference counterpart: KEMField has no adjoint).
//
// Two vector-Jacobian products, both under the numeric integrator policy with the quadrature
// regime of every pair frozen at the one the forward pass used:
//   field term   d/dV of  sum_k g_k phi(P_k) = sum_k g_k sum_j G(P_k; T_j) sigma_j  (and d/dsigma)
//   matrix term  d/dV of  sum_ij w_i A_ij u_j  with A_ij = G(c_i; T_j), c_i the centroid of T_i
// G(P; T) is the potential of triangle T carrying unit charge density.  For a cubature pair
// (numeric rule with nodes Q_q = sum_m beta_qm v_m, weights w_q, prefactor K * area):
//   dG/dv_m = K (d area/dv_m) S + K area sum_q w_q beta_qm (P - Q_q)/R_q^3,  S = sum_q w_q / R_q
//   dG/dP   = -K area sum_q w_q (P - Q_q)/R_q^3
// with d area/dv_0 = n x (v_2 - v_1)/2 and cyclic.  A closed-form (near) pair is differentiated
// by forward-mode dual numbers through the closed-form potential
//   G = K (sum_e t_e ln((R_b + l_b)/(R_a + l_a)) + h Omega)
// (t_e: distance of P's projection to edge e, l: positions of the edge ends along it, h:
// height over the plane, Omega: Van Oosterom-Strackee solid angle).  Near pairs run in kernels
// of their own, so the register footprint of the dual numbers stays out of the far kernels.
//
// Layouts: one lane per source element j (its vertices and accumulators in registers), the
// points / targets wave-uniform.  Element gradients are per element and vertex corner:
// dvert[j*9 + 3*m + c] = d/d(v_m)_c, v_0 = P0, v_1 = P0 + A N1, v_2 = P0 + B N2 of its record.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "kem_device.hpp"
#include "kem_kernels.h"

namespace kem {
namespace {

constexpr int kBlock = 256;
constexpr int kChunk = 64;  // targets per LDS reduction round

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------------------ dual numbers
template <int N>
struct Dual {
    double v;
    double d[N];
};

template <int N>
__device__ __forceinline__ Dual<N> cst(double a)
{
    Dual<N> r;
    r.v = a;
    for (int k = 0; k < N; k++) r.d[k] = 0.;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator+(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v + b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator-(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v - b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    r.v = a.v * b.v;
    for (int k = 0; k < N; k++) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator*(double s, const Dual<N>& a)
{
    Dual<N> r;
    r.v = s * a.v;
    for (int k = 0; k < N; k++) r.d[k] = s * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> operator/(const Dual<N>& a, const Dual<N>& b)
{
    Dual<N> r;
    const double ib = 1. / b.v;
    r.v = a.v * ib;
    for (int k = 0; k < N; k++) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dsqrt(const Dual<N>& a)
{
    Dual<N> r;
    r.v = sqrt(a.v);
    const double f = 0.5 / r.v;
    for (int k = 0; k < N; k++) r.d[k] = f * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> dlog(const Dual<N>& a)
{
    Dual<N> r;
    r.v = log(a.v);
    const double f = 1. / a.v;
    for (int k = 0; k < N; k++) r.d[k] = f * a.d[k];
    return r;
}
template <int N>
__device__ __forceinline__ Dual<N> datan2(const Dual<N>& y, const Dual<N>& x)
{
    Dual<N> r;
    r.v = atan2(y.v, x.v);
    const double f = 1. / (x.v * x.v + y.v * y.v);
    for (int k = 0; k < N; k++) r.d[k] = (x.v * y.d[k] - y.v * x.d[k]) * f;
    return r;
}

template <class S>
__device__ __forceinline__ S dot3(const S* a, const S* b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}
template <class S>
__device__ __forceinline__ void cross3(const S* a, const S* b, S* c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}

// Closed-form potential of triangle V (unit charge density) at P.  The edge terms keep their
// t-derivative when P's projection lies on an edge line (t_e = 0, log finite); only a point on
// the line beyond an edge end, where the logarithm diverges, drops the (vanishing) term.
template <int N>
__device__ Dual<N> tri_potential_cf(const Dual<N>* P, const Dual<N> (*V)[3])
{
    Dual<N> e1[3], e2[3], nr[3], n[3], r[3][3], R[3];
    for (int k = 0; k < 3; k++) {
        e1[k] = V[1][k] - V[0][k];
        e2[k] = V[2][k] - V[0][k];
    }
    cross3(e1, e2, nr);
    const Dual<N> nn = dsqrt(dot3(nr, nr));
    for (int k = 0; k < 3; k++) n[k] = nr[k] / nn;
    for (int m = 0; m < 3; m++) {
        for (int k = 0; k < 3; k++) r[m][k] = V[m][k] - P[k];
        R[m] = dsqrt(dot3(r[m], r[m]));
    }
    const Dual<N> h = cst<N>(0.) - dot3(n, r[0]);
    Dual<N> sum = cst<N>(0.);
    for (int e = 0; e < 3; e++) {
        const int a = e, b = (e + 1) % 3;
        Dual<N> ed[3], u[3], mv[3];
        for (int k = 0; k < 3; k++) ed[k] = V[b][k] - V[a][k];
        const Dual<N> L = dsqrt(dot3(ed, ed));
        for (int k = 0; k < 3; k++) u[k] = ed[k] / L;
        cross3(u, n, mv);
        const Dual<N> t = dot3(mv, r[a]);
        const Dual<N> la = dot3(u, r[a]), lb = dot3(u, r[b]);
        // R + l without cancellation: for l < 0, R + l = rho^2 / (R - l) with rho^2 = t^2 + h^2
        // the squared distance of P from the edge line
        const Dual<N> rho2 = t * t + h * h;
        const Dual<N> fa = la.v >= 0. ? R[a] + la : rho2 / (R[a] - la);
        const Dual<N> fb = lb.v >= 0. ? R[b] + lb : rho2 / (R[b] - lb);
        // P on the edge line beyond an end: the term and its limit vanish (t = 0)
        if (!(fa.v > 1e-300) || !(fb.v > 1e-300)) continue;
        sum = sum + t * dlog(fb / fa);
    }
    Dual<N> c12[3];
    cross3(r[1], r[2], c12);
    const Dual<N> trip = dot3(r[0], c12);
    const Dual<N> den = R[0] * R[1] * R[2] + dot3(r[0], r[1]) * R[2] + dot3(r[0], r[2]) * R[1] +
                        dot3(r[1], r[2]) * R[0];
    const Dual<N> om = 2. * datan2(trip, den);
    return dev::kK * (sum + h * om);
}

// ------------------------------------------------------------------ per-element lane state
struct Elem {
    double v[3][3];   // vertices
    double da[3][3];  // K d(area)/dv_m
    double pref;      // K area
    double sig;       // charge density
    bool tri;
};

__device__ __forceinline__ void load_elem(const double* __restrict__ g, double sig, Elem& E)
{
    E.tri = (int)g[KG_TYPE] == KEM_TRIANGLE;
    for (int m = 0; m < 3; m++)
        for (int c = 0; c < 3; c++) E.v[m][c] = g[KG_V + 3 * m + c];
    const double n[3] = {g[KG_NX], g[KG_NY], g[KG_NZ]};
    for (int m = 0; m < 3; m++) {
        // n x (v_{m+2} - v_{m+1}) / 2: the opposite edge, cyclic
        const int p = (m + 1) % 3, q = (m + 2) % 3;
        const double ed[3] = {E.v[q][0] - E.v[p][0], E.v[q][1] - E.v[p][1], E.v[q][2] - E.v[p][2]};
        double c[3];
        cross3(n, ed, c);
        for (int k = 0; k < 3; k++) E.da[m][k] = 0.5 * dev::kK * c[k];
    }
    E.pref = g[KG_PREF];
    E.sig = sig;
}

// cubature derivatives of one far pair: value G, dG/dv (9), dG/dP (3)
__device__ __forceinline__ void far_derivs(const Elem& E, const double* __restrict__ rule, int slot0, int nq,
                                           double px, double py, double pz, double& G, double* dv, double* dP)
{
    double S = 0., F[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};
    for (int q = 0; q < nq; q++) {
        const double* t = rule + 4 * (slot0 + q);
        const double b0 = t[0], b1 = t[1], b2 = t[2], w = t[3];
        const double qx = b0 * E.v[0][0] + b1 * E.v[1][0] + b2 * E.v[2][0];
        const double qy = b0 * E.v[0][1] + b1 * E.v[1][1] + b2 * E.v[2][1];
        const double qz = b0 * E.v[0][2] + b1 * E.v[1][2] + b2 * E.v[2][2];
        const double rx = px - qx, ry = py - qy, rz = pz - qz;
        const double ir = dev::rsqrt_acc(rx * rx + ry * ry + rz * rz);
        S += w * ir;
        const double t3 = w * ir * ir * ir;
        const double bb[3] = {b0, b1, b2};
        for (int m = 0; m < 3; m++) {
            const double s = t3 * bb[m];
            F[m][0] += s * rx;
            F[m][1] += s * ry;
            F[m][2] += s * rz;
        }
    }
    G = E.pref * S;
    for (int m = 0; m < 3; m++)
        for (int c = 0; c < 3; c++) dv[3 * m + c] = E.da[m][c] * S + E.pref * F[m][c];
    for (int c = 0; c < 3; c++) dP[c] = -E.pref * (F[0][c] + F[1][c] + F[2][c]);
}

// regime of a triangle pair as the forward assembly / pair evaluation chooses it (element_eval)
__device__ __forceinline__ int tri_regime(const double* __restrict__ g, double px, double py, double pz)
{
    const double dr = dev::distance_ratio(g, px, py, pz);
    if (dr > dev::kTriCub12) return 7;
    if (dr > dev::kTriCub33) return 12;
    if (dr > dev::kTriRWG) return 33;
    return 0;
}

// regime of a triangle pair as the forward field pass chooses it (squared thresholds)
__device__ __forceinline__ int tri_regime_field(const double* __restrict__ g, double px, double py, double pz)
{
    const double r2 = dev::dist2(g, px, py, pz);
    if (!(r2 > g[KG_NEAR2])) return 0;
    if (r2 > g[KG_REF7_2]) return 7;
    if (r2 > g[KG_REF12_2]) return 12;
    return 33;
}

__device__ __forceinline__ int rule_slot(int nq) { return nq == 7 ? KEM_N7 : nq == 12 ? KEM_N12 : KEM_N33; }

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// ------------------------------------------------------------------ field term
// partV[s][j][10] = (sum_k g_k G_kj, sum_k g_k sigma_j dG_kj/dv) over the points of split s
template <bool NEAR>
__global__ __launch_bounds__(kBlock) void k_vjp_field(const double* __restrict__ geom,
                                                      const double* __restrict__ sigma,
                                                      const double* __restrict__ rule, int64_t n,
                                                      const double* __restrict__ pts, const double* __restrict__ gk,
                                                      int64_t npts, int64_t per_split, double* __restrict__ partV)
{
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    const double* g = geom + j * KG_STRIDE;
    Elem E;
    load_elem(g, sigma[j], E);
    const int64_t k0 = (int64_t)blockIdx.y * per_split;
    const int64_t k1 = std::min<int64_t>(npts, k0 + per_split);
    double acc[10];
    double* out = partV + ((int64_t)blockIdx.y * n + j) * 10;
    for (int c = 0; c < 10; c++) acc[c] = NEAR ? out[c] : 0.;
    if (E.tri) {
        for (int64_t k = k0; k < k1; k++) {
            const double px = pts[3 * k], py = pts[3 * k + 1], pz = pts[3 * k + 2], wk = gk[k];
            const int reg = tri_regime_field(g, px, py, pz);
            if (NEAR) {
                if (reg != 0) continue;
                Dual<9> V[3][3], P[3];
                for (int m = 0; m < 3; m++)
                    for (int c = 0; c < 3; c++) {
                        V[m][c] = cst<9>(E.v[m][c]);
                        V[m][c].d[3 * m + c] = 1.;
                    }
                P[0] = cst<9>(px);
                P[1] = cst<9>(py);
                P[2] = cst<9>(pz);
                const Dual<9> G = tri_potential_cf<9>(P, V);
                acc[0] += wk * G.v;
                const double ws = wk * E.sig;
                for (int c = 0; c < 9; c++) acc[1 + c] += ws * G.d[c];
            } else {
                if (reg == 0) continue;
                double G, dv[9], dP[3];
                far_derivs(E, rule, rule_slot(reg), reg, px, py, pz, G, dv, dP);
                acc[0] += wk * G;
                const double ws = wk * E.sig;
                for (int c = 0; c < 9; c++) acc[1 + c] += ws * dv[c];
            }
        }
    }
    for (int c = 0; c < 10; c++) out[c] = acc[c];
}

// ------------------------------------------------------------------ matrix term
// lanes: sources j; targets i of split blockIdx.y, wave-uniform.  Weight w_ij = lw_i * u_j.
// partV[s][j][1..9] += sum_i w_ij dA_ij/dv_j;  partP[bx][i][3] (+)= sum_{j in block} w_ij dA_ij/dc_i
template <bool NEAR>
__global__ __launch_bounds__(kBlock) void k_vjp_matrix(const double* __restrict__ geom,
                                                       const double* __restrict__ targets,
                                                       const double* __restrict__ rule, int64_t n,
                                                       const double* __restrict__ lw, const double* __restrict__ u,
                                                       int64_t per_split, double* __restrict__ partV,
                                                       double* __restrict__ partP)
{
    __shared__ double sh[kBlock / 64][kChunk][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int64_t jr = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const bool live = jr < n;
    const int64_t j = live ? jr : n - 1;
    const double* g = geom + j * KG_STRIDE;
    Elem E;
    load_elem(g, live ? u[j] : 0., E);
    const bool act = live && E.tri;
    const int64_t i0 = (int64_t)blockIdx.y * per_split;
    const int64_t i1 = std::min<int64_t>(n, i0 + per_split);
    double acc[9];
    double* outV = partV + ((int64_t)blockIdx.y * n + j) * 10;
    for (int c = 0; c < 9; c++) acc[c] = (NEAR && live) ? outV[1 + c] : 0.;
    for (int64_t ib = i0; ib < i1; ib += kChunk) {
        const int cnt = (int)std::min<int64_t>(kChunk, i1 - ib);
        for (int ii = 0; ii < cnt; ii++) {
            const int64_t i = ib + ii;
            const double* t = targets + i * KT_STRIDE;
            const double px = t[KT_CX], py = t[KT_CY], pz = t[KT_CZ];
            const double w = lw[i] * E.sig;
            double gp[3] = {0., 0., 0.};
            const int reg = act ? tri_regime(g, px, py, pz) : -1;
            if (NEAR) {
                if (reg == 0) {
                    Dual<12> V[3][3], P[3];
                    for (int m = 0; m < 3; m++)
                        for (int c = 0; c < 3; c++) {
                            V[m][c] = cst<12>(E.v[m][c]);
                            V[m][c].d[3 * m + c] = 1.;
                        }
                    const double pc[3] = {px, py, pz};
                    for (int c = 0; c < 3; c++) {
                        P[c] = cst<12>(pc[c]);
                        P[c].d[9 + c] = 1.;
                    }
                    const Dual<12> G = tri_potential_cf<12>(P, V);
                    for (int c = 0; c < 9; c++) acc[c] += w * G.d[c];
                    for (int c = 0; c < 3; c++) gp[c] = w * G.d[9 + c];
                }
            } else if (reg > 0) {
                double G, dv[9], dP[3];
                far_derivs(E, rule, rule_slot(reg), reg, px, py, pz, G, dv, dP);
                for (int c = 0; c < 9; c++) acc[c] += w * dv[c];
                for (int c = 0; c < 3; c++) gp[c] = w * dP[c];
            }
            for (int c = 0; c < 3; c++) {
                const double s = wave_sum(gp[c]);
                if (lane == 0) sh[wave][ii][c] = s;
            }
        }
        __syncthreads();
        for (int e = threadIdx.x; e < cnt * 3; e += kBlock) {
            const int ii = e / 3, c = e % 3;
            const double s = ((sh[0][ii][c] + sh[1][ii][c]) + sh[2][ii][c]) + sh[3][ii][c];
            double* dst = partP + ((int64_t)blockIdx.x * n + ib + ii) * 3 + c;
            *dst = NEAR ? *dst + s : s;
        }
        __syncthreads();
    }
    if (live)
        for (int c = 0; c < 9; c++) outV[1 + c] = acc[c];
}

// dsig[j] = sum_s partV[s][j][0];  dvert[j*9 + c] = sum_s partV[s][j][1 + c]
__global__ __launch_bounds__(kBlock) void k_vjp_reduce_v(const double* __restrict__ partV, int nsplit, int64_t n,
                                                         double* __restrict__ dsig, double* __restrict__ dvert)
{
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= n) return;
    double s[10];
    for (int c = 0; c < 10; c++) s[c] = 0.;
    for (int k = 0; k < nsplit; k++)
        for (int c = 0; c < 10; c++) s[c] += partV[((int64_t)k * n + j) * 10 + c];
    if (dsig) dsig[j] = s[0];
    for (int c = 0; c < 9; c++) dvert[j * 9 + c] = s[1 + c];
}

// dvert[i*9 + 3m + c] += (sum_bx partP[bx][i][c]) / 3 for each corner m (c_i = mean of the vertices)
__global__ __launch_bounds__(kBlock) void k_vjp_reduce_p(const double* __restrict__ partP, int nbx, int64_t n,
                                                         double* __restrict__ dvert)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    double s[3] = {0., 0., 0.};
    for (int b = 0; b < nbx; b++)
        for (int c = 0; c < 3; c++) s[c] += partP[((int64_t)b * n + i) * 3 + c];
    for (int m = 0; m < 3; m++)
        for (int c = 0; c < 3; c++) dvert[i * 9 + 3 * m + c] += s[c] * (1. / 3.);
}

int vjp_splits(int64_t nbx, int64_t work)
{
    int64_t s = (2048 + nbx - 1) / std::max<int64_t>(nbx, 1);
    s = std::min<int64_t>(s, std::max<int64_t>(1, work / 16));
    return (int)std::max<int64_t>(1, std::min<int64_t>(s, 256));
}

}  // namespace

size_t vjp_workspace_doubles(int64_t n, int64_t npts)
{
    const int64_t nbx = (n + kBlock - 1) / kBlock;
    const int sf = vjp_splits(nbx, npts), sm = vjp_splits(nbx, n);
    return (size_t)std::max<int64_t>(sf, sm) * n * 10 + (size_t)nbx * n * 3;
}

hipError_t launch_field_vjp(hipStream_t st, const double* geom, const double* sigma, const double* rule, int64_t n,
                            const double* pts, const double* gk, int64_t npts, double* dsig, double* dvert,
                            double* ws, size_t ws_doubles)
{
    if (n <= 0) return hipSuccess;
    if (ws_doubles < vjp_workspace_doubles(n, npts)) return hipErrorInvalidValue;
    const int64_t nbx = nblocks(n);
    const int s = npts > 0 ? vjp_splits(nbx, npts) : 1;
    const int64_t per = npts > 0 ? (npts + s - 1) / s : 1;
    const int ns = npts > 0 ? (int)((npts + per - 1) / per) : 1;
    dim3 grid((unsigned)nbx, (unsigned)ns);
    k_vjp_field<false><<<grid, kBlock, 0, st>>>(geom, sigma, rule, n, pts, gk, npts, per, ws);
    k_vjp_field<true><<<grid, kBlock, 0, st>>>(geom, sigma, rule, n, pts, gk, npts, per, ws);
    k_vjp_reduce_v<<<(unsigned)nbx, kBlock, 0, st>>>(ws, ns, n, dsig, dvert);
    return hipGetLastError();
}

hipError_t launch_matrix_vjp(hipStream_t st, const double* geom, const double* targets, const double* rule,
                             int64_t n, const double* lw, const double* u, double* dvert, double* ws,
                             size_t ws_doubles)
{
    if (n <= 0) return hipSuccess;
    if (ws_doubles < vjp_workspace_doubles(n, 0)) return hipErrorInvalidValue;
    const int64_t nbx = nblocks(n);
    const int s = vjp_splits(nbx, n);
    const int64_t per = (n + s - 1) / s;
    const int ns = (int)((n + per - 1) / per);
    double* partV = ws;
    double* partP = ws + (size_t)s * n * 10;
    dim3 grid((unsigned)nbx, (unsigned)ns);
    k_vjp_matrix<false><<<grid, kBlock, 0, st>>>(geom, targets, rule, n, lw, u, per, partV, partP);
    k_vjp_matrix<true><<<grid, kBlock, 0, st>>>(geom, targets, rule, n, lw, u, per, partV, partP);
    k_vjp_reduce_v<<<(unsigned)nbx, kBlock, 0, st>>>(partV, ns, n, nullptr, dvert);
    k_vjp_reduce_p<<<(unsigned)nbx, kBlock, 0, st>>>(partP, (int)nbx, n, dvert);
    return hipGetLastError();
}

}  // namespace kem
