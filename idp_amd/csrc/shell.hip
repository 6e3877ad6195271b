// Discrete-shell terms of the step on the device: the flow variant's
// -h/2 x^T L x smoothing term over the shell triangles
// (Library/FEM/Shell/INC_POTENTIAL.h:53-73, 191-211, 323-339) and the
// normal-flow body force (Library/FEM/Shell/DISCRETE_SHELL.h:359-395).
//
// Shell triangles are elements after the tets: triangle e is element nT + e
// and owns the local slots of an element, 12 gradient doubles at
// gl[12 (nT + e)] and 16 vertex blocks at Hl[144 (nT + e)] of which it writes
// the 3 x 3 leading ones, so the static BSR pattern and the gathers treat
// tets and triangles alike.
#include <hip/hip_runtime.h>

#include "geom_dev.hpp"
#include "kernels.hpp"
#include "proj_dev.hpp"

namespace idp {

// per triangle: sum_i sum_d (-h/2) x_i[d] (vol/6 (x_{i+1}[d] + x_{i+2}[d] - 2 x_i[d])),
// the reference's per-vertex terms x . Lx split by triangle; partials per block
__global__ void __launch_bounds__(RB) k_flow_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                    const int* __restrict__ tris, const double* __restrict__ vol,
                                                    const double* __restrict__ X, double h, double* __restrict__ part)
{
    __shared__ double sh[RB];
    const int e = lo[blockIdx.x] + threadIdx.x;
    double acc = 0.0;
    if (e < hi[blockIdx.x]) {
        const int t[3] = {tris[3 * e], tris[3 * e + 1], tris[3 * e + 2]};
        const double w = vol[e] / 6;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double* a = X + 3 * t[i];
            const double* b = X + 3 * t[(i + 1) % 3];
            const double* c = X + 3 * t[(i + 2) % 3];
#pragma unroll
            for (int d = 0; d < 3; ++d) acc += -h * 0.5 * a[d] * (w * ((b[d] + c[d]) - 2 * a[d]));
        }
    }
    const double s = block_sum<RB>(acc, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// gl[12 (nT + e) + 3 i + d] = -h vol/6 (x_{i+1} + x_{i+2} - 2 x_i)[d]
__global__ void k_flow_grad(int nE, int nT, const int* __restrict__ tris, const double* __restrict__ vol,
                            const double* __restrict__ X, double h, double* __restrict__ gl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    const int t[3] = {tris[3 * e], tris[3 * e + 1], tris[3 * e + 2]};
    const double w = vol[e] / 6;
    double* o = gl + 12 * (size_t)(nT + e);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        const double* a = X + 3 * t[i];
        const double* b = X + 3 * t[(i + 1) % 3];
        const double* c = X + 3 * t[(i + 2) % 3];
#pragma unroll
        for (int d = 0; d < 3; ++d) o[3 * i + d] = -h * (w * ((b[d] + c[d]) - 2 * a[d]));
    }
}

// blocks (a, b) of triangle e at Hl[144 (nT + e) + 9 (4 a + b)]: 2 h vol/6 I
// on the diagonal, -h vol/6 I off it (constant: the term is quadratic)
__global__ void k_flow_hess(int nE, int nT, const double* __restrict__ vol, double h, double* __restrict__ Hl)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;  // one lane per (triangle, block)
    if (k >= 9 * nE) return;
    const int e = k / 9, ab = k % 9, a = ab / 3, b = ab % 3;
    const double s = a == b ? 2 * h * vol[e] / 6 : -h * vol[e] / 6;
    double* o = Hl + 144 * (size_t)(nT + e) + 9 * (4 * a + b);
#pragma unroll
    for (int q = 0; q < 9; ++q) o[q] = (q % 4 == 0) ? s : 0.0;
}

// Update_Normal_Flow_Neumann per vertex: n = sum over the vertex's triangles,
// in triangle order, of cross(X2 - X1, X3 - X1); b = magnitude m n / |n|.
// Evaluated as written (no contraction) so it equals the sequential loop.
__global__ void k_normal_flow(int nV, const int* __restrict__ vt_ptr, const int* __restrict__ vt,
                              const int* __restrict__ tris, const double* __restrict__ X,
                              const double* __restrict__ mass, double magnitude, double* __restrict__ b)
{
#pragma clang fp contract(off)
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double n0 = 0.0, n1 = 0.0, n2 = 0.0;
    for (int j = vt_ptr[v]; j < vt_ptr[v + 1]; ++j) {
        const int* t = tris + 3 * vt[j];
        const double* p = X + 3 * t[0];
        const double* q = X + 3 * t[1];
        const double* r = X + 3 * t[2];
        const double u0 = q[0] - p[0], u1 = q[1] - p[1], u2 = q[2] - p[2];
        const double w0 = r[0] - p[0], w1 = r[1] - p[1], w2 = r[2] - p[2];
        n0 += u1 * w2 - u2 * w1;
        n1 += u2 * w0 - u0 * w2;
        n2 += u0 * w1 - u1 * w0;
    }
    const double sq = n0 * n0 + n1 * n1 + n2 * n2;
    const double w = magnitude * mass[v] / sqrt(sq);
    b[3 * v] = n0 * w;
    b[3 * v + 1] = n1 * w;
    b[3 * v + 2] = n2 * w;
}

// ---------------------------------------------------------------------------
// Non-flow shell terms: the neo-Hookean membrane of each shell triangle
// (Shell/MEMBRANE.h:8-312) and the discrete hinge bending of each interior
// edge (Shell/BENDING.h:9-80, 150-214, 440-497; the dihedral angle of
// Math/DIHEDRAL_ANGLE.h:10-24 and its gradient in the closed form of Grinspun
// et al. 2003 / Tamstorf and Grinspun 2013).  A hinge is element
// nT + nE + k (four vertices, like a tet).  Elements whose vertices are all
// Dirichlet nodes contribute nothing.  Hessians are projected in reduced
// coordinates as the barrier's are: membrane y = (x1 - x0, x2 - x0) (order 6,
// one per lane), hinge y = (x0 - x1, x2 - x1, x3 - x1) (order 9, lane
// triples).
// ---------------------------------------------------------------------------
namespace {

struct MemTerms {
    double tr, lnJ, detA, A[4], IA[4], IB[4];  // 2 x 2 row-major
    bool ok;
};

// the shared quantities of MEMBRANE.h for edges e01, e02 and the rest form
// (a, b, d) = (|E01|^2, E01.E02, |E02|^2)
__device__ inline MemTerms mem_terms(d3 e01, d3 e02, const double* rest)
{
    MemTerms T;
    const double a = rest[0], b = rest[1], d = rest[2];
    const double det = a * d - b * b;
    T.ok = det != 0.0;  // MEMBRANE.h:31-32
    if (!T.ok) return T;
    T.IB[0] = d / det;
    T.IB[1] = T.IB[2] = -b / det;
    T.IB[3] = a / det;
    T.A[0] = sqn(e01);
    T.A[1] = T.A[2] = dot(e01, e02);
    T.A[3] = sqn(e02);
    T.detA = T.A[0] * T.A[3] - T.A[1] * T.A[2];
    T.IA[0] = T.A[3] / T.detA;
    T.IA[1] = T.IA[2] = -T.A[1] / T.detA;
    T.IA[3] = T.A[0] / T.detA;
    const double detIB = T.IB[0] * T.IB[3] - T.IB[1] * T.IB[2];
    T.lnJ = log(T.detA * detIB) / 2;
    T.tr = (T.IB[0] * T.A[0] + T.IB[1] * T.A[2]) + (T.IB[2] * T.A[1] + T.IB[3] * T.A[3]);
    return T;
}

__device__ inline bool all_dbc3(const unsigned char* dbc, int a, int b, int c) { return dbc[a] && dbc[b] && dbc[c]; }

// dihedral angle of (x0, x1, x2, x3) around x1 - x2 (DIHEDRAL_ANGLE.h:10-24)
__device__ inline double dihedral(d3 v0, d3 v1, d3 v2, d3 v3)
{
    const d3 n1 = cross(v1 - v0, v2 - v0), n2 = cross(v2 - v3, v1 - v3);
    const double c = dot(n1, n2) / sqrt(sqn(n1) * sqn(n2));
    double da = acos(fmax(-1.0, fmin(1.0, c)));
    if (dot(cross(n2, n1), v1 - v2) < 0) da = -da;
    return da;
}

// first-order dual number (value, derivative along one direction)
struct Dual {
    double v, d;
};
__device__ inline Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ inline Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ inline Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ inline Dual operator/(Dual a, Dual b) { return {a.v / b.v, (a.d * b.v - a.v * b.d) / (b.v * b.v)}; }
__device__ inline Dual dsqrt(Dual a)
{
    const double r = sqrt(a.v);
    return {r, a.d / (2.0 * r)};
}
__device__ inline double dval(double a) { return a; }
__device__ inline double dval(Dual a) { return a.v; }
__device__ inline double dsqrt(double a) { return sqrt(a); }

template <class S>
struct V3S {
    S x, y, z;
};
template <class S>
__device__ inline V3S<S> vsub(V3S<S> a, V3S<S> b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class S>
__device__ inline S vdot(V3S<S> a, V3S<S> b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
template <class S>
__device__ inline V3S<S> vcross(V3S<S> a, V3S<S> b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <class S>
__device__ inline V3S<S> vscale(S s, V3S<S> a) { return {s * a.x, s * a.y, s * a.z}; }

// gradient of the dihedral angle (order x0, x1, x2, x3; the shared edge
// e0 = x2 - x1, n1 = e0 x (x0 - x1), n2 = (x3 - x1) x e0)
template <class S>
__device__ inline void dihedral_grad(V3S<S> x0, V3S<S> x1, V3S<S> x2, V3S<S> x3, V3S<S> (&g)[4])
{
    const V3S<S> e0 = vsub(x2, x1), e1 = vsub(x0, x1), e2 = vsub(x3, x1), e3 = vsub(x0, x2), e4 = vsub(x3, x2);
    const V3S<S> n1 = vcross(e0, e1), n2 = vcross(e2, e0);
    const S n1s = vdot(n1, n1), n2s = vdot(n2, n2), e0n = dsqrt(vdot(e0, e0));
    S zero = e0n - e0n;
    g[0] = vscale(zero - e0n / n1s, n1);
    {
        const S a = vdot(e0, e3) / (e0n * n1s), b = vdot(e0, e4) / (e0n * n2s);
        g[1] = vsub(vscale(zero - a, n1), vscale(b, n2));
    }
    {
        const S a = vdot(e0, e1) / (e0n * n1s), b = vdot(e0, e2) / (e0n * n2s);
        V3S<S> t = vscale(a, n1), u = vscale(b, n2);
        g[2] = {t.x + u.x, t.y + u.y, t.z + u.z};
    }
    g[3] = vscale(zero - e0n / n2s, n2);
}

__device__ inline V3S<double> vs(d3 a) { return {a.x, a.y, a.z}; }

}  // namespace

__global__ void __launch_bounds__(RB) k_membrane_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                        const int* __restrict__ tris, const double* __restrict__ tIB,
                                                        const double* __restrict__ vol, const double* __restrict__ lam,
                                                        const double* __restrict__ mu,
                                                        const unsigned char* __restrict__ dbc,
                                                        const double* __restrict__ X, double w,
                                                        double* __restrict__ part)
{
    __shared__ double sh[RB];
    const int e = lo[blockIdx.x] + threadIdx.x;
    double E = 0.0;
    if (e < hi[blockIdx.x]) {
        const int t0 = tris[3 * e], t1 = tris[3 * e + 1], t2 = tris[3 * e + 2];
        if (!all_dbc3(dbc, t0, t1, t2)) {
            const d3 x0 = ld3(X, t0);
            const MemTerms T = mem_terms(ld3(X, t1) - x0, ld3(X, t2) - x0, tIB + 3 * e);
            if (T.ok) E = w * vol[e] * (mu[e] / 2 * (T.tr - 2 - 2 * T.lnJ) + lam[e] / 2 * T.lnJ * T.lnJ);
        }
    }
    const double s = block_sum<RB>(E, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// gl[12 (nT + e) + 3 a + d]: the MEMBRANE.h:57-121 gradient through
// temp = w vol (mu/2 IB^-1 + (lam lnJ - mu)/2 A^-1)
__global__ void k_membrane_grad(int nE, int nT, const int* __restrict__ tris, const double* __restrict__ tIB,
                                const double* __restrict__ vol, const double* __restrict__ lam,
                                const double* __restrict__ mu, const unsigned char* __restrict__ dbc,
                                const double* __restrict__ X, double w, double* __restrict__ gl)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    double* o = gl + 12 * (size_t)(nT + e);
    const int t0 = tris[3 * e], t1 = tris[3 * e + 1], t2 = tris[3 * e + 2];
    const d3 x0 = ld3(X, t0), e01 = ld3(X, t1) - x0, e02 = ld3(X, t2) - x0;
    const MemTerms T = mem_terms(e01, e02, tIB + 3 * e);
    if (all_dbc3(dbc, t0, t1, t2) || !T.ok) {
        for (int k = 0; k < 9; ++k) o[k] = 0.0;
        return;
    }
    const double c = w * vol[e], t1c = (-mu[e] + lam[e] * T.lnJ) / 2;
    double tp[4];
    for (int k = 0; k < 4; ++k) tp[k] = c * (mu[e] / 2 * T.IB[k] + t1c * T.IA[k]);
    const double off = tp[1] + tp[2];
    const d3 g1 = 2.0 * tp[0] * e01 + off * e02, g2 = off * e01 + 2.0 * tp[3] * e02;
    const d3 g0 = -1.0 * (g1 + g2);
    const d3 gg[3] = {g0, g1, g2};
    for (int a = 0; a < 3; ++a) {
        o[3 * a] = gg[a].x;
        o[3 * a + 1] = gg[a].y;
        o[3 * a + 2] = gg[a].z;
    }
}

// projected membrane Hessian, one triangle per lane, into Hl[144 (nT + e)]
// blocks (a, b) at 9 (4 a + b): H_y of MEMBRANE.h:231-262 in y = (e01, e02),
// K = C H_y C, cyclic Jacobi, clamp, C^-1 K+ C^-1, B^T H_y B
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_membrane_hess(int nE, int nT, const int* __restrict__ tris, const double* __restrict__ tIB,
                const double* __restrict__ vol, const double* __restrict__ lam, const double* __restrict__ mu,
                const unsigned char* __restrict__ dbc, const double* __restrict__ X, double w, int project,
                ReducedRoots R, double* __restrict__ Hl)
{
#pragma clang fp contract(fast)
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nE) return;
    double* o = Hl + 144 * (size_t)(nT + e);
    const int t0 = tris[3 * e], t1 = tris[3 * e + 1], t2 = tris[3 * e + 2];
    const d3 x0 = ld3(X, t0), y0 = ld3(X, t1) - x0, y1 = ld3(X, t2) - x0;
    const MemTerms T = mem_terms(y0, y1, tIB + 3 * e);
    if (all_dbc3(dbc, t0, t1, t2) || !T.ok) {
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b)
                for (int q = 0; q < 9; ++q) o[9 * (4 * a + b) + q] = 0.0;
        return;
    }
    const double yv[6] = {y0.x, y0.y, y0.z, y1.x, y1.y, y1.z};
    // dA rows over y (A00, A10, A01, A11)
    double dA[4][6];
    for (int k = 0; k < 3; ++k) {
        dA[0][k] = 2 * yv[k];
        dA[0][3 + k] = 0.0;
        dA[1][k] = dA[2][k] = yv[3 + k];
        dA[1][3 + k] = dA[2][3 + k] = yv[k];
        dA[3][k] = 0.0;
        dA[3][3 + k] = 2 * yv[3 + k];
    }
    double ainv[6];
    for (int i = 0; i < 6; ++i)
        ainv[i] = dA[0][i] * T.IA[0] + dA[1][i] * T.IA[2] + dA[2][i] * T.IA[1] + dA[3][i] * T.IA[3];
    const double term1 = (-mu[e] + lam[e] * T.lnJ) / 2;
    const double c1 = -term1 + lam[e] / 4, c2 = term1 / T.detA, sc = w * vol[e];
    double coef[4];  // ahess weights, (i, j) -> i + 2 j
    for (int k = 0; k < 4; ++k) coef[k] = term1 * T.IA[(k & 1) * 2 + (k >> 1)] + mu[e] / 2 * T.IB[(k & 1) * 2 + (k >> 1)];
    double H[21];
#pragma unroll
    for (int i = 0; i < 6; ++i)
#pragma unroll
        for (int j = i; j < 6; ++j) {
            double v = c1 * ainv[i] * ainv[j];
            v += c2 * ((dA[3][i] * dA[0][j] + dA[0][i] * dA[3][j]) - dA[1][i] * dA[1][j] - dA[2][i] * dA[2][j]);
            const int bi = i / 3, bj = j / 3;
            if (i % 3 == j % 3) {  // second derivatives of A: 2I (y0,y0), I (y0,y1), I (y1,y0), 2I (y1,y1)
                if (bi == 0 && bj == 0) v += 2 * coef[0];
                else if (bi == 1 && bj == 1) v += 2 * coef[3];
                else v += coef[1] + coef[2];
            }
            H[sx<6>(i, j)] = sc * v;
        }
    double Hp[21];
#pragma unroll
    for (int q = 0; q < 21; ++q) Hp[q] = H[q];
    if (project) {
        double K[21], V[36];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = a; b < 2; ++b)
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int j = 0; j < 3; ++j) {
                        if (a == b && j < i) continue;
                        double acc = 0.0;
#pragma unroll
                        for (int g = 0; g < 2; ++g)
#pragma unroll
                            for (int hh = 0; hh < 2; ++hh) acc += R.S2[a * 2 + g] * H[sx<6>(3 * g + i, 3 * hh + j)] * R.S2[hh * 2 + b];
                        K[sx<6>(3 * a + i, 3 * b + j)] = acc;
                    }
        lane_jacobi<6>(K, V);
        double lmin = K[0];
#pragma unroll
        for (int i = 1; i < 6; ++i) lmin = fmin(lmin, K[sx<6>(i, i)]);
        if (lmin < 0.0) {
            double L[6], Kp[21];
#pragma unroll
            for (int i = 0; i < 6; ++i) L[i] = fmax(K[sx<6>(i, i)], 0.0);
#pragma unroll
            for (int i = 0; i < 6; ++i)
#pragma unroll
                for (int j = i; j < 6; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < 6; ++k) acc += V[6 * i + k] * L[k] * V[6 * j + k];
                    Kp[sx<6>(i, j)] = acc;
                }
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = a; b < 2; ++b)
#pragma unroll
                    for (int i = 0; i < 3; ++i)
#pragma unroll
                        for (int j = 0; j < 3; ++j) {
                            if (a == b && j < i) continue;
                            double acc = 0.0;
#pragma unroll
                            for (int g = 0; g < 2; ++g)
#pragma unroll
                                for (int hh = 0; hh < 2; ++hh)
                                    acc += R.Si2[a * 2 + g] * Kp[sx<6>(3 * g + i, 3 * hh + j)] * R.Si2[hh * 2 + b];
                            Hp[sx<6>(3 * a + i, 3 * b + j)] = acc;
                        }
        }
    }
    // B = [[-1, 1, 0], [-1, 0, 1]]: block (a, b) = sum_rs B[r][a] B[s][b] Hp[(r, s)]
    constexpr double B[2][3] = {{-1, 1, 0}, {-1, 0, 1}};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int r = 0; r < 2; ++r)
#pragma unroll
                        for (int s2 = 0; s2 < 2; ++s2)
                            if (B[r][a] * B[s2][b] != 0.0) acc += B[r][a] * B[s2][b] * Hp[sx<6>(3 * r + i, 3 * s2 + j)];
                    o[9 * (4 * a + b) + 3 * i + j] = acc;
                }
}

__global__ void __launch_bounds__(RB) k_hinge_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                     const int4* __restrict__ hinge, const double* __restrict__ rest,
                                                     const unsigned char* __restrict__ dbc,
                                                     const double* __restrict__ X, double wk,
                                                     double* __restrict__ part)
{
    __shared__ double sh[RB];
    const int k = lo[blockIdx.x] + threadIdx.x;
    double E = 0.0;
    if (k < hi[blockIdx.x]) {
        const int4 s = hinge[k];
        if (!(dbc[s.x] && dbc[s.y] && dbc[s.z] && dbc[s.w])) {
            const double* r = rest + 3 * k;
            const double th = dihedral(ld3(X, s.x), ld3(X, s.y), ld3(X, s.z), ld3(X, s.w));
            E = wk * (th - r[0]) * (th - r[0]) * r[1] / r[2];
        }
    }
    const double s = block_sum<RB>(E, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// gl[12 (base + k) + 3 a + d] = w k 2 (theta - theta_rest) |e| / h grad theta
__global__ void k_hinge_grad(int nH, int base, const int4* __restrict__ hinge, const double* __restrict__ rest,
                             const unsigned char* __restrict__ dbc, const double* __restrict__ X, double wk,
                             double* __restrict__ gl)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nH) return;
    double* o = gl + 12 * (size_t)(base + k);
    const int4 s = hinge[k];
    if (dbc[s.x] && dbc[s.y] && dbc[s.z] && dbc[s.w]) {
        for (int q = 0; q < 12; ++q) o[q] = 0.0;
        return;
    }
    const d3 x0 = ld3(X, s.x), x1 = ld3(X, s.y), x2 = ld3(X, s.z), x3 = ld3(X, s.w);
    const double* r = rest + 3 * k;
    const double th = dihedral(x0, x1, x2, x3);
    V3S<double> g[4];
    dihedral_grad<double>(vs(x0), vs(x1), vs(x2), vs(x3), g);
    const double c = wk * 2 * (th - r[0]) * r[1] / r[2];
    for (int a = 0; a < 4; ++a) {
        o[3 * a] = c * g[a].x;
        o[3 * a + 1] = c * g[a].y;
        o[3 * a + 2] = c * g[a].z;
    }
}

// stage 1 of the hinge Hessian, one hinge per lane: H_y (y = (x0 - x1,
// x2 - x1, x3 - x1)) = c1 Hess(theta) + c2 grad grad^T with Hess(theta) from
// directional derivatives of the closed-form gradient, then K = C H_y C;
// record F + 90 k as the barrier's form stage writes it (packed K, packed
// H_y); ok[k] = 0 for a skipped hinge
__global__ void __launch_bounds__(256) k_hinge_form(int nH, const int4* __restrict__ hinge,
                                                    const double* __restrict__ rest,
                                                    const unsigned char* __restrict__ dbc,
                                                    const double* __restrict__ X, double wk, ReducedRoots R,
                                                    double* __restrict__ F, unsigned char* __restrict__ ok)
{
#pragma clang fp contract(fast)
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nH) return;
    const int4 s = hinge[k];
    if (dbc[s.x] && dbc[s.y] && dbc[s.z] && dbc[s.w]) {
        ok[k] = 0;
        return;
    }
    const d3 x0 = ld3(X, s.x), x1 = ld3(X, s.y), x2 = ld3(X, s.z), x3 = ld3(X, s.w);
    const double* r = rest + 3 * k;
    const double th = dihedral(x0, x1, x2, x3);
    const double c1 = wk * 2 * (th - r[0]) * r[1] / r[2], c2 = wk * 2 * r[1] / r[2];
    const d3 y[3] = {x0 - x1, x2 - x1, x3 - x1};
    // gradient in y: (g_x0, g_x2, g_x3)
    double gy[9];
    {
        V3S<double> g[4];
        const V3S<double> o = {0.0, 0.0, 0.0};
        dihedral_grad<double>(vs(y[0]), o, vs(y[1]), vs(y[2]), g);
        const V3S<double> gg[3] = {g[0], g[2], g[3]};
        for (int a = 0; a < 3; ++a) {
            gy[3 * a] = gg[a].x;
            gy[3 * a + 1] = gg[a].y;
            gy[3 * a + 2] = gg[a].z;
        }
    }
    double H[45];
    for (int q = 0; q < 45; ++q) H[q] = 0.0;
    for (int dir = 0; dir < 9; ++dir) {  // column dir of Hess(theta)
        V3S<Dual> yd[3];
        for (int a = 0; a < 3; ++a) {
            yd[a].x = {y[a].x, 3 * a == dir ? 1.0 : 0.0};
            yd[a].y = {y[a].y, 3 * a + 1 == dir ? 1.0 : 0.0};
            yd[a].z = {y[a].z, 3 * a + 2 == dir ? 1.0 : 0.0};
        }
        const Dual z = {0.0, 0.0};
        V3S<Dual> g[4];
        dihedral_grad<Dual>(yd[0], {z, z, z}, yd[1], yd[2], g);
        const V3S<Dual> gg[3] = {g[0], g[2], g[3]};
        double col[9];
        for (int a = 0; a < 3; ++a) {
            col[3 * a] = gg[a].x.d;
            col[3 * a + 1] = gg[a].y.d;
            col[3 * a + 2] = gg[a].z.d;
        }
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            // symmetric: (i, dir) and (dir, i) each get half
            if (i <= dir) H[sx<9>(i, dir)] += 0.5 * c1 * col[i];
            if (i >= dir) H[sx<9>(dir, i)] += 0.5 * c1 * col[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = i; j < 9; ++j) H[sx<9>(i, j)] += c2 * gy[i] * gy[j];
    double* Fi = F + 90 * (size_t)k;
#pragma unroll
    for (int q = 0; q < 45; ++q) Fi[45 + q] = H[q];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = a; b < 3; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if (a == b && j < i) continue;
                    double acc = 0.0;
#pragma unroll
                    for (int g = 0; g < 3; ++g)
#pragma unroll
                        for (int hh = 0; hh < 3; ++hh) acc += R.S3[a * 3 + g] * H[sx<9>(3 * g + i, 3 * hh + j)] * R.S3[hh * 3 + b];
                    Fi[sx<9>(3 * a + i, 3 * b + j)] = acc;
                }
    ok[k] = 1;
}

// stage 2 of the hinge Hessian on lane triples (the barrier's order-9
// projection) into Hl[144 (base + k)], blocks (a, b) at 9 (4 a + b)
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_hinge_proj(int nH, int base, const double* __restrict__ F, const unsigned char* __restrict__ ok_in, int project,
             ReducedRoots R, double* __restrict__ Hl)
{
    constexpr int PER_WAVE = 21;
    const int wl = threadIdx.x & 63, g = wl / 3, r = wl % 3, lb = 3 * g;
    const long waves = (long)gridDim.x * (blockDim.x / 64);
    const long stride = waves * PER_WAVE;
    const long wave_id = (long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const long iters = (nH + stride - 1) / stride;
    for (long it = 0; it < iters; ++it) {
        const long k = it * stride + wave_id * PER_WAVE + g;
        if (g >= PER_WAVE || k >= nH) continue;
        double* o = Hl + 144 * (size_t)(base + k);
        if (!ok_in[k]) {
            for (int q = r; q < 144; q += 3) o[q] = 0.0;
            continue;
        }
        const double* Fi = F + 90 * (size_t)k;
        double L[9], V[27];
        bool clamped = false;
        if (project) {
            double K[45];
#pragma unroll
            for (int q = 0; q < 45; ++q) K[q] = Fi[q];
            tri_jacobi(K, V, r);
            double lmin = K[0];
#pragma unroll
            for (int q = 1; q < 9; ++q) lmin = fmin(lmin, K[sx<9>(q, q)]);
            clamped = lmin < 0.0;
#pragma unroll
            for (int q = 0; q < 9; ++q) L[q] = fmax(K[sx<9>(q, q)], 0.0);
        }
        tri_blocks<true>(r, lb, 1, clamped, L, V, R.Si3, Fi + 45, o);
    }
}

void launch_flow_energy(hipStream_t s, const MeshDev& m, const double* X, double h, double* part)
{
    const SegDev& g = m.inst.sblk;
    IDP_LAUNCH(k_flow_energy, g.nblk, RB, 0, s, g.lo.p, g.hi.p, m.tris.p, m.tvol.p, X, h, part);
}

void launch_flow_grad(hipStream_t s, const MeshDev& m, const double* X, double h, double* gl)
{
    IDP_LAUNCH(k_flow_grad, grid_for(m.nE, 256), 256, 0, s, m.nE, m.nT, m.tris.p, m.tvol.p, X, h, gl);
}

void launch_flow_hess(hipStream_t s, const MeshDev& m, double h, double* Hl)
{
    IDP_LAUNCH(k_flow_hess, grid_for(9L * m.nE, 256), 256, 0, s, m.nE, m.nT, m.tvol.p, h, Hl);
}

void launch_shell_energy(hipStream_t s, const MeshDev& m, const double* X, double w, double* pm, double* ph)
{
    const SegDev& g = m.inst.sblk;
    IDP_LAUNCH(k_membrane_energy, g.nblk, RB, 0, s, g.lo.p, g.hi.p, m.tris.p, m.tIB.p, m.tvol.p, m.tlam.p, m.tmu.p,
               m.dbc.p, X, w, pm);
    const SegDev& hb = m.inst.hblk;
    IDP_LAUNCH(k_hinge_energy, hb.nblk, RB, 0, s, hb.lo.p, hb.hi.p, (const int4*)m.hinge.p, m.hrest.p, m.dbc.p, X,
               w * m.bend_k, ph);
}

void launch_shell_grad(hipStream_t s, const MeshDev& m, const double* X, double w, double* gl)
{
    IDP_LAUNCH(k_membrane_grad, grid_for(m.nE, 256), 256, 0, s, m.nE, m.nT, m.tris.p, m.tIB.p, m.tvol.p, m.tlam.p,
               m.tmu.p, m.dbc.p, X, w, gl);
    IDP_LAUNCH(k_hinge_grad, grid_for(m.nH, 256), 256, 0, s, m.nH, m.nT + m.nE, (const int4*)m.hinge.p, m.hrest.p,
               m.dbc.p, X, w * m.bend_k, gl);
}

void launch_shell_hess(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* F,
                       unsigned char* ok, double* Hl)
{
    static const ReducedRoots R = reduced_roots();
    IDP_LAUNCH(k_membrane_hess, grid_for(m.nE, 256), 256, 0, s, m.nE, m.nT, m.tris.p, m.tIB.p, m.tvol.p, m.tlam.p,
               m.tmu.p, m.dbc.p, X, w, project, R, Hl);
    if (m.nH > 0) {
        IDP_LAUNCH(k_hinge_form, grid_for(m.nH, 256), 256, 0, s, m.nH, (const int4*)m.hinge.p, m.hrest.p, m.dbc.p,
                   X, w * m.bend_k, R, F, ok);
        IDP_LAUNCH(k_hinge_proj, (int)std::min<long>(grid_for(m.nH, 4 * 21), 4096), 256, 0, s, m.nH, m.nT + m.nE, F,
                   ok, project, R, Hl);
    }
}

void launch_normal_flow(hipStream_t s, const MeshDev& m, const double* X, double magnitude, double* b)
{
    IDP_LAUNCH(k_normal_flow, grid_for(m.nV, 256), 256, 0, s, m.nV, m.vt_ptr.p, m.vt.p, m.tris.p, X, m.mass.p,
               magnitude, b);
}

}  // namespace idp
