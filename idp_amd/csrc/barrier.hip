// IPC barrier energy, gradient and PSD-projected Hessian per constraint
// (Library/FEM/IPC.h:742-1731, BARRIER.h, UTILS.h makePD, EDGE_EDGE_MOLLIFIER.h).
//
// Formulation.  Every barrier term is translation invariant, so it is a
// function of k-1 difference vectors y = (b (x) I3) x of its k vertices:
//   PT   (p,t0,t1,t2): y = (p-t0, t1-t0, t2-t0)          d = (a.(u x v))^2/|u x v|^2
//   EE   (a0,a1,b0,b1): y = (b0-a0, a1-a0, b1-b0)         same plane form
//   PE   (p,e0,e1):     y = (p-e0, e1-e0)                 d = |a|^2 - (a.u)^2/|u|^2
//   PP   (p,q):         y = p-q                           d = |a|^2
//   mollified EE/PE/PP use the EE coordinates and multiply by e(|u x v|^2).
// Gradient and Hessian are evaluated in closed form in y (dimension 9/6/3),
// then H_x = B^T H_y B.  The PSD clamp of H_x is computed exactly in the
// reduced space: with C = (b b^T)^(1/2) (x) I3, the nonzero spectrum of
// H_x equals that of K = C H_y C, and makePD(H_x) = B^T C^-1 makePD(K) C^-1 B.
// K is diagonalised by cyclic Jacobi, one constraint per lane (packed K in
// registers, eigenvectors in LDS).
#include "common.hpp"
#include "geom_dev.hpp"
#include "kernels.hpp"

namespace idp {

__constant__ double c_S3pt[9], c_Si3pt[9];  // PT  (b b^T)^(1/2), inverse
__constant__ double c_S3ee[9], c_Si3ee[9];  // EE / mollified
__constant__ double c_S2[4], c_Si2[4];      // PE
__constant__ double c_bpt[12], c_bee[12], c_bpe[6];

namespace {

// host: symmetric square root of a small SPD matrix via Jacobi
void host_sqrtm(int n, const double* A, double* S, double* Si)
{
    double a[9], v[9];
    for (int i = 0; i < n * n; ++i) a[i] = A[i];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) v[i * n + j] = (i == j);
    for (int sw = 0; sw < 50; ++sw)
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = a[p * n + q];
                if (apq == 0) continue;
                double th = (a[q * n + q] - a[p * n + p]) / (2 * apq);
                double t = (th >= 0 ? 1 : -1) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double x = a[k * n + p], y = a[k * n + q];
                    a[k * n + p] = c * x - s * y;
                    a[k * n + q] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    double x = a[p * n + k], y = a[q * n + k];
                    a[p * n + k] = c * x - s * y;
                    a[q * n + k] = s * x + c * y;
                }
                for (int k = 0; k < n; ++k) {
                    double x = v[k * n + p], y = v[k * n + q];
                    v[k * n + p] = c * x - s * y;
                    v[k * n + q] = s * x + c * y;
                }
            }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0, si = 0;
            for (int k = 0; k < n; ++k) {
                double l = sqrt(a[k * n + k]);
                s += v[i * n + k] * l * v[j * n + k];
                si += v[i * n + k] / l * v[j * n + k];
            }
            S[i * n + j] = s;
            Si[i * n + j] = si;
        }
}

void bbt(int r, int c, const double* b, double* o)
{
    for (int i = 0; i < r; ++i)
        for (int j = 0; j < r; ++j) {
            double s = 0;
            for (int k = 0; k < c; ++k) s += b[i * c + k] * b[j * c + k];
            o[i * r + j] = s;
        }
}

}  // namespace

void barrier_init_constants()
{
    const double bpt[12] = {1, -1, 0, 0, 0, -1, 1, 0, 0, -1, 0, 1};
    const double bee[12] = {-1, 0, 1, 0, -1, 1, 0, 0, 0, 0, -1, 1};
    const double bpe[6] = {1, -1, 0, 0, -1, 1};
    double M[9], S[9], Si[9];
    bbt(3, 4, bpt, M);
    host_sqrtm(3, M, S, Si);
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_S3pt), S, sizeof(S)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_Si3pt), Si, sizeof(Si)));
    bbt(3, 4, bee, M);
    host_sqrtm(3, M, S, Si);
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_S3ee), S, sizeof(S)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_Si3ee), Si, sizeof(Si)));
    bbt(2, 3, bpe, M);
    host_sqrtm(2, M, S, Si);
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_S2), S, 4 * sizeof(double)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_Si2), Si, 4 * sizeof(double)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_bpt), bpt, sizeof(bpt)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_bee), bee, sizeof(bee)));
    IDP_HIP(hipMemcpyToSymbol(HIP_SYMBOL(c_bpe), bpe, sizeof(bpe)));
}

// ---------------------------------------------------------------------------
// constraint decoding
// ---------------------------------------------------------------------------
enum Kind { K_PT = 0, K_EE = 1, K_PE = 2, K_PP = 3, K_MEE = 4, K_MPE = 5, K_MPP = 6 };

struct Decoded {
    int kind;
    int nv;      // vertices carrying the local matrix
    int v[4];    // vertex ids in the reference's emission order
    double mult; // multiplicity (PE/PP merged duplicates)
};

__device__ inline Decoded decode(int4 c)
{
    Decoded d;
    d.mult = 1.0;
    if (c.x >= 0) {
        if (c.z >= 0 && c.w >= 0) {
            d.kind = K_EE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = c.w;
        } else if (c.w >= 0) {
            d.kind = K_MEE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = c.y; d.v[2] = -c.z - 1; d.v[3] = c.w;
        } else if (c.z >= 0) {
            d.kind = K_MPE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = -c.w - 1; d.v[2] = c.y; d.v[3] = c.z;
        } else {
            d.kind = K_MPP;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = -c.z - 1; d.v[2] = c.y; d.v[3] = -c.w - 1;
        }
    } else {
        int p = -c.x - 1;
        if (c.w >= 0) {
            d.kind = K_PT;
            d.nv = 4;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = c.w;
        } else if (c.z >= 0) {
            d.kind = K_PE;
            d.nv = 3;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = -1;
            d.mult = (double)(-c.w);
        } else {
            d.kind = K_PP;
            d.nv = 2;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = -1; d.v[3] = -1;
            d.mult = (double)(-c.w);
        }
    }
    return d;
}

// squared distance exactly as Compute_Barrier / Compute_Min_Dist2 evaluate it
__device__ inline double constraint_d2(const Decoded& k, const double* X)
{
    switch (k.kind) {
    case K_PT: return d2_pt(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_EE:
    case K_MEE: return d2_ee(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_PE: return d2_pe(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]));
    case K_MPE: return d2_pe(ld3(X, k.v[0]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_PP: return d2_pp(ld3(X, k.v[0]), ld3(X, k.v[1]));
    default: return d2_pp(ld3(X, k.v[0]), ld3(X, k.v[2]));
    }
}

__device__ inline double mollifier_value(const Decoded& k, const double* X, const double* X0, double& q,
                                         double& eps)
{
    q = ee_cross_sqn(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    eps = mollifier_eps(ld3(X0, k.v[0]), ld3(X0, k.v[1]), ld3(X0, k.v[2]), ld3(X0, k.v[3]));
    if (q < eps) {
        double r = q / eps;
        return (2.0 - r) * r;
    }
    return 1.0;
}

// ---------------------------------------------------------------------------
// energy and minimum distance
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(RB) k_barrier_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                       const int4* __restrict__ cs, const double* __restrict__ X,
                                                       const double* __restrict__ X0, BarrierParams bp,
                                                       double* __restrict__ part, int* __restrict__ err)
{
    __shared__ double sh[RB];
    const int i = lo[blockIdx.x] + threadIdx.x;
    double e = 0.0;
    if (i < hi[blockIdx.x]) {
        Decoded k = decode(cs[i]);
        const int inst = bp.inst_of(k.v[0]);
        double d = constraint_d2(k, X) - bp.th2[inst];
        if (d <= 0) {
            atomicOr(err + inst, 1);
        } else {
            e = barrier_b(d, bp.dh[inst], bp.kappa[inst]);
            if (k.kind >= K_MEE) {
                double q, eps;
                e *= mollifier_value(k, X, X0, q, eps);
            } else if (k.kind == K_PE || k.kind == K_PP) {
                if (k.mult > 1.0) e *= k.mult;
            }
        }
    }
    double s = block_sum<RB>(e, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// dmin: one running minimum per instance (vinst null: a single one)
__global__ void __launch_bounds__(RB) k_min_dist2(long n, const int4* __restrict__ cs,
                                                  const double* __restrict__ X, const int* __restrict__ vinst,
                                                  double* __restrict__ dist2, unsigned long long* __restrict__ dmin)
{
    long i = (long)blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    double d = constraint_d2(k, X);
    if (dist2) dist2[i] = d;
    unsigned long long* m = dmin + (vinst ? vinst[k.v[0]] : 0);
    // squared distances are >= 0, so their IEEE bit patterns order like integers
    if (d >= 0) atomicMin(m, (unsigned long long)__double_as_longlong(d));
    else atomicMin(m, 0ull);
}

// ---------------------------------------------------------------------------
// reduced-coordinate derivatives, specialised per constraint kind
// ---------------------------------------------------------------------------
template <int KIND>
struct KT;
template <> struct KT<K_PT> { static constexpr int M = 3, NV = 4; static constexpr bool MOLL = false; };
template <> struct KT<K_EE> { static constexpr int M = 3, NV = 4; static constexpr bool MOLL = false; };
template <> struct KT<K_PE> { static constexpr int M = 2, NV = 3; static constexpr bool MOLL = false; };
template <> struct KT<K_PP> { static constexpr int M = 1, NV = 2; static constexpr bool MOLL = false; };
template <> struct KT<K_MEE> { static constexpr int M = 3, NV = 4; static constexpr bool MOLL = true; };
template <> struct KT<K_MPE> { static constexpr int M = 3, NV = 4; static constexpr bool MOLL = true; };
template <> struct KT<K_MPP> { static constexpr int M = 3, NV = 4; static constexpr bool MOLL = true; };

// Symmetric matrices are held as packed upper triangles: entry (i, j) of an
// N x N matrix at sx<N>(i, j).  All loops over entries are fully unrolled, so
// every index is a compile-time constant and the arrays stay in registers.
template <int N>
__device__ constexpr int sx(int i, int j)
{
    return (i <= j) ? (i * (2 * N - i + 1)) / 2 + (j - i) : (j * (2 * N - j + 1)) / 2 + (i - j);
}

// entry (r, c) of [w]x  ([w]x z = w x z)
__device__ inline double skew_entry(int r, int c, d3 w)
{
    if (r == 0 && c == 1) return -w.z;
    if (r == 0 && c == 2) return w.y;
    if (r == 1 && c == 0) return w.z;
    if (r == 1 && c == 2) return -w.x;
    if (r == 2 && c == 0) return -w.y;
    if (r == 2 && c == 1) return w.x;
    return 0.0;
}

__device__ inline double comp(d3 w, int i) { return i == 0 ? w.x : (i == 1 ? w.y : w.z); }

// q = |u x v|^2 over y = (a,u,v): gradient (a-block zero)
__device__ inline void q_grad(d3 u, d3 v, double uu, double vv, double uv, double* gq)
{
#pragma clang fp contract(fast)
    gq[0] = gq[1] = gq[2] = 0.0;
    gq[3] = 2 * vv * u.x - 2 * uv * v.x;
    gq[4] = 2 * vv * u.y - 2 * uv * v.y;
    gq[5] = 2 * vv * u.z - 2 * uv * v.z;
    gq[6] = 2 * uu * v.x - 2 * uv * u.x;
    gq[7] = 2 * uu * v.y - 2 * uv * u.y;
    gq[8] = 2 * uu * v.z - 2 * uv * u.z;
}

// entry (i, j) of the Hessian of q (a-blocks zero)
__device__ inline double q_hess(int i, int j, d3 u, d3 v, double uu, double vv, double uv)
{
#pragma clang fp contract(fast)
    const int bi = i / 3, bj = j / 3, ii = i % 3, jj = j % 3;
    const double id = (ii == jj) ? 1.0 : 0.0;
    if (bi == 1 && bj == 1) return 2 * vv * id - 2 * comp(v, ii) * comp(v, jj);
    if (bi == 2 && bj == 2) return 2 * uu * id - 2 * comp(u, ii) * comp(u, jj);
    if (bi == 1 && bj == 2) return 4 * comp(u, ii) * comp(v, jj) - 2 * comp(v, ii) * comp(u, jj) - 2 * uv * id;
    if (bi == 2 && bj == 1) return 4 * comp(u, jj) * comp(v, ii) - 2 * comp(v, jj) * comp(u, ii) - 2 * uv * id;
    return 0.0;
}

// plane form d = s^2/q, s = a.(u x v), q = |u x v|^2 over y=(a,u,v):
// gradient on construction, Hessian entries on demand (no stored matrix)
struct PlaneForm {
    d3 a, u, v;
    double gs[9], gq[9];
    double uu, vv, uv, c1, c2, css, csq, cqq;
    __device__ PlaneForm(d3 a_, d3 u_, d3 v_, double* gd) : a(a_), u(u_), v(v_)
    {
#pragma clang fp contract(fast)
        uu = sqn(u);
        vv = sqn(v);
        uv = dot(u, v);
        q_grad(u, v, uu, vv, uv, gq);
        const d3 n = cross(u, v);
        const double s = dot(a, n);
        const double q = sqn(n);
        const d3 gsu = cross(v, a), gsv = cross(a, u);
        gs[0] = n.x; gs[1] = n.y; gs[2] = n.z;
        gs[3] = gsu.x; gs[4] = gsu.y; gs[5] = gsu.z;
        gs[6] = gsv.x; gs[7] = gsv.y; gs[8] = gsv.z;
        const double iq = 1.0 / q;
        c1 = 2 * s * iq;        // d(d)/ds
        c2 = -s * s * iq * iq;  // d(d)/dq
        for (int i = 0; i < 9; ++i) gd[i] = c1 * gs[i] + c2 * gq[i];
        css = 2 * iq;
        csq = -2 * s * iq * iq;
        cqq = 2 * s * s * iq * iq * iq;
    }
    __device__ double h(int i, int j) const
    {
#pragma clang fp contract(fast)
        double r = css * gs[i] * gs[j] + csq * (gs[i] * gq[j] + gq[i] * gs[j]) + cqq * gq[i] * gq[j] +
                   c2 * q_hess(i, j, u, v, uu, vv, uv);
        // + (2s/q) Hess s: [a,u] = -[v]x, [a,v] = [u]x, [u,v] = -[a]x and transposes
        const int bi = i / 3, bj = j / 3;
        if (bi == 0 && bj == 1) r += -c1 * skew_entry(i % 3, j % 3, v);
        if (bi == 1 && bj == 0) r += c1 * skew_entry(i % 3, j % 3, v);
        if (bi == 0 && bj == 2) r += c1 * skew_entry(i % 3, j % 3, u);
        if (bi == 2 && bj == 0) r += -c1 * skew_entry(i % 3, j % 3, u);
        if (bi == 1 && bj == 2) r += -c1 * skew_entry(i % 3, j % 3, a);
        if (bi == 2 && bj == 1) r += c1 * skew_entry(i % 3, j % 3, a);
        return r;
    }
};

// point-edge form d = |a|^2 - (a.w)^2/|w|^2 in blocks IA < IW of a 9-wide layout
template <int IA, int IW>
struct PEForm {
    double ac[3], wc[3], c, iw2;
    __device__ PEForm(d3 a, d3 w, double* gd)
    {
#pragma clang fp contract(fast)
        c = dot(a, w);
        iw2 = 1.0 / sqn(w);
        ac[0] = a.x; ac[1] = a.y; ac[2] = a.z;
        wc[0] = w.x; wc[1] = w.y; wc[2] = w.z;
        for (int i = 0; i < 3; ++i) {
            gd[IA + i] = 2 * ac[i] - 2 * c * iw2 * wc[i];
            gd[IW + i] = -2 * c * iw2 * ac[i] + 2 * c * c * iw2 * iw2 * wc[i];
        }
    }
    __device__ double aw(int i, int j) const
    {
#pragma clang fp contract(fast)
        const double id = (i == j) ? 1.0 : 0.0;
        return -2 * iw2 * wc[i] * ac[j] + 4 * c * iw2 * iw2 * wc[i] * wc[j] - 2 * c * iw2 * id;
    }
    __device__ double h(int i, int j) const
    {
#pragma clang fp contract(fast)
        const bool ia = i >= IA && i < IA + 3, iw = i >= IW && i < IW + 3;
        const bool ja = j >= IA && j < IA + 3, jw = j >= IW && j < IW + 3;
        if (ia && ja) {
            const int r = i - IA, s = j - IA;
            return 2 * (r == s ? 1.0 : 0.0) - 2 * iw2 * wc[r] * wc[s];
        }
        if (iw && jw) {
            const int r = i - IW, s = j - IW;
            return -2 * iw2 * ac[r] * ac[s] + 4 * c * iw2 * iw2 * (ac[r] * wc[s] + wc[r] * ac[s]) -
                   8 * c * c * iw2 * iw2 * iw2 * wc[r] * wc[s] + 2 * c * c * iw2 * iw2 * (r == s ? 1.0 : 0.0);
        }
        if (ia && jw) return aw(i - IA, j - IW);
        if (iw && ja) return aw(j - IA, i - IW);
        return 0.0;
    }
};

// point-point d = |a|^2
struct PPForm {
    __device__ PPForm(d3 a, double* gd)
    {
        gd[0] = 2 * a.x;
        gd[1] = 2 * a.y;
        gd[2] = 2 * a.z;
    }
    __device__ double h(int i, int j) const { return (i == j && i < 3) ? 2.0 : 0.0; }
};

// Gradient g (3M) and, when WANT_H, Hessian H (3M x 3M, packed sx<3M>) in y
// of f = mult * b(d - th2) [* e(q)].  Returns false when d - th2 <= 0.
template <int KIND, bool WANT_H, class Form>
__device__ inline void assemble_reduced(const Form& F, const double* gd, const Decoded& k, const double* X,
                                        const double* X0, double b, double b1, double b2, d3 y1, d3 y2, double* g,
                                        double* H, double& energy)
{
#pragma clang fp contract(fast)
    constexpr int N = 3 * KT<KIND>::M;
    if (!KT<KIND>::MOLL) {
        const double m = k.mult;
        energy = (m > 1.0) ? b * m : b;
#pragma unroll
        for (int i = 0; i < N; ++i) g[i] = m * b1 * gd[i];
        if (WANT_H)
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = i; j < N; ++j) H[sx<N>(i, j)] = m * (b2 * gd[i] * gd[j] + b1 * F.h(i, j));
        return;
    }
    double q, eps;
    const double e = mollifier_value(k, X, X0, q, eps);
    energy = b * e;
    if (!(q < eps)) {  // e == 1, grad e == 0
#pragma unroll
        for (int i = 0; i < 9; ++i) g[i] = b1 * gd[i];
        if (WANT_H)
#pragma unroll
            for (int i = 0; i < 9; ++i)
#pragma unroll
                for (int j = i; j < 9; ++j) H[sx<9>(i, j)] = b2 * gd[i] * gd[j] + b1 * F.h(i, j);
        return;
    }
    const double de = 2.0 / eps * (1.0 - q / eps);
    const double d2e = -2.0 / (eps * eps);
    const double uu = sqn(y1), vv = sqn(y2), uv = dot(y1, y2);
    double gq[9];
    q_grad(y1, y2, uu, vv, uv, gq);
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = e * b1 * gd[i] + b * de * gq[i];
    if (WANT_H)
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = i; j < 9; ++j) {
                const double He = de * q_hess(i, j, y1, y2, uu, vv, uv) + d2e * gq[i] * gq[j];
                H[sx<9>(i, j)] = e * b2 * gd[i] * gd[j] + e * b1 * F.h(i, j) +
                                 b1 * de * (gd[i] * gq[j] + gq[i] * gd[j]) + b * He;
            }
}

template <int KIND, bool WANT_H>
__device__ inline bool reduced_derivs(const Decoded& k, const double* X, const double* X0, double dh,
                                      double kappa, double th2, double* g, double* H, double& energy)
{
#pragma clang fp contract(fast)
    d3 y0, y1, y2;
    if (KIND == K_PT) {
        d3 p = ld3(X, k.v[0]), t0 = ld3(X, k.v[1]);
        y0 = p - t0; y1 = ld3(X, k.v[2]) - t0; y2 = ld3(X, k.v[3]) - t0;
    } else if (KIND == K_PE) {
        d3 e0 = ld3(X, k.v[1]);
        y0 = ld3(X, k.v[0]) - e0; y1 = ld3(X, k.v[2]) - e0; y2 = {0, 0, 0};
    } else if (KIND == K_PP) {
        y0 = ld3(X, k.v[0]) - ld3(X, k.v[1]); y1 = {0, 0, 0}; y2 = {0, 0, 0};
    } else {  // (ea0, ea1, eb0, eb1)
        d3 a0 = ld3(X, k.v[0]), b0 = ld3(X, k.v[2]);
        y0 = b0 - a0; y1 = ld3(X, k.v[1]) - a0; y2 = ld3(X, k.v[3]) - b0;
    }
    // the distance value exactly as the energy kernel evaluates it
    const double d = constraint_d2(k, X) - th2;
    if (!(d > 0)) return false;
    const double b = barrier_b(d, dh, kappa);
    const double b1 = barrier_db(d, dh, kappa);
    const double b2 = WANT_H ? barrier_d2b(d, dh, kappa) : 0.0;
    double gd[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) gd[i] = 0.0;
    if (KIND == K_PT || KIND == K_EE || KIND == K_MEE) {
        PlaneForm F(y0, y1, y2, gd);
        assemble_reduced<KIND, WANT_H>(F, gd, k, X, X0, b, b1, b2, y1, y2, g, H, energy);
    } else if (KIND == K_PE) {
        PEForm<0, 3> F(y0, y1, gd);
        assemble_reduced<KIND, WANT_H>(F, gd, k, X, X0, b, b1, b2, y1, y2, g, H, energy);
    } else if (KIND == K_MPE) {
        PEForm<0, 6> F(y0, y2, gd);
        assemble_reduced<KIND, WANT_H>(F, gd, k, X, X0, b, b1, b2, y1, y2, g, H, energy);
    } else {
        PPForm F(y0, gd);
        assemble_reduced<KIND, WANT_H>(F, gd, k, X, X0, b, b1, b2, y1, y2, g, H, energy);
    }
    return true;
}

template <int KIND>
__device__ inline const double* bmat()
{
    if (KIND == K_PT) return c_bpt;
    if (KIND == K_PE) return c_bpe;
    return c_bee;
}

// ---------------------------------------------------------------------------
// Cyclic Jacobi eigen-decomposition of a symmetric N x N matrix.  A is held in
// registers as the packed upper triangle; the eigenvector matrix W lives in
// LDS (element (k, p) of this lane at W[(9 k + p) * WS]).
// ---------------------------------------------------------------------------
constexpr int WS = 64;  // lanes per block of the Hessian kernel = LDS stride

template <int N>
__device__ inline void jacobi(double* A, double* W)
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) W[(i * 9 + j) * WS] = (i == j) ? 1.0 : 0.0;
#pragma nounroll
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dg += A[sx<N>(i, i)] * A[sx<N>(i, i)];
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += A[sx<N>(i, j)] * A[sx<N>(i, j)];
        }
        if (!(off > 1e-34 * dg)) break;
#pragma unroll
        for (int p = 0; p < N - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[sx<N>(p, q)];
                if (apq != 0.0) {
                    const double theta = (A[sx<N>(q, q)] - A[sx<N>(p, p)]) / (2.0 * apq);
                    const double at = fabs(theta);
                    const double t = (at > 1e150) ? 0.5 / theta
                                                  : copysign(1.0, theta) / (at + sqrt(theta * theta + 1.0));
                    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                    for (int k2 = 0; k2 < N; ++k2) {
                        if (k2 == p || k2 == q) continue;
                        const double akp = A[sx<N>(k2, p)], akq = A[sx<N>(k2, q)];
                        A[sx<N>(k2, p)] = c * akp - s * akq;
                        A[sx<N>(k2, q)] = s * akp + c * akq;
                    }
                    A[sx<N>(p, p)] = A[sx<N>(p, p)] - t * apq;
                    A[sx<N>(q, q)] = A[sx<N>(q, q)] + t * apq;
                    A[sx<N>(p, q)] = 0.0;
#pragma unroll
                    for (int k2 = 0; k2 < N; ++k2) {
                        const double wkp = W[(k2 * 9 + p) * WS], wkq = W[(k2 * 9 + q) * WS];
                        W[(k2 * 9 + p) * WS] = c * wkp - s * wkq;
                        W[(k2 * 9 + q) * WS] = s * wkp + c * wkq;
                    }
                }
                // keep each rotation's LDS traffic in place (no hoisting of later
                // rotations' loads, which only raises register pressure)
                __builtin_amdgcn_sched_barrier(0);
            }
    }
}

// makePD in the reduced space; H (packed sx<3M>) is replaced by its
// projection.  W: this lane's LDS eigenvector workspace.
template <int KIND>
__device__ inline bool project_reduced(double* H, double* W)
{
#pragma clang fp contract(fast)
    constexpr int M = KT<KIND>::M;
    constexpr int N = 3 * M;
    const double* S = (M == 3) ? ((KIND == K_PT) ? c_S3pt : c_S3ee) : c_S2;
    const double* Si = (M == 3) ? ((KIND == K_PT) ? c_Si3pt : c_Si3ee) : c_Si2;
    double K[N * (N + 1) / 2];
    // K = C H C, C = S (x) I3  (for M == 1, C = sqrt(2) I); upper triangle
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = a; b < M; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if (a == b && j < i) continue;
                    double acc = 0.0;
                    if (M == 1) {
                        acc = 2.0 * H[sx<N>(i, j)];
                    } else {
#pragma unroll
                        for (int g = 0; g < M; ++g)
#pragma unroll
                            for (int h = 0; h < M; ++h)
                                acc += S[a * M + g] * H[sx<N>(3 * g + i, 3 * h + j)] * S[h * M + b];
                    }
                    K[sx<N>(3 * a + i, 3 * b + j)] = acc;
                }
    jacobi<N>(K, W);
    double lmin = K[sx<N>(0, 0)];
#pragma unroll
    for (int i = 1; i < N; ++i) lmin = fmin(lmin, K[sx<N>(i, i)]);
    if (lmin >= 0.0) return false;  // UTILS.h:12-15: nothing to clamp
    double lam[N];
#pragma unroll
    for (int i = 0; i < N; ++i) lam[i] = fmax(K[sx<N>(i, i)], 0.0);
#pragma unroll
    for (int i = 0; i < N; ++i) {
        double wi[N];
#pragma unroll
        for (int k2 = 0; k2 < N; ++k2) wi[k2] = W[(i * 9 + k2) * WS] * lam[k2];
#pragma unroll
        for (int j = i; j < N; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < N; ++k2) acc += wi[k2] * W[(j * 9 + k2) * WS];
            K[sx<N>(i, j)] = acc;
        }
    }
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = a; b < M; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    if (a == b && j < i) continue;
                    double acc = 0.0;
                    if (M == 1) {
                        acc = 0.5 * K[sx<N>(i, j)];
                    } else {
#pragma unroll
                        for (int g = 0; g < M; ++g)
#pragma unroll
                            for (int h = 0; h < M; ++h)
                                acc += Si[a * M + g] * K[sx<N>(3 * g + i, 3 * h + j)] * Si[h * M + b];
                    }
                    H[sx<N>(3 * a + i, 3 * b + j)] = acc;
                }
    return true;
}

// vertex gradient g_x = B^T g_y (3*NV values)
template <int KIND>
__device__ inline void expand_grad(const double* g, double* gx)
{
    constexpr int M = KT<KIND>::M, NV = KT<KIND>::NV;
    if (M == 1) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            gx[i] = g[i];
            gx[3 + i] = -g[i];
        }
        return;
    }
    const double* B = bmat<KIND>();
#pragma unroll
    for (int vtx = 0; vtx < NV; ++vtx)
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < M; ++a) acc += B[a * NV + vtx] * g[3 * a + i];
            gx[3 * vtx + i] = acc;
        }
}

// vertex block (vi, vj) of B^T H B
template <int KIND>
__device__ inline void expand_block(const double* H, int vi, int vj, double* blk)
{
#pragma clang fp contract(fast)
    constexpr int M = KT<KIND>::M, NV = KT<KIND>::NV;
    if (M == 1) {
        const double s = (vi == vj) ? 1.0 : -1.0;
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) blk[3 * i + j] = s * H[sx<3>(i, j)];
        return;
    }
    const double* B = bmat<KIND>();
#pragma unroll
    for (int i = 0; i < 9; ++i) blk[i] = 0.0;
#pragma unroll
    for (int a = 0; a < M; ++a) {
        const double ba = B[a * NV + vi];
#pragma unroll
        for (int b = 0; b < M; ++b) {
            const double s = ba * B[b * NV + vj];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) blk[3 * i + j] += s * H[sx<3 * M>(3 * a + i, 3 * b + j)];
        }
    }
}

template <int KIND>
__device__ inline void grad_body(const Decoded& k, const double* X, const double* X0, double dh, double kappa,
                                 double th2, double* out, int* err)
{
    double g[9], H[1], e;
    double gx[12];
    if (!reduced_derivs<KIND, false>(k, X, X0, dh, kappa, th2, g, H, e)) {
        atomicOr(err, 1);
#pragma unroll
        for (int j = 0; j < 12; ++j) out[j] = 0.0;
        return;
    }
    expand_grad<KIND>(g, gx);
    constexpr int NV = KT<KIND>::NV;
#pragma unroll
    for (int j = 0; j < 12; ++j) out[j] = (j < 3 * NV) ? gx[j] : 0.0;
}

template <int KIND>
__device__ inline void hess_body(const Decoded& k, const double* X, const double* X0, double dh, double kappa,
                                 double th2, int project_spd, int mode, double* out, int* err, int* counts,
                                 double* W)
{
    constexpr int NV = KT<KIND>::NV;
    double g[9], H[45], e;
    if (!reduced_derivs<KIND, true>(k, X, X0, dh, kappa, th2, g, H, e)) {
        atomicOr(err, 1);
        for (int j = 0; j < 144; ++j) out[j] = 0.0;
        return;
    }
    bool clamped = false;
    if (project_spd) clamped = project_reduced<KIND>(H, W);
    if (counts) {
        atomicAdd(counts + 2 * KIND, 1);
        if (clamped) atomicAdd(counts + 2 * KIND + 1, 1);
    }
    double blk[9];
#pragma nounroll
    for (int a = 0; a < NV; ++a)
#pragma nounroll
        for (int b = 0; b < NV; ++b) {
            expand_block<KIND>(H, a, b, blk);
            if (mode == 0) {
#pragma unroll
                for (int r = 0; r < 3; ++r)
#pragma unroll
                    for (int c = 0; c < 3; ++c) out[(3 * a + r) * (3 * NV) + 3 * b + c] = blk[3 * r + c];
            } else {
#pragma unroll
                for (int r = 0; r < 9; ++r) out[9 * (a * NV + b) + r] = blk[r];
            }
        }
    if (mode == 0)
        for (int j = 9 * NV * NV; j < 144; ++j) out[j] = 0.0;
}

#define IDP_KIND_SWITCH(kind, F, ...)                                                                    \
    switch (kind) {                                                                                      \
    case K_PT: F<K_PT>(__VA_ARGS__); break;                                                               \
    case K_EE: F<K_EE>(__VA_ARGS__); break;                                                               \
    case K_PE: F<K_PE>(__VA_ARGS__); break;                                                               \
    case K_PP: F<K_PP>(__VA_ARGS__); break;                                                               \
    case K_MEE: F<K_MEE>(__VA_ARGS__); break;                                                             \
    case K_MPE: F<K_MPE>(__VA_ARGS__); break;                                                             \
    default: F<K_MPP>(__VA_ARGS__); break;                                                                \
    }

// gradient: per-constraint vertex contributions (12 doubles; ids from decode)
__global__ void __launch_bounds__(256) k_barrier_grad(long n, const int4* __restrict__ cs,
                                                      const double* __restrict__ X,
                                                      const double* __restrict__ X0, BarrierParams bp,
                                                      double* __restrict__ gout, int* __restrict__ err)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    const int inst = bp.inst_of(k.v[0]);
    if (bp.skip(inst)) return;
    IDP_KIND_SWITCH(k.kind, grad_body, k, X, X0, bp.dh[inst], bp.kappa[inst], bp.th2[inst], gout + 12 * i,
                    err + inst)
}

// Hessian: mode 0 writes the dense (3nv)^2 local matrix (144 per constraint,
// row-major, vertex order of decode); mode 1 writes nv*nv vertex blocks of 9
// (block (a,b) at 9*(a*nv+b)).
__global__ void __launch_bounds__(WS) k_barrier_hess(long n, const int4* __restrict__ cs,
                                                     const double* __restrict__ X,
                                                     const double* __restrict__ X0, BarrierParams bp,
                                                     int project_spd, int mode,
                                                     double* __restrict__ out, int* __restrict__ vid,
                                                     int* __restrict__ err, int* __restrict__ counts)
{
    __shared__ double wsh[81 * WS];
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    const int inst = bp.inst_of(k.v[0]);
    if (bp.skip(inst)) return;
    IDP_KIND_SWITCH(k.kind, hess_body, k, X, X0, bp.dh[inst], bp.kappa[inst], bp.th2[inst], project_spd, mode,
                    out + 144 * i, err + inst, counts,
                    wsh + threadIdx.x)
    if (vid)
        for (int j = 0; j < 4; ++j) vid[4 * i + j] = (j < k.nv) ? k.v[j] : -1;
}

// vertex ids of each constraint in decode order (-1 padded)
__global__ void k_constraint_vids(long n, const int4* __restrict__ cs, int* __restrict__ vid, int* __restrict__ nv)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    for (int j = 0; j < 4; ++j) vid[4 * i + j] = (j < k.nv) ? k.v[j] : -1;
    nv[i] = k.nv;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
void launch_barrier_energy(hipStream_t s, int nblk, const int* lo, const int* hi, const int* cs, const double* X,
                           const double* X0, const BarrierParams& bp, double* part, int* err)
{
    IDP_LAUNCH(k_barrier_energy, nblk, RB, 0, s, lo, hi, (const int4*)cs, X, X0, bp, part, err);
}

void launch_min_dist2(hipStream_t s, long n, const int* cs, const double* X, const int* vinst, double* dist2,
                      unsigned long long* dmin)
{
    IDP_LAUNCH(k_min_dist2, grid_for(n, RB), RB, 0, s, n, (const int4*)cs, X, vinst, dist2, dmin);
}

void launch_barrier_grad(hipStream_t s, long n, const int* cs, const double* X, const double* X0,
                         const BarrierParams& bp, double* gout, int* err)
{
    IDP_LAUNCH(k_barrier_grad, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, X, X0, bp, gout, err);
}

void launch_barrier_hess(hipStream_t s, long n, const int* cs, const double* X, const double* X0,
                         const BarrierParams& bp, int project_spd, int mode, double* out, int* vid, int* err,
                         int* counts)
{
    IDP_LAUNCH(k_barrier_hess, grid_for(n, WS), WS, 0, s, n, (const int4*)cs, X, X0, bp, project_spd, mode, out,
               vid, err, counts);
}

void launch_constraint_vids(hipStream_t s, long n, const int* cs, int* vid, int* nv)
{
    IDP_LAUNCH(k_constraint_vids, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, vid, nv);
}

}  // namespace idp
