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
// The Hessian runs in two stages over the constraints grouped by reduced
// order: k_barrier_form (one per lane) forms H_y and K; then K of order 6
// (PE) and 3 (PP) is diagonalised by cyclic Jacobi one constraint per lane
// with K and its eigenvectors in registers (k_barrier_hess_lane), and K of
// order 9 (PT, EE, mollified) by a round-robin parallel Jacobi on a team of
// lanes in LDS (k_barrier_hess).
#include "common.hpp"
#include "constraint_dev.hpp"
#include "geom_dev.hpp"
#include "kernels.hpp"
#include "proj_dev.hpp"

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

ReducedRoots reduced_roots()
{
    ReducedRoots r;
    const double bpt[12] = {1, -1, 0, 0, 0, -1, 1, 0, 0, -1, 0, 1};
    const double bpe[6] = {1, -1, 0, 0, -1, 1};
    double M[9];
    bbt(3, 4, bpt, M);
    host_sqrtm(3, M, r.S3, r.Si3);
    bbt(2, 3, bpe, M);
    host_sqrtm(2, M, r.S2, r.Si2);
    return r;
}

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

// Range [x, y) of the constraints whose reduced order is N (9: PT, EE and the
// mollified kinds; 6: PE; 3: PP) in the grouped order, from the kind counts
__device__ inline int2 group_range(const unsigned* __restrict__ cnt, int N)
{
    const int n9 = (int)(cnt[K_PT] + cnt[K_EE] + cnt[K_MEE] + cnt[K_MPE] + cnt[K_MPP]);
    const int n6 = (int)cnt[K_PE], n3 = (int)cnt[K_PP];
    if (N == 9) return make_int2(0, n9);
    if (N == 6) return make_int2(n9, n9 + n6);
    return make_int2(n9 + n6, n9 + n6 + n3);
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

// ---------------------------------------------------------------------------
// Barrier Hessian with the PSD projection, one constraint per team of TL
// lanes (HB / TL teams per wavefront-sized workgroup; TL = 8, 16, 32 or 64 is
// a launch parameter, and no sum depends on it, so neither do the results).  The team's first lane
// forms the reduced Hessian H_y (N = 3M) and K = C H_y C in the team's LDS;
// the team then diagonalises K by cyclic Jacobi, each rotation's row and
// eigenvector updates spread over the lanes (lane k updates row k), and
// spreads the clamp V max(L, 0) V^T, the map back C^-1 (.) C^-1 and the
// vertex blocks B^T H_y B over its lanes as well.  The loops index LDS, so
// the kernel stays small enough for the instruction cache, and a constraint
// takes TL lanes instead of one, which shortens the serial chain per
// launch.
// ---------------------------------------------------------------------------
constexpr int HB = 64;  // lanes per workgroup: one wavefront

struct TeamLds {
    double K[2][81];  // the matrix being diagonalised, row-major, stride 9 (double buffer)
    double V[2][81];  // eigenvectors (columns), stride 9 (double buffer)
    double H[81];     // H_y, stride 9; projected in place
    double rc[10], rs[10], rt[10];  // this round's rotation seen from index i: c, signed s, t
    int rp[10];                      // ... and i's partner (i itself when i does not rotate)
    double lam[9];
    int n, kind, ok, clamped;  // ok: 1 formed, 0 d <= 0, 2 nothing to do
    int done;                  // Jacobi converged
};

// lanes of one team see each other's LDS writes in program order (the team
// is inside one wavefront): keep the compiler from moving LDS accesses
// across this point
__device__ inline void team_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// runtime index of (i, j), i <= j, in a packed upper triangle of order n
__device__ inline int pk(int i, int j, int n) { return (i * (2 * n - i + 1)) / 2 + (j - i); }

// Stage 1, one constraint per lane: the reduced Hessian H_y (N = 3M) and
// K = C H_y C (C = S (x) I3; M == 1: sqrt(2) I), both as packed upper
// triangles of order N (45 doubles each) at F + 90 i; ok[i] = 0 when the
// distance is not positive (the constraint's blocks are zero).
template <int KIND>
__device__ inline void form_body(const Decoded& k, const double* X, const double* X0, double dh, double kappa,
                                 double th2, double* F, unsigned char* ok, int* err)
{
#pragma clang fp contract(fast)
    constexpr int M = KT<KIND>::M;
    constexpr int N = 3 * M;
    double g[9], H[45], e;
    if (!reduced_derivs<KIND, true>(k, X, X0, dh, kappa, th2, g, H, e)) {
        atomicOr(err, 1);
        *ok = 0;
        return;
    }
    const double* S = (M == 3) ? ((KIND == K_PT) ? c_S3pt : c_S3ee) : c_S2;
#pragma unroll
    for (int q = 0; q < N * (N + 1) / 2; ++q) F[45 + q] = H[q];
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
                        for (int gg = 0; gg < M; ++gg)
#pragma unroll
                            for (int hh = 0; hh < M; ++hh)
                                acc += S[a * M + gg] * H[sx<N>(3 * gg + i, 3 * hh + j)] * S[hh * M + b];
                    }
                    F[sx<N>(3 * a + i, 3 * b + j)] = acc;
                }
    *ok = 1;
}

__global__ void __launch_bounds__(256) k_barrier_form(long n, const int* __restrict__ perm,
                                                      const int4* __restrict__ cs, const double* __restrict__ X,
                                                      const double* __restrict__ X0, BarrierParams bp,
                                                      double* __restrict__ F, unsigned char* __restrict__ ok,
                                                      int* __restrict__ err)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const long i = perm ? (long)perm[t] : t;
    Decoded k = decode(cs[i]);
    const int inst = bp.inst_of(k.v[0]);
    if (bp.skip(inst)) {
        ok[t] = 2;
        return;
    }
    IDP_KIND_SWITCH(k.kind, form_body, k, X, X0, bp.dh[inst], bp.kappa[inst], bp.th2[inst], F + 90 * t, ok + t,
                    err + inst)
}

// Partner of index i in round r of the circle schedule over m players
// (m even): player m-1 stays, the others turn; every pair meets once in
// m-1 rounds.
__device__ inline int rr_partner(int i, int r, int m)
{
    if (i == m - 1) return r;
    if (i == r) return m - 1;
    const int q = (2 * r - i) % (m - 1);
    return q < 0 ? q + (m - 1) : q;
}

// Parallel (round-robin) Jacobi on K (n x n): each round applies the
// rotations of n/2 disjoint index pairs at once, K <- J^T K J and V <- V J,
// every entry from the previous round's matrices (double buffer); each
// rotation is the classical one (Golub & Van Loan 8.5), with the rotated
// pair's diagonal updated as app - t apq, aqq + t apq and apq set to 0.  A
// lane owns rows of K and V.  Entries are evaluated without contraction and
// grouped so that (i, j) and (j, i) round alike: K stays exactly symmetric.
// Returns the buffer holding the result.
template <int TL>
__device__ inline int team_jacobi(TeamLds& T, int tl, int n)
{
#pragma clang fp contract(off)
    const int m = n + (n & 1);  // players of the schedule (an odd n gets an idle one)
    int cur = 0;
    for (int idx = tl; idx < 81; idx += TL) T.V[0][idx] = (idx / 9 == idx % 9) ? 1.0 : 0.0;
    team_sync();
    for (int sweep = 0; sweep < 16; ++sweep) {
        // convergence test, summed by one lane in a fixed order
        if (tl == 0) {
            const double* K = T.K[cur];
            double off = 0.0, dg = 0.0;
            for (int i = 0; i < n; ++i) {
                dg += K[10 * i] * K[10 * i];
                for (int j = i + 1; j < n; ++j) off += K[9 * i + j] * K[9 * i + j];
            }
            T.done = !(off > 1e-34 * dg);
        }
        team_sync();
        if (T.done) break;
        for (int r = 0; r < m - 1; ++r) {
            const double* K0 = T.K[cur];
            const double* V0 = T.V[cur];
            double* K1 = T.K[cur ^ 1];
            double* V1 = T.V[cur ^ 1];
            // this round's rotations, from the lane of each pair's smaller index
            for (int i = tl; i < n; i += TL) {
                const int q = rr_partner(i, r, m);
                if (q >= n) {
                    T.rp[i] = i;
                    T.rc[i] = 1.0;
                    T.rs[i] = 0.0;
                    T.rt[i] = 0.0;
                } else if (q > i) {
                    const double apq = K0[9 * i + q];
                    double c = 1.0, sn = 0.0, t = 0.0;
                    if (apq != 0.0) {
                        const double dd = K0[10 * q] - K0[10 * i];
                        const double rt = sqrt(dd * dd + 4.0 * apq * apq);
                        t = copysign(1.0, dd) * (2.0 * apq) / (fabs(dd) + rt);
                        c = rsqrt(t * t + 1.0);
                        sn = t * c;
                    }
                    T.rp[i] = q;
                    T.rc[i] = c;
                    T.rs[i] = -sn;
                    T.rt[i] = t;
                    T.rp[q] = i;
                    T.rc[q] = c;
                    T.rs[q] = sn;
                    T.rt[q] = t;
                }
            }
            team_sync();
            for (int i = tl; i < n; i += TL) {
                const int pi = T.rp[i];
                const double ci = T.rc[i], si = T.rs[i];
                for (int j = 0; j < n; ++j) {
                    const int pj = T.rp[j];
                    const double cj = T.rc[j], sj = T.rs[j];
                    V1[9 * i + j] = cj * V0[9 * i + j] + sj * V0[9 * i + pj];
                    double v;
                    if (pi != i && j == pi) {
                        v = 0.0;  // the rotated pair's off-diagonal entry
                    } else if (pi != i && j == i) {
                        const int lo = min(i, pi), hi = max(i, pi);
                        const double apq = K0[9 * lo + hi];
                        v = (i == lo) ? K0[10 * i] - T.rt[i] * apq : K0[10 * i] + T.rt[i] * apq;
                    } else {
                        const double cc = ci * cj, ss = si * sj;
                        const double a1 = (ci * sj) * K0[9 * i + pj], a2 = (si * cj) * K0[9 * pi + j];
                        v = (cc * K0[9 * i + j] + ss * K0[9 * pi + pj]) + (a1 + a2);
                    }
                    K1[9 * i + j] = v;
                }
            }
            cur ^= 1;
            team_sync();
        }
    }
    return cur;
}

// makePD (UTILS.h:9-27) of H_y through K: returns whether anything was clamped
template <int TL>
__device__ inline bool team_project(TeamLds& T, int tl, int n, int M, const double* Si)
{
#pragma clang fp contract(fast)
    const int cur = team_jacobi<TL>(T, tl, n);
    double* K = T.K[cur];
    const double* V = T.V[cur];
    double lmin = K[0];
    for (int i = 1; i < n; ++i) lmin = fmin(lmin, K[10 * i]);
    if (lmin >= 0.0) return false;  // UTILS.h:12-15: nothing to clamp
    for (int i = tl; i < n; i += TL) T.lam[i] = fmax(K[10 * i], 0.0);  // TL may be below n
    team_sync();
    // K+ = V max(L, 0) V^T (upper triangle, mirrored)
    for (int idx = tl; idx < n * n; idx += TL) {
        const int i = idx / n, j = idx % n;
        if (j < i) continue;
        double acc = 0.0;
        for (int k2 = 0; k2 < n; ++k2) acc += V[9 * i + k2] * T.lam[k2] * V[9 * j + k2];
        K[9 * i + j] = acc;
        K[9 * j + i] = acc;
    }
    team_sync();
    // H_y = C^-1 K+ C^-1
    for (int idx = tl; idx < n * n; idx += TL) {
        const int r = idx / n, cc = idx % n;
        if (cc < r) continue;
        const int a = r / 3, i = r % 3, b = cc / 3, j = cc % 3;
        double acc = 0.0;
        if (M == 1) {
            acc = 0.5 * K[9 * i + j];
        } else {
            for (int gg = 0; gg < M; ++gg)
                for (int hh = 0; hh < M; ++hh) acc += Si[a * M + gg] * K[9 * (3 * gg + i) + 3 * hh + j] * Si[hh * M + b];
        }
        T.H[9 * r + cc] = acc;
        T.H[9 * cc + r] = acc;
    }
    team_sync();
    return true;
}

// Stage 2, one constraint per team: the projection and the vertex blocks.
// mode 0 writes the dense (3nv)^2 local matrix (144 per constraint,
// row-major, vertex order of decode); mode 1 writes nv*nv vertex blocks of 9
// (block (a,b) at 9*(a*nv+b)).  perm (null: identity) lists the constraints
// grouped by kind (both stages walk it in the same order), so that the lanes
// of a wavefront take the same branch of the kind switch in stage 1; every
// result is written at the constraint's own index.
// counts (null: none): per kind, evaluations and clamped ones, one atomic
// per wavefront and kind.
// one constraint (t < n; t == n: an idle team that still takes part in the
// wavefront's ballots)
template <int TL>
__device__ inline void team_body(TeamLds& T, int tl, long t, long n, const int* __restrict__ perm,
                                 const int4* __restrict__ cs, const double* __restrict__ F,
                                 const unsigned char* __restrict__ ok_in, int project_spd, int mode,
                                 double* __restrict__ out, int* __restrict__ vid, int* __restrict__ counts)
{
#pragma clang fp contract(fast)
    // the team loads stage 1's packed K and H_y into LDS as full matrices
    if (tl == 0) {
        T.ok = 2;
        T.clamped = 0;
        if (t < n) {
            const long i = perm ? (long)perm[t] : t;
            const Decoded k = decode(cs[i]);
            T.ok = ok_in[t];
            T.kind = k.kind;
            T.n = (k.kind == K_PE) ? 6 : ((k.kind == K_PP) ? 3 : 9);
            if (vid && T.ok != 2)
                for (int j = 0; j < 4; ++j) vid[4 * i + j] = (j < k.nv) ? k.v[j] : -1;
        }
    }
    team_sync();
    if (T.ok == 1) {
        const int nn = T.n;
        const double* Fi = F + 90 * t;
        for (int idx = tl; idx < nn * nn; idx += TL) {
            const int r = idx / nn, cc = idx % nn;
            const int q = (r <= cc) ? pk(r, cc, nn) : pk(cc, r, nn);
            T.K[0][9 * r + cc] = Fi[q];
            T.H[9 * r + cc] = Fi[45 + q];
        }
    }
    team_sync();
    const int ok = T.ok;
    const long ci = (t < n) ? (perm ? (long)perm[t] : t) : 0;
    double* o = out + 144 * ci;
    if (ok == 0) {
        for (int j = tl; j < 144; j += TL) o[j] = 0.0;
    } else if (ok == 1) {
        const int kind = T.kind, nn = T.n;
        const int M = nn / 3;
        const int NV = (kind == K_PE) ? 3 : ((kind == K_PP) ? 2 : 4);
        const double* Si = (M == 3) ? ((kind == K_PT) ? c_Si3pt : c_Si3ee) : c_Si2;
        bool clamped = false;
        if (project_spd) clamped = team_project<TL>(T, tl, nn, M, Si);
        if (tl == 0) T.clamped = clamped;
        // vertex block (vi, vj) of B^T H_y B: one block per lane
        const double* B = (kind == K_PT) ? c_bpt : ((kind == K_PE) ? c_bpe : c_bee);
        for (int bk = tl; bk < NV * NV; bk += TL) {
            const int vi = bk / NV, vj = bk % NV;
            double blk[9];
            if (M == 1) {
                const double sg = (vi == vj) ? 1.0 : -1.0;
                for (int r = 0; r < 9; ++r) blk[r] = sg * T.H[9 * (r / 3) + r % 3];
            } else {
                for (int r = 0; r < 9; ++r) blk[r] = 0.0;
                for (int a = 0; a < M; ++a) {
                    const double ba = B[a * NV + vi];
                    for (int b = 0; b < M; ++b) {
                        const double sc = ba * B[b * NV + vj];
                        for (int r = 0; r < 3; ++r)
                            for (int c2 = 0; c2 < 3; ++c2) blk[3 * r + c2] += sc * T.H[9 * (3 * a + r) + 3 * b + c2];
                    }
                }
            }
            if (mode == 0) {
                for (int r = 0; r < 3; ++r)
                    for (int c2 = 0; c2 < 3; ++c2) o[(3 * vi + r) * (3 * NV) + 3 * vj + c2] = blk[3 * r + c2];
            } else {
                for (int r = 0; r < 9; ++r) o[9 * (vi * NV + vj) + r] = blk[r];
            }
        }
        if (mode == 0)
            for (int j = 9 * NV * NV + tl; j < 144; j += TL) o[j] = 0.0;
    }
    if (counts) {
        const int lane = threadIdx.x & 63;
        const int kind = (tl == 0 && ok == 1) ? T.kind : -1;
        const bool cl = tl == 0 && ok == 1 && T.clamped;
        for (int q = 0; q < 7; ++q) {
            const unsigned long long m = __ballot(kind == q), mc = __ballot(kind == q && cl);
            if (m && lane == __ffsll((long long)m) - 1) {
                atomicAdd(counts + 2 * q, __popcll(m));
                if (mc) atomicAdd(counts + 2 * q + 1, __popcll(mc));
            }
        }
    }
    team_sync();  // T is reused by the team's next constraint
}

template <int TL>
__global__ void __launch_bounds__(HB) k_barrier_hess(long n, const unsigned* __restrict__ kcnt,
                                                     const int* __restrict__ perm, const int4* __restrict__ cs,
                                                     const double* __restrict__ F,
                                                     const unsigned char* __restrict__ ok_in, int project_spd,
                                                     int mode, double* __restrict__ out, int* __restrict__ vid,
                                                     int* __restrict__ counts)
{
#pragma clang fp contract(fast)
    constexpr int TEAMS = HB / TL;
    __shared__ TeamLds lds[TEAMS];
    const int team = threadIdx.x / TL, tl = threadIdx.x % TL;
    TeamLds& T = lds[team];
    // with the kinds grouped (kcnt), this kernel takes the order-9 group;
    // grid-stride, the same number of passes for every team
    const long hi = kcnt ? (long)group_range(kcnt, 9).y : n;
    const long stride = (long)gridDim.x * TEAMS;
    const long iters = (hi + stride - 1) / stride;
    for (long it = 0; it < iters; ++it) {
        const long t = it * stride + (long)blockIdx.x * TEAMS + team;
        team_body<TL>(T, tl, t < hi ? t : n, n, perm, cs, F, ok_in, project_spd, mode, out, vid, counts);
    }
}

// ---------------------------------------------------------------------------
// Stage 2, one constraint per lane: cyclic Jacobi on the packed K (order N =
// 9, 6 or 3) with the eigenvectors, both in registers (every index a
// compile-time constant), then the clamp, C^-1 K+ C^-1 and the vertex blocks.
// Each rotation is the classical one (Golub & Van Loan 8.4): t from
// theta = (aqq - app) / (2 apq), app - t apq, aqq + t apq, apq = 0; sweeps run
// in row order (p < q) until off(K)^2 <= 1e-34 sum K_ii^2 (at most 16).
// ---------------------------------------------------------------------------
// projected H_y of one constraint (packed, order N = 3 M) from the form
// stage's record Fi (packed K, then packed H_y); returns whether anything
// was clamped
template <int N>
__device__ inline bool lane_project(const double* __restrict__ Fi, const double* Si, int project_spd,
                                    double (&H)[N * (N + 1) / 2])
{
#pragma clang fp contract(fast)
    constexpr int P = N * (N + 1) / 2, M = N / 3;
    if (project_spd) {
        double K[P], V[N * N];
#pragma unroll
        for (int q = 0; q < P; ++q) K[q] = Fi[q];
        lane_jacobi<N>(K, V);
        double lmin = K[0];
#pragma unroll
        for (int i = 1; i < N; ++i) lmin = fmin(lmin, K[sx<N>(i, i)]);
        if (lmin < 0.0) {  // UTILS.h:12-15: clamp only when an eigenvalue is negative
            double L[N];
#pragma unroll
            for (int i = 0; i < N; ++i) L[i] = fmax(K[sx<N>(i, i)], 0.0);
            double Kp[P];
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = i; j < N; ++j) {
                    double acc = 0.0;
#pragma unroll
                    for (int k = 0; k < N; ++k) acc += V[N * i + k] * L[k] * V[N * j + k];
                    Kp[sx<N>(i, j)] = acc;
                }
            // H_y = C^-1 K+ C^-1 with C^-1 = Si (x) I3 (order 3: C = sqrt(2) I)
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
                                acc = 0.5 * Kp[sx<N>(i, j)];
                            } else {
#pragma unroll
                                for (int gg = 0; gg < M; ++gg)
#pragma unroll
                                    for (int hh = 0; hh < M; ++hh)
                                        acc += Si[a * M + gg] * Kp[sx<N>(3 * gg + i, 3 * hh + j)] * Si[hh * M + b];
                            }
                            H[sx<N>(3 * a + i, 3 * b + j)] = acc;
                        }
            return true;
        }
    }
#pragma unroll
    for (int q = 0; q < P; ++q) H[q] = Fi[45 + q];
    return false;
}

// vertex blocks of B^T H_y B into o (mode 0: dense (3 NV)^2 row-major in 144
// slots, zero filled; mode 1: block (a, b) at 9 (a NV + b))
template <int N, bool PT>
__device__ inline void lane_blocks(const double (&H)[N * (N + 1) / 2], int mode, double* __restrict__ o)
{
#pragma clang fp contract(fast)
    constexpr int M = N / 3, NV = M + 1;
    constexpr const double* B = (M == 3) ? (PT ? kBpt : kBee) : kBpe;
#pragma unroll
    for (int vi = 0; vi < NV; ++vi)
#pragma unroll
        for (int vj = 0; vj < NV; ++vj)
#pragma unroll
            for (int r = 0; r < 3; ++r)
#pragma unroll
                for (int c2 = 0; c2 < 3; ++c2) {
                    double acc = 0.0;
                    if (M == 1) {
                        acc = (vi == vj ? 1.0 : -1.0) * H[sx<N>(r, c2)];
                    } else {
#pragma unroll
                        for (int a = 0; a < M; ++a)
#pragma unroll
                            for (int b = 0; b < M; ++b) {
                                const double sc = B[a * NV + vi] * B[b * NV + vj];
                                if (sc != 0.0) acc += sc * H[sx<N>(3 * a + r, 3 * b + c2)];
                            }
                    }
                    if (mode == 0) o[(3 * vi + r) * (3 * NV) + 3 * vj + c2] = acc;
                    else o[9 * (vi * NV + vj) + 3 * r + c2] = acc;
                }
    if (mode == 0)
#pragma unroll
        for (int j = 9 * NV * NV; j < 144; ++j) o[j] = 0.0;
}

template <int N>
__device__ inline bool lane_hess(int kind, const double* __restrict__ Fi, int project_spd, int mode,
                                 double* __restrict__ o)
{
    constexpr int M = N / 3;
    const double* Si = (M == 3) ? ((kind == K_PT) ? c_Si3pt : c_Si3ee) : c_Si2;
    double H[N * (N + 1) / 2];
    const bool clamped = lane_project<N>(Fi, Si, project_spd, H);
    if (M == 3 && kind == K_PT) lane_blocks<N, true>(H, mode, o);
    else lane_blocks<N, false>(H, mode, o);
    return clamped;
}

// Stage 2 for the constraints of reduced order N (kinds grouped: kcnt gives
// the group's range), one per lane, grid-stride; an idle lane still takes part
// in the wavefront's ballots.
template <int N>
__device__ inline void lane_body(long t, long n, const int* __restrict__ perm, const int4* __restrict__ cs,
                                 const double* __restrict__ F, const unsigned char* __restrict__ ok_in,
                                 int project_spd, int mode, double* __restrict__ out, int* __restrict__ vid,
                                 int* __restrict__ counts)
{
    int kind = -1, ok = 2;
    bool clamped = false;
    if (t < n) {
        const long i = perm[t];
        const Decoded k = decode(cs[i]);
        ok = ok_in[t];
        if (vid && ok != 2)
            for (int j = 0; j < 4; ++j) vid[4 * i + j] = (j < k.nv) ? k.v[j] : -1;
        double* o = out + 144 * i;
        if (ok == 0) {
            for (int j = 0; j < 144; ++j) o[j] = 0.0;
        } else if (ok == 1) {
            kind = k.kind;
            clamped = lane_hess<N>(kind, F + 90 * t, project_spd, mode, o);
        }
    }
    if (counts) {
        const int lane = threadIdx.x & 63;
        for (int q = 0; q < 7; ++q) {
            const unsigned long long m = __ballot(kind == q), mc = __ballot(kind == q && clamped);
            if (m && lane == __ffsll((long long)m) - 1) {
                atomicAdd(counts + 2 * q, __popcll(m));
                if (mc) atomicAdd(counts + 2 * q + 1, __popcll(mc));
            }
        }
    }
}

template <int N>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(N == 3 ? 4 : 2, 8)))
k_barrier_hess_lane(long n, const unsigned* __restrict__ kcnt, const int* __restrict__ perm,
                    const int4* __restrict__ cs, const double* __restrict__ F, const unsigned char* __restrict__ ok_in,
                    int project_spd, int mode, double* __restrict__ out, int* __restrict__ vid,
                    int* __restrict__ counts)
{
    const int2 r = group_range(kcnt, N);
    const long stride = (long)gridDim.x * blockDim.x;
    const long span = r.y - r.x;
    const long iters = (span + stride - 1) / stride;  // the same for every lane
    for (long it = 0; it < iters; ++it) {
        const long t = r.x + it * stride + (long)blockIdx.x * blockDim.x + threadIdx.x;
        lane_body<N>(t < r.y ? t : n, n, perm, cs, F, ok_in, project_spd, mode, out, vid, counts);
    }
}

// ---------------------------------------------------------------------------
// Order 9 (PT, EE and the mollified kinds): three lanes per constraint, lane
// r = 0, 1, 2 of a triple owning coordinate component r.  Every lane of the
// triple runs the same cyclic Jacobi on the whole packed K (identical
// arithmetic, so the three copies stay equal and the triple stays in step)
// and rotates only the eigenvector rows (a, r), a = 0..2, of its component.
// U = C^-1 V mixes rows of one component only, so each lane forms
// Z_r[v][k] = sum_a B[a][v] U[(a, r)][k] for the 4 vertices; the vertex
// blocks are sum_k max(L_k, 0) Z_r[vi][k] Z_c[vj][k], with Z_c from lane c of
// the triple by shuffles.  Lane r writes the entries of row component r.
// 21 triples per wavefront (lane 63 idles).
// ---------------------------------------------------------------------------
constexpr int TRI_PER_WAVE = 21;

__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 8)))
k_barrier_hess_tri(long n, const unsigned* __restrict__ kcnt, const int* __restrict__ perm,
                   const int4* __restrict__ cs, const double* __restrict__ F, const unsigned char* __restrict__ ok_in,
                   int project_spd, int mode, double* __restrict__ out, int* __restrict__ vid,
                   int* __restrict__ counts)
{
    const int2 rg = group_range(kcnt, 9);
    const int wl = threadIdx.x & 63, g = wl / 3, r = wl % 3, base = 3 * g;
    const long waves = (long)gridDim.x * (blockDim.x / 64);
    const long stride = waves * TRI_PER_WAVE;
    const long wave_id = (long)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const long span = rg.y - rg.x;
    const long iters = (span + stride - 1) / stride;
    for (long it = 0; it < iters; ++it) {
        const long t = rg.x + it * stride + wave_id * TRI_PER_WAVE + g;
        int kind = -1, ok = 2;
        bool clamped = false;
        if (g < TRI_PER_WAVE && t < rg.y) {
            const long i = perm[t];
            const Decoded k = decode(cs[i]);
            ok = ok_in[t];
            if (r == 0 && vid && ok != 2)
                for (int j = 0; j < 4; ++j) vid[4 * i + j] = (j < k.nv) ? k.v[j] : -1;
            double* o = out + 144 * i;
            if (ok == 0) {
                if (r == 0)
                    for (int j = 0; j < 144; ++j) o[j] = 0.0;
            } else if (ok == 1) {
                kind = k.kind;
                const double* Fi = F + 90 * t;
                const double* Si = (kind == K_PT) ? c_Si3pt : c_Si3ee;
                double L[9], V[27];
                if (project_spd) {
                    double K[45];
#pragma unroll
                    for (int q = 0; q < 45; ++q) K[q] = Fi[q];
                    tri_jacobi(K, V, r);
                    double lmin = K[0];
#pragma unroll
                    for (int q = 1; q < 9; ++q) lmin = fmin(lmin, K[sx<9>(q, q)]);
                    clamped = lmin < 0.0;  // UTILS.h:12-15
#pragma unroll
                    for (int q = 0; q < 9; ++q) L[q] = fmax(K[sx<9>(q, q)], 0.0);
                }
                // unclamped: the packed H_y of the record (Fi + 45) is the result
                if (kind == K_PT) tri_blocks<true>(r, base, mode, clamped, L, V, Si, Fi + 45, o);
                else tri_blocks<false>(r, base, mode, clamped, L, V, Si, Fi + 45, o);
            }
        }
        if (counts) {
            const bool lead = r == 0;
            for (int q = 0; q < 7; ++q) {
                const unsigned long long m = __ballot(lead && kind == q), mc = __ballot(lead && kind == q && clamped);
                if (m && wl == __ffsll((long long)m) - 1) {
                    atomicAdd(counts + 2 * q, __popcll(m));
                    if (mc) atomicAdd(counts + 2 * q + 1, __popcll(mc));
                }
            }
        }
    }
}

// position of a kind in the grouped order (reduced order 9, then 6, then 3)
__device__ inline int kind_rank(int kind)
{
    return kind == K_PE ? 5 : (kind == K_PP ? 6 : (kind <= K_EE ? kind : kind - 2));
}

// Constraint indices grouped by kind (a counting sort, stable, no atomics):
// k_kind_count writes each block's count per kind (bcnt[8 b + q]);
// k_kind_scan turns them into each block's offset inside its kind (boff) and
// the totals (cnt[0..7)); k_kind_scatter places every index at its group's
// offset (kinds in kind_rank order) + its block's offset + its rank among the
// block's lanes of that kind.  The Hessian writes by constraint index, so the
// order only groups the work.
constexpr int KIND_BLOCK = 256;

__device__ inline void block_kind_counts(int kind, unsigned (&wc)[KIND_BLOCK / 64][8])
{
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const unsigned long long m = __ballot(kind == q);
        if (lane == 0) wc[w][q] = (unsigned)__popcll(m);
    }
}

__global__ void __launch_bounds__(KIND_BLOCK) k_kind_count(long n, const int4* __restrict__ cs,
                                                           unsigned* __restrict__ bcnt)
{
    __shared__ unsigned wc[KIND_BLOCK / 64][8];
    const long i = (long)blockIdx.x * KIND_BLOCK + threadIdx.x;
    block_kind_counts(i < n ? decode(cs[i]).kind : -1, wc);
    __syncthreads();
    if (threadIdx.x < 7) {
        unsigned c = 0;
        for (int w = 0; w < KIND_BLOCK / 64; ++w) c += wc[w][threadIdx.x];
        bcnt[8 * (long)blockIdx.x + threadIdx.x] = c;
    }
}

// one workgroup: exclusive scan over the blocks for every kind at once;
// thread t sums a contiguous run of blocks, the 256 run totals are scanned in
// LDS (one Hillis-Steele pass for the 7 kinds), and each thread then writes
// its run's offsets
__global__ void __launch_bounds__(256) k_kind_scan(int nb, const unsigned* __restrict__ bcnt,
                                                   unsigned* __restrict__ boff, unsigned* __restrict__ cnt)
{
    __shared__ unsigned sh[7][256];
    const int per = (nb + 255) / 256;
    const int b0 = min(nb, (int)threadIdx.x * per), b1 = min(nb, b0 + per);
    unsigned run[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) run[q] = 0;
    for (int b = b0; b < b1; ++b)
#pragma unroll
        for (int q = 0; q < 7; ++q) run[q] += bcnt[8 * (long)b + q];
#pragma unroll
    for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] = run[q];
    __syncthreads();
    for (int d = 1; d < 256; d <<= 1) {  // inclusive scan of the run totals
        unsigned a[7];
#pragma unroll
        for (int q = 0; q < 7; ++q) a[q] = threadIdx.x >= (unsigned)d ? sh[q][threadIdx.x - d] : 0u;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 7; ++q) sh[q][threadIdx.x] += a[q];
        __syncthreads();
    }
    unsigned acc[7];
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] = sh[q][threadIdx.x] - run[q];
    for (int b = b0; b < b1; ++b)
#pragma unroll
        for (int q = 0; q < 7; ++q) {
            boff[8 * (long)b + q] = acc[q];
            acc[q] += bcnt[8 * (long)b + q];
        }
    if (threadIdx.x < 7) cnt[threadIdx.x] = sh[threadIdx.x][255];
}

__global__ void __launch_bounds__(KIND_BLOCK) k_kind_scatter(long n, const int4* __restrict__ cs,
                                                             const unsigned* __restrict__ boff,
                                                             const unsigned* __restrict__ cnt, int* __restrict__ perm)
{
    __shared__ unsigned wc[KIND_BLOCK / 64][8];
    const long i = (long)blockIdx.x * KIND_BLOCK + threadIdx.x;
    const int kind = i < n ? decode(cs[i]).kind : -1;
    block_kind_counts(kind, wc);
    __syncthreads();
    if (kind < 0) return;
    const int lane = threadIdx.x & 63, w = threadIdx.x / 64;
    // groups in the order PT, EE, MEE, MPE, MPP (reduced order 9), PE (6),
    // PP (3): the lanes of a wavefront then share the projection's size
    unsigned off = 0;
    for (int q = 0; q < 7; ++q) off += (kind_rank(q) < kind_rank(kind)) ? cnt[q] : 0u;
    off += boff[8 * (long)blockIdx.x + kind];
    for (int v = 0; v < w; ++v) off += wc[v][kind];
    unsigned long long mine = 0;  // lanes of this wavefront with this lane's kind
#pragma unroll
    for (int q = 0; q < 7; ++q) {
        const unsigned long long m = __ballot(kind == q);
        if (kind == q) mine = m;
    }
    perm[off + (unsigned)__popcll(mine & ((1ull << lane) - 1))] = (int)i;
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
                         double* form, unsigned char* form_ok, int* counts, int* perm, unsigned* kcnt, int team)
{
    if (n <= 0) return;
    if (perm && kcnt) {
        const int nb = grid_for(n, KIND_BLOCK);
        unsigned* bcnt = kcnt + 16;
        unsigned* boff = bcnt + 8 * (size_t)nb;
        IDP_LAUNCH(k_kind_count, nb, KIND_BLOCK, 0, s, n, (const int4*)cs, bcnt);
        IDP_LAUNCH(k_kind_scan, 1, 256, 0, s, nb, bcnt, boff, kcnt);
        IDP_LAUNCH(k_kind_scatter, nb, KIND_BLOCK, 0, s, n, (const int4*)cs, boff, kcnt, perm);
    } else {
        perm = nullptr;
    }
    IDP_LAUNCH(k_barrier_form, grid_for(n, 256), 256, 0, s, n, perm, (const int4*)cs, X, X0, bp, form, form_ok, err);
    // grouped: order 9 on teams of `team` lanes, orders 6 and 3 one per lane;
    // every kernel walks its group's range (read on the device) grid-stride
    const unsigned* cnt = perm ? kcnt : nullptr;
    const int gl = (int)std::min<long>(grid_for(n, 256), 2048);
    if (cnt) {
        IDP_LAUNCH(k_barrier_hess_lane<6>, gl, 256, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok, project_spd,
                   mode, out, vid, counts);
        IDP_LAUNCH(k_barrier_hess_lane<3>, gl, 256, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok, project_spd,
                   mode, out, vid, counts);
    }
    if (cnt && team <= 3) {  // order 9 on lane triples
        const int gt = (int)std::min<long>(grid_for(n, 4 * TRI_PER_WAVE), 4096);
        IDP_LAUNCH(k_barrier_hess_tri, gt, 256, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok, project_spd,
                   mode, out, vid, counts);
        return;
    }
    auto team_grid = [&](int lanes) { return (int)std::min<long>(grid_for(n, HB / lanes), 16384); };
    switch (team) {
    case 8:
        IDP_LAUNCH(k_barrier_hess<8>, team_grid(8), HB, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok,
                   project_spd, mode, out, vid, counts);
        break;
    case 32:
        IDP_LAUNCH(k_barrier_hess<32>, team_grid(32), HB, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok,
                   project_spd, mode, out, vid, counts);
        break;
    case 64:
        IDP_LAUNCH(k_barrier_hess<64>, team_grid(64), HB, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok,
                   project_spd, mode, out, vid, counts);
        break;
    default:
        IDP_LAUNCH(k_barrier_hess<16>, team_grid(16), HB, 0, s, n, cnt, perm, (const int4*)cs, form, form_ok,
                   project_spd, mode, out, vid, counts);
        break;
    }
}

void launch_constraint_vids(hipStream_t s, long n, const int* cs, int* vid, int* nv)
{
    IDP_LAUNCH(k_constraint_vids, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, vid, nv);
}

}  // namespace idp
