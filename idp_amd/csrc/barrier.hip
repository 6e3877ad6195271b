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
// K is diagonalised by cyclic Jacobi in registers, one constraint per lane.
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
__global__ void __launch_bounds__(RB) k_barrier_energy(long n, const int4* __restrict__ cs,
                                                       const double* __restrict__ X,
                                                       const double* __restrict__ X0, double dh, double kappa,
                                                       double th2, double* __restrict__ part,
                                                       int* __restrict__ err)
{
    __shared__ double sh[RB];
    long i = (long)blockIdx.x * RB + threadIdx.x;
    double e = 0.0;
    if (i < n) {
        Decoded k = decode(cs[i]);
        double d = constraint_d2(k, X) - th2;
        if (d <= 0) {
            atomicOr(err, 1);
        } else {
            e = barrier_b(d, dh, kappa);
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

__global__ void __launch_bounds__(RB) k_min_dist2(long n, const int4* __restrict__ cs,
                                                  const double* __restrict__ X, double* __restrict__ dist2,
                                                  unsigned long long* __restrict__ dmin)
{
    long i = (long)blockIdx.x * RB + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    double d = constraint_d2(k, X);
    if (dist2) dist2[i] = d;
    // squared distances are >= 0, so their IEEE bit patterns order like integers
    if (d >= 0) atomicMin(dmin, (unsigned long long)__double_as_longlong(d));
    else atomicMin(dmin, 0ull);
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

__device__ inline void skew_add(double* H, int r0, int c0, d3 w, double s)
{
    // H[r0.., c0..] += s * [w]x  ([w]x z = w x z)
    H[(r0 + 0) * 9 + c0 + 1] += -s * w.z;
    H[(r0 + 0) * 9 + c0 + 2] += s * w.y;
    H[(r0 + 1) * 9 + c0 + 0] += s * w.z;
    H[(r0 + 1) * 9 + c0 + 2] += -s * w.x;
    H[(r0 + 2) * 9 + c0 + 0] += -s * w.y;
    H[(r0 + 2) * 9 + c0 + 1] += s * w.x;
}

// q = |u x v|^2 over y = (a,u,v): gradient and Hessian (a-blocks zero)
__device__ inline void q_derivs(d3 u, d3 v, double* gq, double* Hq)
{
#pragma clang fp contract(fast)
    const double uu = sqn(u), vv = sqn(v), uv = dot(u, v);
    gq[0] = gq[1] = gq[2] = 0.0;
    gq[3] = 2 * vv * u.x - 2 * uv * v.x;
    gq[4] = 2 * vv * u.y - 2 * uv * v.y;
    gq[5] = 2 * vv * u.z - 2 * uv * v.z;
    gq[6] = 2 * uu * v.x - 2 * uv * u.x;
    gq[7] = 2 * uu * v.y - 2 * uv * u.y;
    gq[8] = 2 * uu * v.z - 2 * uv * u.z;
    const double uc[3] = {u.x, u.y, u.z}, vc[3] = {v.x, v.y, v.z};
#pragma unroll
    for (int i = 0; i < 81; ++i) Hq[i] = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double id = (i == j) ? 1.0 : 0.0;
            Hq[(3 + i) * 9 + 3 + j] = 2 * vv * id - 2 * vc[i] * vc[j];
            Hq[(6 + i) * 9 + 6 + j] = 2 * uu * id - 2 * uc[i] * uc[j];
            const double huv = 4 * uc[i] * vc[j] - 2 * vc[i] * uc[j] - 2 * uv * id;
            Hq[(3 + i) * 9 + 6 + j] = huv;
            Hq[(6 + j) * 9 + 3 + i] = huv;
        }
}

// plane form d = s^2/q, s = a.(u x v), q = |u x v|^2 over y=(a,u,v)
template <bool WANT_H>
__device__ inline void plane_derivs(d3 a, d3 u, d3 v, double* gd, double* Hd)
{
#pragma clang fp contract(fast)
    double gq[9], Hq[81];
    q_derivs(u, v, gq, Hq);
    const d3 n = cross(u, v);
    const double s = dot(a, n);
    const double q = sqn(n);
    const d3 gsu = cross(v, a), gsv = cross(a, u);
    const double gs[9] = {n.x, n.y, n.z, gsu.x, gsu.y, gsu.z, gsv.x, gsv.y, gsv.z};
    const double iq = 1.0 / q;
    const double c1 = 2 * s * iq;        // d(d)/ds
    const double c2 = -s * s * iq * iq;  // d(d)/dq
#pragma unroll
    for (int i = 0; i < 9; ++i) gd[i] = c1 * gs[i] + c2 * gq[i];
    if (!WANT_H) return;
    const double css = 2 * iq;
    const double csq = -2 * s * iq * iq;
    const double cqq = 2 * s * s * iq * iq * iq;
#pragma unroll
    for (int i = 0; i < 9; ++i)
#pragma unroll
        for (int j = 0; j < 9; ++j)
            Hd[i * 9 + j] = css * gs[i] * gs[j] + csq * (gs[i] * gq[j] + gq[i] * gs[j]) + cqq * gq[i] * gq[j] +
                            c2 * Hq[i * 9 + j];
    // + (2s/q) Hess s : blocks [a,u]=-[v]x [u,a]=[v]x [a,v]=[u]x [v,a]=-[u]x [u,v]=-[a]x [v,u]=[a]x
    skew_add(Hd, 0, 3, v, -c1);
    skew_add(Hd, 3, 0, v, c1);
    skew_add(Hd, 0, 6, u, c1);
    skew_add(Hd, 6, 0, u, -c1);
    skew_add(Hd, 3, 6, a, -c1);
    skew_add(Hd, 6, 3, a, c1);
}

// point-edge form d = |a|^2 - (a.w)^2/|w|^2, blocks IA, IW of a 9-wide layout
template <bool WANT_H, int IA, int IW>
__device__ inline void pe_derivs(d3 a, d3 w, double* gd, double* Hd)
{
#pragma clang fp contract(fast)
    const double c = dot(a, w), ww = sqn(w), iw2 = 1.0 / ww;
    const double ac[3] = {a.x, a.y, a.z}, wc[3] = {w.x, w.y, w.z};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        gd[IA + i] = 2 * ac[i] - 2 * c * iw2 * wc[i];
        gd[IW + i] = -2 * c * iw2 * ac[i] + 2 * c * c * iw2 * iw2 * wc[i];
    }
    if (!WANT_H) return;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double id = (i == j) ? 1.0 : 0.0;
            Hd[(IA + i) * 9 + IA + j] = 2 * id - 2 * iw2 * wc[i] * wc[j];
            const double haw = -2 * iw2 * wc[i] * ac[j] + 4 * c * iw2 * iw2 * wc[i] * wc[j] - 2 * c * iw2 * id;
            Hd[(IA + i) * 9 + IW + j] = haw;
            Hd[(IW + j) * 9 + IA + i] = haw;
            Hd[(IW + i) * 9 + IW + j] = -2 * iw2 * ac[i] * ac[j] +
                                        4 * c * iw2 * iw2 * (ac[i] * wc[j] + wc[i] * ac[j]) -
                                        8 * c * c * iw2 * iw2 * iw2 * wc[i] * wc[j] + 2 * c * c * iw2 * iw2 * id;
        }
}

// Gradient g (3M) and, when WANT_H, Hessian H (3M x 3M, ld 9) in y of
// f = mult * b(d - th2) [* e(q)].  Returns false when d - th2 <= 0.
template <int KIND, bool WANT_H>
__device__ inline bool reduced_derivs(const Decoded& k, const double* X, const double* X0, double dh,
                                      double kappa, double th2, double* g, double* H, double& energy)
{
#pragma clang fp contract(fast)
    constexpr int M = KT<KIND>::M;
    constexpr int N = 3 * M;
    double gd[9], Hd[81];
#pragma unroll
    for (int i = 0; i < 9; ++i) gd[i] = 0.0;
    if (WANT_H)
#pragma unroll
        for (int i = 0; i < 81; ++i) Hd[i] = 0.0;
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
    if (KIND == K_PT || KIND == K_EE || KIND == K_MEE) plane_derivs<WANT_H>(y0, y1, y2, gd, Hd);
    else if (KIND == K_PE) pe_derivs<WANT_H, 0, 3>(y0, y1, gd, Hd);
    else if (KIND == K_MPE) pe_derivs<WANT_H, 0, 6>(y0, y2, gd, Hd);
    else {  // PP and mollified PP: d = |a|^2
        gd[0] = 2 * y0.x; gd[1] = 2 * y0.y; gd[2] = 2 * y0.z;
        if (WANT_H) Hd[0] = Hd[10] = Hd[20] = 2.0;
    }
    const double b = barrier_b(d, dh, kappa);
    const double b1 = barrier_db(d, dh, kappa);
    const double b2 = WANT_H ? barrier_d2b(d, dh, kappa) : 0.0;
    if (!KT<KIND>::MOLL) {
        const double m = k.mult;
        energy = (m > 1.0) ? b * m : b;
#pragma unroll
        for (int i = 0; i < N; ++i) g[i] = m * b1 * gd[i];
        if (WANT_H)
#pragma unroll
            for (int i = 0; i < N; ++i)
#pragma unroll
                for (int j = 0; j < N; ++j) H[i * 9 + j] = m * (b2 * gd[i] * gd[j] + b1 * Hd[i * 9 + j]);
        return true;
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
                for (int j = 0; j < 9; ++j) H[i * 9 + j] = b2 * gd[i] * gd[j] + b1 * Hd[i * 9 + j];
        return true;
    }
    const double de = 2.0 / eps * (1.0 - q / eps);
    const double d2e = -2.0 / (eps * eps);
    double gq[9], Hq[81];
    q_derivs(y1, y2, gq, Hq);
#pragma unroll
    for (int i = 0; i < 9; ++i) g[i] = e * b1 * gd[i] + b * de * gq[i];
    if (WANT_H)
#pragma unroll
        for (int i = 0; i < 9; ++i)
#pragma unroll
            for (int j = 0; j < 9; ++j) {
                const double He = de * Hq[i * 9 + j] + d2e * gq[i] * gq[j];
                H[i * 9 + j] = e * b2 * gd[i] * gd[j] + e * b1 * Hd[i * 9 + j] +
                               b1 * de * (gd[i] * gq[j] + gq[i] * gd[j]) + b * He;
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
// Jacobi eigen-decomposition of a symmetric N x N matrix (ld 9) in registers.
// ---------------------------------------------------------------------------
template <int N>
__device__ inline void jacobi(double* A, double* W)
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) W[i * 9 + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dg += A[i * 9 + i] * A[i * 9 + i];
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += A[i * 9 + j] * A[i * 9 + j];
        }
        if (!(off > 1e-34 * dg)) break;
#pragma unroll
        for (int p = 0; p < N - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[p * 9 + q];
                if (apq != 0.0) {
                    const double theta = (A[q * 9 + q] - A[p * 9 + p]) / (2.0 * apq);
                    const double at = fabs(theta);
                    const double t = (at > 1e150) ? 0.5 / theta
                                                  : copysign(1.0, theta) / (at + sqrt(theta * theta + 1.0));
                    const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
                    for (int k2 = 0; k2 < N; ++k2) {
                        if (k2 == p || k2 == q) continue;
                        const double akp = A[k2 * 9 + p], akq = A[k2 * 9 + q];
                        const double np = c * akp - s * akq, nq = s * akp + c * akq;
                        A[k2 * 9 + p] = np;
                        A[p * 9 + k2] = np;
                        A[k2 * 9 + q] = nq;
                        A[q * 9 + k2] = nq;
                    }
                    A[p * 9 + p] = A[p * 9 + p] - t * apq;
                    A[q * 9 + q] = A[q * 9 + q] + t * apq;
                    A[p * 9 + q] = 0.0;
                    A[q * 9 + p] = 0.0;
#pragma unroll
                    for (int k2 = 0; k2 < N; ++k2) {
                        const double wkp = W[k2 * 9 + p], wkq = W[k2 * 9 + q];
                        W[k2 * 9 + p] = c * wkp - s * wkq;
                        W[k2 * 9 + q] = s * wkp + c * wkq;
                    }
                }
            }
    }
}

// makePD in the reduced space; H (ld 9, M blocks) is replaced by its projection.
template <int KIND>
__device__ inline bool project_reduced(double* H)
{
#pragma clang fp contract(fast)
    constexpr int M = KT<KIND>::M;
    constexpr int N = 3 * M;
    const double* S = (M == 3) ? ((KIND == K_PT) ? c_S3pt : c_S3ee) : c_S2;
    const double* Si = (M == 3) ? ((KIND == K_PT) ? c_Si3pt : c_Si3ee) : c_Si2;
    double K[81], W[81];
    // K = C H C, C = S (x) I3  (for M == 1, C = sqrt(2) I)
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
                    if (M == 1) {
                        acc = 2.0 * H[i * 9 + j];
                    } else {
#pragma unroll
                        for (int g = 0; g < M; ++g)
#pragma unroll
                            for (int h = 0; h < M; ++h)
                                acc += S[a * M + g] * H[(3 * g + i) * 9 + 3 * h + j] * S[h * M + b];
                    }
                    K[(3 * a + i) * 9 + 3 * b + j] = acc;
                }
    jacobi<N>(K, W);
    double lmin = K[0];
#pragma unroll
    for (int i = 1; i < N; ++i) lmin = fmin(lmin, K[i * 9 + i]);
    if (lmin >= 0.0) return false;  // UTILS.h:12-15: nothing to clamp
    double lam[N];
#pragma unroll
    for (int i = 0; i < N; ++i) lam[i] = fmax(K[i * 9 + i], 0.0);
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = i; j < N; ++j) {
            double acc = 0.0;
#pragma unroll
            for (int k2 = 0; k2 < N; ++k2) acc += W[i * 9 + k2] * lam[k2] * W[j * 9 + k2];
            K[i * 9 + j] = acc;
            K[j * 9 + i] = acc;
        }
#pragma unroll
    for (int a = 0; a < M; ++a)
#pragma unroll
        for (int b = 0; b < M; ++b)
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int j = 0; j < 3; ++j) {
                    double acc = 0.0;
                    if (M == 1) {
                        acc = 0.5 * K[i * 9 + j];
                    } else {
#pragma unroll
                        for (int g = 0; g < M; ++g)
#pragma unroll
                            for (int h = 0; h < M; ++h)
                                acc += Si[a * M + g] * K[(3 * g + i) * 9 + 3 * h + j] * Si[h * M + b];
                    }
                    H[(3 * a + i) * 9 + 3 * b + j] = acc;
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
            for (int j = 0; j < 3; ++j) blk[3 * i + j] = s * H[i * 9 + j];
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
                for (int j = 0; j < 3; ++j) blk[3 * i + j] += s * H[(3 * a + i) * 9 + 3 * b + j];
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
                                 double th2, int project_spd, int mode, double* out, int* err, int* counts)
{
    constexpr int NV = KT<KIND>::NV;
    double g[9], H[81], e;
    if (!reduced_derivs<KIND, true>(k, X, X0, dh, kappa, th2, g, H, e)) {
        atomicOr(err, 1);
        for (int j = 0; j < 144; ++j) out[j] = 0.0;
        return;
    }
    bool clamped = false;
    if (project_spd) clamped = project_reduced<KIND>(H);
    if (counts) {
        atomicAdd(counts + 2 * KIND, 1);
        if (clamped) atomicAdd(counts + 2 * KIND + 1, 1);
    }
    double blk[9];
#pragma unroll
    for (int a = 0; a < NV; ++a)
#pragma unroll
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
                                                      const double* __restrict__ X0, double dh, double kappa,
                                                      double th2, double* __restrict__ gout, int* __restrict__ err)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    IDP_KIND_SWITCH(k.kind, grad_body, k, X, X0, dh, kappa, th2, gout + 12 * i, err)
}

// Hessian: mode 0 writes the dense (3nv)^2 local matrix (144 per constraint,
// row-major, vertex order of decode); mode 1 writes nv*nv vertex blocks of 9
// (block (a,b) at 9*(a*nv+b)).
__global__ void __launch_bounds__(64) k_barrier_hess(long n, const int4* __restrict__ cs,
                                                     const double* __restrict__ X,
                                                     const double* __restrict__ X0, double dh, double kappa,
                                                     double th2, int project_spd, int mode,
                                                     double* __restrict__ out, int* __restrict__ vid,
                                                     int* __restrict__ err, int* __restrict__ counts)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Decoded k = decode(cs[i]);
    IDP_KIND_SWITCH(k.kind, hess_body, k, X, X0, dh, kappa, th2, project_spd, mode, out + 144 * i, err, counts)
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
void launch_barrier_energy(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                           double kappa, double th2, double* part, int* err)
{
    IDP_LAUNCH(k_barrier_energy, grid_for(n, RB), RB, 0, s, n, (const int4*)cs, X, X0, dh, kappa, th2, part, err);
}
int barrier_energy_blocks(long n) { return grid_for(n, RB); }

void launch_min_dist2(hipStream_t s, long n, const int* cs, const double* X, double* dist2,
                      unsigned long long* dmin)
{
    IDP_LAUNCH(k_min_dist2, grid_for(n, RB), RB, 0, s, n, (const int4*)cs, X, dist2, dmin);
}

void launch_barrier_grad(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                         double kappa, double th2, double* gout, int* err)
{
    IDP_LAUNCH(k_barrier_grad, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, X, X0, dh, kappa, th2, gout, err);
}

void launch_barrier_hess(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                         double kappa, double th2, int project_spd, int mode, double* out, int* vid, int* err,
                         int* counts)
{
    IDP_LAUNCH(k_barrier_hess, grid_for(n, 64), 64, 0, s, n, (const int4*)cs, X, X0, dh, kappa, th2, project_spd,
               mode, out, vid, err, counts);
}

void launch_constraint_vids(hipStream_t s, long n, const int* cs, int* vid, int* nv)
{
    IDP_LAUNCH(k_constraint_vids, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, vid, nv);
}

}  // namespace idp
