// Tet Neo-Hookean elasticity and the inertia (target-matching) term.
//   F = Ds * IB            Library/FEM/DEFORMATION_GRADIENT.h:113-143
//   psi, P, dP/dF          Library/Physics/NEOHOOKEAN.h:78-301
//   nodal forces / 12x12   Library/FEM/ELEM_TO_NODE.h:90-277
//   0.5 (x-x~)^T M (x-x~)  Library/FEM/Shell/INC_POTENTIAL.h:131-144
// One tet per lane; results are written per tet and reduced per vertex /
// per BSR block by gather kernels (no atomics).
#include "common.hpp"
#include "kernels.hpp"

namespace idp {

namespace {

// F = Ds * IB with the fused chain of the reference's column product
// (VECTOR.h:307-324), so det F <= 0 decisions match the CPU restatement.
__device__ inline void def_grad(const double* __restrict__ X, int4 t, const double* __restrict__ ib, double* F)
{
    double Ds[9];
    const int v[4] = {t.x, t.y, t.z, t.w};
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i) Ds[3 * i + k] = X[3 * v[k + 1] + i] - X[3 * v[0] + i];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] = fma(Ds[3 * i + 2], ib[6 + j], fma(Ds[3 * i + 1], ib[3 + j], Ds[3 * i] * ib[j]));
}

__device__ inline double det3(const double* m)
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}

__device__ inline void cof3(const double* m, double* c)
{
    c[0] = m[4] * m[8] - m[5] * m[7];
    c[1] = m[5] * m[6] - m[3] * m[8];
    c[2] = m[3] * m[7] - m[4] * m[6];
    c[3] = m[2] * m[7] - m[1] * m[8];
    c[4] = m[0] * m[8] - m[2] * m[6];
    c[5] = m[1] * m[6] - m[0] * m[7];
    c[6] = m[1] * m[5] - m[2] * m[4];
    c[7] = m[2] * m[3] - m[0] * m[5];
    c[8] = m[0] * m[4] - m[1] * m[3];
}

// one-sided Jacobi SVD of a 3x3 (F = U diag(s) V^T, U,V rotations when det F > 0)
__device__ inline void svd3(const double* F, double* U, double* s, double* V)
{
#pragma clang fp contract(fast)
    double G[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        G[i] = F[i];
        V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 20; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                double a = 0, b = 0, c = 0;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    a += G[3 * i + p] * G[3 * i + p];
                    b += G[3 * i + q] * G[3 * i + q];
                    c += G[3 * i + p] * G[3 * i + q];
                }
                if (c == 0.0) continue;
                off = fmax(off, fabs(c) / sqrt(a * b));
                double z = (b - a) / (2 * c);
                double t = copysign(1.0, z) / (fabs(z) + sqrt(1 + z * z));
                double cs = 1 / sqrt(1 + t * t), sn = cs * t;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double gp = G[3 * i + p], gq = G[3 * i + q];
                    G[3 * i + p] = cs * gp - sn * gq;
                    G[3 * i + q] = sn * gp + cs * gq;
                    double vp = V[3 * i + p], vq = V[3 * i + q];
                    V[3 * i + p] = cs * vp - sn * vq;
                    V[3 * i + q] = sn * vp + cs * vq;
                }
            }
        if (off < 1e-16) break;
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double n = sqrt(G[c] * G[c] + G[3 + c] * G[3 + c] + G[6 + c] * G[6 + c]);
        s[c] = n;
        double inv = n > 0 ? 1.0 / n : 0.0;
#pragma unroll
        for (int r = 0; r < 3; ++r) U[3 * r + c] = G[3 * r + c] * inv;
    }
    // rotation-variant convention: det V = +1, U = rotation, sign on s[2]
    if (det3(V) < 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) V[3 * r + 2] = -V[3 * r + 2];
#pragma unroll
        for (int r = 0; r < 3; ++r) U[3 * r + 2] = -U[3 * r + 2];
    }
    if (det3(U) < 0) {
#pragma unroll
        for (int r = 0; r < 3; ++r) U[3 * r + 2] = -U[3 * r + 2];
        s[2] = -s[2];
    }
}

__device__ inline void make_pd2(double* B)  // 2x2 [[l+r, l-r],[l-r, l+r]] has eigen (2l, 2r)
{
    double l = 0.5 * (B[0] + B[1]), r = 0.5 * (B[0] - B[1]);
    double e1 = fmax(2 * l, 0.0), e2 = fmax(2 * r, 0.0);
    B[0] = B[3] = 0.5 * (e1 + e2);
    B[1] = B[2] = 0.5 * (e1 - e2);
}

__device__ inline void make_pd3(double* A)
{
#pragma clang fp contract(fast)
    double a[9], w[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        a[i] = A[i];
        w[i] = (i % 4 == 0) ? 1.0 : 0.0;
    }
    for (int sweep = 0; sweep < 20; ++sweep) {
        double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        double dg = a[0] * a[0] + a[4] * a[4] + a[8] * a[8];
        if (!(off > 1e-34 * dg)) break;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int q = p + 1; q < 3; ++q) {
                double apq = a[3 * p + q];
                if (apq == 0.0) continue;
                double th = (a[3 * q + q] - a[3 * p + p]) / (2 * apq);
                double t = copysign(1.0, th) / (fabs(th) + sqrt(th * th + 1));
                double c = 1 / sqrt(t * t + 1), s = t * c;
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double x = a[3 * k + p], y = a[3 * k + q];
                    a[3 * k + p] = c * x - s * y;
                    a[3 * k + q] = s * x + c * y;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double x = a[3 * p + k], y = a[3 * q + k];
                    a[3 * p + k] = c * x - s * y;
                    a[3 * q + k] = s * x + c * y;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    double x = w[3 * k + p], y = w[3 * k + q];
                    w[3 * k + p] = c * x - s * y;
                    w[3 * k + q] = s * x + c * y;
                }
            }
    }
    double l[3] = {a[0], a[4], a[8]};
    if (l[0] >= 0 && l[1] >= 0 && l[2] >= 0) return;
#pragma unroll
    for (int k = 0; k < 3; ++k) l[k] = fmax(l[k], 0.0);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) A[3 * i + j] = w[3 * i] * l[0] * w[3 * j] + w[3 * i + 1] * l[1] * w[3 * j + 1] +
                                                   w[3 * i + 2] * l[2] * w[3 * j + 2];
}

}  // namespace

// psi per tet (w * vol * psi), reduced per block; invalid flag if det F <= 0
__global__ void __launch_bounds__(RB) k_tet_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                   const int* __restrict__ binst, const int4* __restrict__ tets,
                                                   const double* __restrict__ IB, const double* __restrict__ vol,
                                                   const double* __restrict__ lam, const double* __restrict__ mu,
                                                   const double* __restrict__ X, double w,
                                                   double* __restrict__ part, int* __restrict__ invalid)
{
    __shared__ double sh[RB];
    const int t = lo[blockIdx.x] + threadIdx.x;
    double e = 0.0;
    if (t < hi[blockIdx.x]) {
        double F[9];
        def_grad(X, tets[t], IB + 9 * (size_t)t, F);
        double J = det3(F);
        if (J <= 0) {
            atomicOr(invalid + binst[blockIdx.x], 1);
        } else {
#pragma clang fp contract(fast)
            double s2 = 0.0;
#pragma unroll
            for (int k = 0; k < 9; ++k) s2 += F[k] * F[k];
            double lj = log(J);
            double m = mu[t], l = lam[t];
            e = (m / 2.0 * (s2 - 3) - (m - l / 2.0 * lj) * lj) * (w * vol[t]);
        }
    }
    double s = block_sum<RB>(e, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// nodal forces per tet: gl[12 t + 3 a + i]
__global__ void __launch_bounds__(256) k_tet_grad(int nT, const int4* __restrict__ tets,
                                                  const double* __restrict__ IB, const double* __restrict__ vol,
                                                  const double* __restrict__ lam, const double* __restrict__ mu,
                                                  const double* __restrict__ X, double w, const int* __restrict__ vinst,
                                                  const unsigned char* __restrict__ act, double* __restrict__ gl)
{
#pragma clang fp contract(fast)
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int4 tt = tets[t];
    if (act && !act[vinst[tt.x]]) return;  // instance not iterating
    double F[9], C[9], P[9];
    const double* ib = IB + 9 * (size_t)t;
    def_grad(X, tt, ib, F);
    double J = det3(F);
    cof3(F, C);
    double lj = log(J), iJ = 1.0 / J, wv = w * vol[t], m = mu[t], l = lam[t];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double fit = C[k] * iJ;
        P[k] = wv * (m * (F[k] - fit) + l * lj * fit);
    }
    double g[12];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int i = 0; i < 3; ++i)
            g[3 * (k + 1) + i] = ib[3 * k] * P[3 * i] + ib[3 * k + 1] * P[3 * i + 1] + ib[3 * k + 2] * P[3 * i + 2];
#pragma unroll
    for (int i = 0; i < 3; ++i) g[i] = -g[3 + i] - g[6 + i] - g[9 + i];
    double* o = gl + 12 * (size_t)t;
#pragma unroll
    for (int k = 0; k < 12; ++k) o[k] = g[k];
}

// projected 12x12 stiffness per tet, written as 16 vertex blocks of 9
// (block (a,b) at 144 t + 9 (4a+b)), or dense row-major when dense != 0
__global__ void __launch_bounds__(128) k_tet_hess(int nT, const int4* __restrict__ tets,
                                                  const double* __restrict__ IB, const double* __restrict__ vol,
                                                  const double* __restrict__ lam, const double* __restrict__ mu,
                                                  const double* __restrict__ X, double w, int project, int dense,
                                                  const int* __restrict__ vinst, const unsigned char* __restrict__ act,
                                                  double* __restrict__ Hl)
{
#pragma clang fp contract(fast)
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nT) return;
    const int4 tt = tets[t];
    if (act && !act[vinst[tt.x]]) return;  // instance not iterating
    double F[9], U[9], s[3], V[9];
    const double* ib = IB + 9 * (size_t)t;
    def_grad(X, tt, ib, F);
    svd3(F, U, s, V);
    const double m = mu[t], l = lam[t], wv = w * vol[t];
    // NEOHOOKEAN.h:154-189: d psi/d sigma, d2 psi/d sigma2, B-block coefficients
    const double lj = log(s[0] * s[1] * s[2]);
    double dE[3], A[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        double inv = 1.0 / s[i];
        dE[i] = m * (s[i] - inv) + l * inv * lj;
        double inv2 = inv * inv;
        A[4 * i] = m * (1.0 + inv2) - l * inv2 * (lj - 1.0);
    }
    A[1] = A[3] = l / (s[0] * s[1]);
    A[5] = A[7] = l / (s[1] * s[2]);
    A[2] = A[6] = l / (s[2] * s[0]);
    if (project) make_pd3(A);
    const double mid = m - l * lj;
    double B[3][4];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const int cp = (c + 1) % 3;
        double left = 0.5 * (m + mid / (s[c] * s[cp]));
        double sum = s[c] + s[cp];
        double right = (dE[c] + dE[cp]) / (2 * (sum < 1e-6 ? 1e-6 : sum));
        B[c][0] = B[c][3] = left + right;
        B[c][1] = B[c][2] = left - right;
        if (project) make_pd2(B[c]);
    }
    // dP/dF = sum M[(a,b),(c,d)] vec(u_a v_b^T) vec(u_c v_d^T)^T, weight wv.
    // The 12x12 block (node α, node β) = D_α^T dPdF D_β with D_k(j) = IB(k-1, j).
    double cw[4][3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        cw[1][j] = ib[j];
        cw[2][j] = ib[3 + j];
        cw[3][j] = ib[6 + j];
        cw[0][j] = -cw[1][j] - cw[2][j] - cw[3][j];
    }
    // Project the node weights onto the V frame: q[node][b] = sum_j cw[node][j] V(j,b)
    double q[4][3];
#pragma unroll
    for (int n = 0; n < 4; ++n)
#pragma unroll
        for (int b = 0; b < 3; ++b) q[n][b] = cw[n][0] * V[b] + cw[n][1] * V[3 + b] + cw[n][2] * V[6 + b];
    // dPdF in the (U,V) frame is sparse: entries M[(a,b),(c,d)] only for
    // (a=b,c=d) [A block] and {(a,b),(b,a)} pairs [B blocks].
    // Block (α,β)_{i,r} = wv * sum_{(a,b),(c,d)} M U(i,a) q[α][b] U(r,c) q[β][d]
    double* o = Hl + 144 * (size_t)t;
#pragma unroll
    for (int al = 0; al < 4; ++al)
#pragma unroll
        for (int be = 0; be < 4; ++be) {
            // G[a][c] = sum over the sparse M couplings of the frame weights
            double G[3][3];
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int c = 0; c < 3; ++c) G[a][c] = A[3 * a + c] * q[al][a] * q[be][c];
            // B blocks couple (a,b)->(a,b) and (a,b)->(b,a)
            // term for M[(a,b),(a,b)] = Bc[0] : U(i,a) q[al][b] U(r,a) q[be][b]
            // term for M[(a,b),(b,a)] = Bc[1] : U(i,a) q[al][b] U(r,b) q[be][a]
            double blk[9];
#pragma unroll
            for (int i = 0; i < 3; ++i)
#pragma unroll
                for (int r = 0; r < 3; ++r) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int c = 0; c < 3; ++c) acc += U[3 * i + a] * G[a][c] * U[3 * r + c];
#pragma unroll
                    for (int c = 0; c < 3; ++c) {
                        const int a = c, b = (c + 1) % 3;  // pair (a,b) of block c
                        const double* Bc = B[c];
                        // basis order of block c is ((a,b),(b,a)) for c = 0,1 and ((b,a),(a,b)) for c = 2
                        // (NEOHOOKEAN.h:258-262); the 2x2 is symmetric with equal diagonals, so only
                        // the diagonal/off-diagonal values matter.
                        const double dgl = Bc[0], odg = Bc[1];
                        acc += dgl * (U[3 * i + a] * q[al][b] * U[3 * r + a] * q[be][b] +
                                      U[3 * i + b] * q[al][a] * U[3 * r + b] * q[be][a]);
                        acc += odg * (U[3 * i + a] * q[al][b] * U[3 * r + b] * q[be][a] +
                                      U[3 * i + b] * q[al][a] * U[3 * r + a] * q[be][b]);
                    }
                    blk[3 * i + r] = wv * acc;
                }
            if (dense) {
#pragma unroll
                for (int i = 0; i < 3; ++i)
#pragma unroll
                    for (int r = 0; r < 3; ++r) o[(3 * al + i) * 12 + 3 * be + r] = blk[3 * i + r];
            } else {
#pragma unroll
                for (int k = 0; k < 9; ++k) o[9 * (4 * al + be) + k] = blk[k];
            }
        }
}

__global__ void __launch_bounds__(RB) k_inertia_energy(const int* __restrict__ lo, const int* __restrict__ hi,
                                                       const double* __restrict__ X, const double* __restrict__ Xt,
                                                       const double* __restrict__ mass, double* __restrict__ part)
{
    __shared__ double sh[RB];
    const int v = lo[blockIdx.x] + threadIdx.x;
    double e = 0.0;
    if (v < hi[blockIdx.x]) {
        double m = mass[v];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            double dx = X[3 * v + d] - Xt[3 * v + d];
            e += (m * dx) * dx;
        }
        e *= 0.5;
    }
    double s = block_sum<RB>(e, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

void launch_tet_energy(hipStream_t s, const MeshDev& m, const SegDev& seg, const double* X, double w, double* part,
                       int* invalid)
{
    IDP_LAUNCH(k_tet_energy, seg.nblk, RB, 0, s, seg.lo.p, seg.hi.p, seg.inst.p, (const int4*)m.tets.p, m.IB.p,
               m.vol.p, m.lam.p, m.mu.p, X, w, part, invalid);
}
void launch_tet_grad(hipStream_t s, const MeshDev& m, const double* X, double w, double* gl, const unsigned char* act)
{
    IDP_LAUNCH(k_tet_grad, grid_for(m.nT, 256), 256, 0, s, m.nT, (const int4*)m.tets.p, m.IB.p, m.vol.p, m.lam.p,
               m.mu.p, X, w, m.inst.vinst.p, act, gl);
}
void launch_tet_hess(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* Hl,
                     const unsigned char* act)
{
    IDP_LAUNCH(k_tet_hess, grid_for(m.nT, 128), 128, 0, s, m.nT, (const int4*)m.tets.p, m.IB.p, m.vol.p, m.lam.p,
               m.mu.p, X, w, project, 0, m.inst.vinst.p, act, Hl);
}
void launch_tet_hess_dense(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* Hl)
{
    IDP_LAUNCH(k_tet_hess, grid_for(m.nT, 128), 128, 0, s, m.nT, (const int4*)m.tets.p, m.IB.p, m.vol.p, m.lam.p,
               m.mu.p, X, w, project, 1, m.inst.vinst.p, (const unsigned char*)nullptr, Hl);
}
void launch_inertia_energy(hipStream_t s, const SegDev& seg, const double* X, const double* Xt, const double* mass,
                           double* part)
{
    IDP_LAUNCH(k_inertia_energy, seg.nblk, RB, 0, s, seg.lo.p, seg.hi.p, X, Xt, mass, part);
}

}  // namespace idp
