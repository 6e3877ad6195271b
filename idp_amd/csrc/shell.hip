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

#include "kernels.hpp"

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

void launch_normal_flow(hipStream_t s, const MeshDev& m, const double* X, double magnitude, double* b)
{
    IDP_LAUNCH(k_normal_flow, grid_for(m.nV, 256), 256, 0, s, m.nV, m.vt_ptr.p, m.vt.p, m.tris.p, X, m.mass.p,
               magnitude, b);
}

}  // namespace idp
