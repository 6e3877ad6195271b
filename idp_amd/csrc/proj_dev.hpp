// Device helpers of the PSD projections in reduced coordinates, shared by the
// barrier Hessian and the shell elasticity Hessians: packed symmetric
// storage, cyclic Jacobi one matrix per lane (order 6 or 3) and on lane
// triples (order 9), and the difference maps B of the reduced coordinates.
#pragma once
#include <hip/hip_runtime.h>

namespace idp {

// Symmetric matrices are held as packed upper triangles: entry (i, j) of an
// N x N matrix at sx<N>(i, j).  All loops over entries are fully unrolled, so
// every index is a compile-time constant and the arrays stay in registers.
template <int N>
__device__ constexpr int sx(int i, int j)
{
    return (i <= j) ? (i * (2 * N - i + 1)) / 2 + (j - i) : (j * (2 * N - j + 1)) / 2 + (i - j);
}

// reduced-coordinate maps y = B x per vertex (rows a, columns vertex), the
// same as the constant tables of barrier_init_constants
__device__ constexpr double kBpt[12] = {1, -1, 0, 0, 0, -1, 1, 0, 0, -1, 0, 1};
__device__ constexpr double kBee[12] = {-1, 0, 1, 0, -1, 1, 0, 0, 0, 0, -1, 1};
__device__ constexpr double kBpe[6] = {1, -1, 0, 0, -1, 1};

template <int N>
__device__ inline void lane_jacobi(double (&K)[N * (N + 1) / 2], double (&V)[N * N])
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int i = 0; i < N; ++i)
#pragma unroll
        for (int j = 0; j < N; ++j) V[N * i + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < N; ++i) {
            dg += K[sx<N>(i, i)] * K[sx<N>(i, i)];
#pragma unroll
            for (int j = i + 1; j < N; ++j) off += K[sx<N>(i, j)] * K[sx<N>(i, j)];
        }
        if (!(off > 1e-34 * dg)) break;
#pragma unroll
        for (int p = 0; p < N - 1; ++p)
#pragma unroll
            for (int q = p + 1; q < N; ++q) {
                const double apq = K[sx<N>(p, q)];
                const double dd = K[sx<N>(q, q)] - K[sx<N>(p, p)];
                const double rt = sqrt(dd * dd + 4.0 * apq * apq);
                // apq == 0: the identity (t = 0), also when dd == 0; no branch, so
                // the whole sweep is one block and its values stay in registers
                const double den = fabs(dd) + rt;
                const double t = copysign(1.0, dd) * (2.0 * apq) / (den != 0.0 ? den : 1.0);
                const double c = rsqrt(t * t + 1.0);
                const double sn = t * c;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    if (k == p || k == q) continue;
                    const double kp = K[sx<N>(k, p)], kq = K[sx<N>(k, q)];
                    K[sx<N>(k, p)] = c * kp - sn * kq;
                    K[sx<N>(k, q)] = sn * kp + c * kq;
                }
                K[sx<N>(p, p)] -= t * apq;
                K[sx<N>(q, q)] += t * apq;
                K[sx<N>(p, q)] = 0.0;
#pragma unroll
                for (int k = 0; k < N; ++k) {
                    const double vp = V[N * k + p], vq = V[N * k + q];
                    V[N * k + p] = c * vp - sn * vq;
                    V[N * k + q] = sn * vp + c * vq;
                }
            }
    }
}

// Order 9 on a lane triple (lane r = 0, 1, 2 owns coordinate component r):
// every lane of the triple runs the same cyclic Jacobi on the whole packed K
// and rotates only the eigenvector rows (a, r), a = 0..2, of its component.
__device__ inline void tri_jacobi(double (&K)[45], double (&V)[27], int r)
{
#pragma clang fp contract(fast)
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int k = 0; k < 9; ++k) V[9 * a + k] = (3 * a + r == k) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 16; ++sweep) {
        double off = 0.0, dg = 0.0;
#pragma unroll
        for (int i = 0; i < 9; ++i) {
            dg += K[sx<9>(i, i)] * K[sx<9>(i, i)];
#pragma unroll
            for (int j = i + 1; j < 9; ++j) off += K[sx<9>(i, j)] * K[sx<9>(i, j)];
        }
        if (!(off > 1e-34 * dg)) break;
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int q = p + 1; q < 9; ++q) {
                const double apq = K[sx<9>(p, q)];
                const double dd = K[sx<9>(q, q)] - K[sx<9>(p, p)];
                const double rt = sqrt(dd * dd + 4.0 * apq * apq);
                const double den = fabs(dd) + rt;  // as in lane_jacobi: branch free
                const double t = copysign(1.0, dd) * (2.0 * apq) / (den != 0.0 ? den : 1.0);
                const double c = rsqrt(t * t + 1.0);
                const double sn = t * c;
#pragma unroll
                for (int k = 0; k < 9; ++k) {
                    if (k == p || k == q) continue;
                    const double kp = K[sx<9>(k, p)], kq = K[sx<9>(k, q)];
                    K[sx<9>(k, p)] = c * kp - sn * kq;
                    K[sx<9>(k, q)] = sn * kp + c * kq;
                }
                K[sx<9>(p, p)] -= t * apq;
                K[sx<9>(q, q)] += t * apq;
                K[sx<9>(p, q)] = 0.0;
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const double vp = V[9 * a + p], vq = V[9 * a + q];
                    V[9 * a + p] = c * vp - sn * vq;
                    V[9 * a + q] = sn * vp + c * vq;
                }
            }
    }
}


// this lane's 48 entries (row component r) of the 4 x 4 vertex blocks
template <bool PT>
__device__ inline void tri_blocks(int r, int base, int mode, bool clamped, const double (&L)[9],
                                  const double (&V)[27], const double* Si, const double* __restrict__ Hy,
                                  double* __restrict__ o)
{
#pragma clang fp contract(fast)
    constexpr const double* B = PT ? kBpt : kBee;
    if (!clamped) {  // B^T H_y B from the packed H_y of the record
#pragma unroll
        for (int vi = 0; vi < 4; ++vi)
#pragma unroll
            for (int vj = 0; vj < 4; ++vj)
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    double acc = 0.0;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) {
                            const double sc = B[a * 4 + vi] * B[b * 4 + vj];
                            if (sc != 0.0) acc += sc * Hy[sx<9>(3 * a + r, 3 * b + c)];
                        }
                    if (mode == 0) o[(3 * vi + r) * 12 + 3 * vj + c] = acc;
                    else o[9 * (vi * 4 + vj) + 3 * r + c] = acc;
                }
        return;
    }
    double Z[4][9];
#pragma unroll
    for (int k = 0; k < 9; ++k) {
        double u[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) u[a] = Si[3 * a] * V[k] + Si[3 * a + 1] * V[9 + k] + Si[3 * a + 2] * V[18 + k];
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            double acc = 0.0;
#pragma unroll
            for (int a = 0; a < 3; ++a)
                if (B[a * 4 + v] != 0.0) acc += B[a * 4 + v] * u[a];
            Z[v][k] = acc;
        }
    }
#pragma unroll
    for (int sft = 0; sft < 3; ++sft) {
        const int c = (r + sft) % 3;
#pragma unroll
        for (int vj = 0; vj < 4; ++vj) {
            double zc[9];  // row vj of Z_c, from lane c of the triple
#pragma unroll
            for (int k = 0; k < 9; ++k) zc[k] = __shfl(Z[vj][k], base + c);
#pragma unroll
            for (int vi = 0; vi < 4; ++vi) {
                double acc = 0.0;
#pragma unroll
                for (int k = 0; k < 9; ++k) acc += L[k] * (Z[vi][k] * zc[k]);
                if (mode == 0) o[(3 * vi + r) * 12 + 3 * vj + c] = acc;
                else o[9 * (vi * 4 + vj) + 3 * r + c] = acc;
            }
            __builtin_amdgcn_sched_barrier(0);
        }
    }
}

}  // namespace idp
