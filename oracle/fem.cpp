// Oracle (test infrastructure only): restatement of IDP's tet elasticity and
// inertia terms.
//   rest basis / volume      Library/FEM/DEFORMATION_GRADIENT.h:39-76
//   deformation gradient     Library/FEM/DEFORMATION_GRADIENT.h:113-143
//   Neo-Hookean psi, P, dP/dF Library/Physics/NEOHOOKEAN.h:78-301
//   nodal scatter            Library/FEM/ELEM_TO_NODE.h:90-277
//   lumped mass              Library/FEM/ELEM_TO_NODE.h:11-87
#include "oracle.h"
#include "geom.hpp"

#include <vector>

using namespace orc;

namespace {

// 3x3 column access helpers; matrices stored row-major m[3*i+j] = M(i,j)
inline double det3(const double* m)
{
    // MATRIX::determinant, VECTOR.h:349-354 (expression order kept)
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}
inline void cofactor3(const double* m, double* c)
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

// F = Ds * IB with the fused multiply-add chain of MATRIX::operator*
// (VECTOR.h:307-324: column d = Ds.col0*IB(0,d) then fmadd col1, col2).
inline void deformation_gradient(const double* X, const int* t, const double* IB, double* F)
{
    double Ds[9];
    for (int k = 0; k < 3; ++k)
        for (int i = 0; i < 3; ++i) Ds[3 * i + k] = X[3 * t[k + 1] + i] - X[3 * t[0] + i];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            F[3 * i + j] =
                std::fma(Ds[3 * i + 2], IB[6 + j], std::fma(Ds[3 * i + 1], IB[3 + j], Ds[3 * i + 0] * IB[j]));
}

// NEOHOOKEAN.h:136-142
inline double psi_sigma(const double s[3], double mu, double lam)
{
    double s2 = s[0] * s[0] + s[1] * s[1] + s[2] * s[2];
    double lj = std::log(s[0] * s[1] * s[2]);
    return mu / 2.0 * (s2 - 3) - (mu - lam / 2.0 * lj) * lj;
}

// SVD through the eigen-decomposition of F^T F (as SINGULAR_VALUE_DECOMPOSITION.h:110-150)
void svd3(const double* F, double* U, double* s, double* V)
{
    double A[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += F[3 * k + i] * F[3 * k + j];
            A[3 * i + j] = a;
        }
    double w[3], Vm[9];
    jacobi_eig(3, A, w, Vm);
    // sort descending
    int idx[3] = {0, 1, 2};
    std::sort(idx, idx + 3, [&](int a, int b) { return w[a] > w[b]; });
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) V[3 * r + c] = Vm[3 * r + idx[c]];
    double lam[3] = {w[idx[0]], w[idx[1]], w[idx[2]]};
    if (det3(V) < 0)
        for (int r = 0; r < 3; ++r) V[3 * r + 2] = -V[3 * r + 2];
    for (int c = 0; c < 3; ++c) s[c] = std::sqrt(std::max(lam[c], 0.0));
    if (det3(F) < 0) s[2] = -s[2];
    double u[3][3];
    for (int c = 0; c < 2; ++c) {
        for (int r = 0; r < 3; ++r) {
            double a = 0;
            for (int k = 0; k < 3; ++k) a += F[3 * r + k] * V[3 * k + c];
            u[c][r] = a;
        }
    }
    double n0 = std::sqrt(u[0][0] * u[0][0] + u[0][1] * u[0][1] + u[0][2] * u[0][2]);
    for (int r = 0; r < 3; ++r) u[0][r] /= n0;
    double d = u[0][0] * u[1][0] + u[0][1] * u[1][1] + u[0][2] * u[1][2];
    for (int r = 0; r < 3; ++r) u[1][r] -= d * u[0][r];
    double n1 = std::sqrt(u[1][0] * u[1][0] + u[1][1] * u[1][1] + u[1][2] * u[1][2]);
    for (int r = 0; r < 3; ++r) u[1][r] /= n1;
    u[2][0] = u[0][1] * u[1][2] - u[0][2] * u[1][1];
    u[2][1] = u[0][2] * u[1][0] - u[0][0] * u[1][2];
    u[2][2] = u[0][0] * u[1][1] - u[0][1] * u[1][0];
    for (int c = 0; c < 3; ++c)
        for (int r = 0; r < 3; ++r) U[3 * r + c] = u[c][r];
}

// Compute_DP_Div_DF, NEOHOOKEAN.h:191-301; out: 9x9, index (j*3+i) for F(i,j)
void dPdF(const double* F, double mu, double lam, double w, bool proj, double* out)
{
    double U[9], s[3], V[9];
    svd3(F, U, s, V);
    double lj = std::log(s[0] * s[1] * s[2]);
    double dE[3], A[9];
    for (int i = 0; i < 3; ++i) {
        double inv = 1.0 / s[i];
        dE[i] = mu * (s[i] - inv) + lam * inv * lj;
        double inv2 = 1.0 / s[i] / s[i];
        A[4 * i] = mu * (1.0 + inv2) - lam * inv2 * (lj - 1.0);
    }
    A[1] = A[3] = lam / s[0] / s[1];
    A[5] = A[7] = lam / s[1] / s[2];
    A[2] = A[6] = lam / s[2] / s[0];
    if (proj) make_pd(3, A);
    double B[3][4];
    const double middle = mu - lam * lj;
    for (int c = 0; c < 3; ++c) {
        int cp = (c + 1) % 3;
        double left = (mu + middle / s[c] / s[cp]) / 2.0;
        double right = dE[c] + dE[cp];
        double sum = s[c] + s[cp];
        const double eps = 1.0e-6;
        right /= (sum < eps) ? 2 * eps : 2 * sum;
        B[c][0] = B[c][3] = left + right;
        B[c][1] = B[c][2] = left - right;
        if (proj) make_pd(2, B[c]);
    }
    // M in the (a,b) = a*3+b basis
    double M[81] = {0};
    for (int a = 0; a < 3; ++a)
        for (int c = 0; c < 3; ++c) M[(a * 4) * 9 + c * 4] = w * A[3 * a + c];
    auto setB = [&](int a, int b, const double* Bc) {
        int i0 = a * 3 + b, i1 = b * 3 + a;
        M[i0 * 9 + i0] = w * Bc[0];
        M[i0 * 9 + i1] = w * Bc[1];
        M[i1 * 9 + i0] = w * Bc[2];
        M[i1 * 9 + i1] = w * Bc[3];
    };
    setB(0, 1, B[0]);
    setB(1, 2, B[1]);
    setB(2, 0, B[2]);
    // Phi column (a,b) = vec(u_a v_b^T) with vec index j*3+i
    double Phi[81];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b)
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) Phi[(j * 3 + i) * 9 + a * 3 + b] = U[3 * i + a] * V[3 * j + b];
    double T[81];
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += Phi[r * 9 + k] * M[k * 9 + c];
            T[r * 9 + c] = acc;
        }
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) {
            double acc = 0;
            for (int k = 0; k < 9; ++k) acc += T[r * 9 + k] * Phi[c * 9 + k];
            out[r * 9 + c] = acc;
        }
}

}  // namespace

// Compute_Vol_And_Inv_Basis + FIXED_COROTATED lambda/mu + lumped mass
extern "C" int orc_fem_init(int nV, const double* X, int nT, const int* tets, double rho, double E, double nu,
                            double* vol, double* IB, double* lam, double* mu, double* mass)
{
    for (int v = 0; v < nV; ++v) mass[v] = 0;
    const double L = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double M = E / (2 * (1 + nu));
    for (int t = 0; t < nT; ++t) {
        const int* e = tets + 4 * t;
        double D[9];
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) D[3 * i + k] = X[3 * e[k + 1] + i] - X[3 * e[0] + i];
        double det = det3(D);
        if (det == 0.0) return ORC_E_DEGENERATE;
        vol[t] = det / 6.0;
        // MATRIX::invert (VECTOR.h:379-405): adjugate / det
        double c[9];
        cofactor3(D, c);
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) IB[9 * t + 3 * i + j] = c[3 * j + i] / det;
        lam[t] = L;
        mu[t] = M;
        double me = vol[t] * rho / 4;
        for (int k = 0; k < 4; ++k) mass[e[k]] += me;
    }
    return 0;
}

// Returns 1 if every tet has det F > 0 (INC_POTENTIAL.h:98-108), 0 otherwise.
extern "C" int orc_elastic_energy(const double* X, int nT, const int* tets, const double* IB, const double* vol,
                                  const double* lam, const double* mu, double w, double* E)
{
    std::vector<double> psi(nT);
    int bad = 0;
#pragma omp parallel for reduction(+ : bad)
    for (int t = 0; t < nT; ++t) {
        double F[9];
        deformation_gradient(X, tets + 4 * t, IB + 9 * t, F);
        if (det3(F) <= 0) {
            ++bad;
            psi[t] = 0;
            continue;
        }
        double U[9], s[3], V[9];
        svd3(F, U, s, V);
        psi[t] = psi_sigma(s, mu[t], lam[t]) * (w * vol[t]);
    }
    if (bad) return 0;
    double s = 0;
    for (int t = 0; t < nT; ++t) s += psi[t];
    *E = s;
    return 1;
}

// P = w vol (mu (F - F^-T) + lam ln J F^-T) scattered with IB (ELEM_TO_NODE.h:90-118, 231-257)
extern "C" void orc_elastic_gradient(const double* X, int nT, const int* tets, const double* IB,
                                     const double* vol, const double* lam, const double* mu, double w, double* g)
{
    std::vector<double> gl(12 * (size_t)nT);
#pragma omp parallel for
    for (int t = 0; t < nT; ++t) {
        double F[9], C[9], P[9];
        deformation_gradient(X, tets + 4 * t, IB + 9 * t, F);
        double J = det3(F);
        cofactor3(F, C);
        double lj = std::log(J);
        for (int k = 0; k < 9; ++k) {
            double finvT = C[k] / J;
            P[k] = w * vol[t] * (mu[t] * (F[k] - finvT) + lam[t] * lj * finvT);
        }
        const double* ib = IB + 9 * t;
        double* o = gl.data() + 12 * (size_t)t;
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i)
                o[3 * (k + 1) + i] = ib[3 * k] * P[3 * i] + ib[3 * k + 1] * P[3 * i + 1] + ib[3 * k + 2] * P[3 * i + 2];
        for (int i = 0; i < 3; ++i) o[i] = -o[3 + i] - o[6 + i] - o[9 + i];
    }
    for (int t = 0; t < nT; ++t)
        for (int k = 0; k < 4; ++k)
            for (int i = 0; i < 3; ++i) g[3 * tets[4 * t + k] + i] += gl[12 * (size_t)t + 3 * k + i];
}

// 12x12 local stiffness per tet, row-major, DOF order (node k, component i) -> 3k+i
extern "C" void orc_elastic_hessian_local(const double* X, int nT, const int* tets, const double* IB,
                                          const double* vol, const double* lam, const double* mu, double w,
                                          int projectSPD, double* H)
{
#pragma omp parallel for
    for (int t = 0; t < nT; ++t) {
        double F[9], D9[81];
        deformation_gradient(X, tets + 4 * t, IB + 9 * t, F);
        dPdF(F, mu[t], lam[t], w * vol[t], projectSPD != 0, D9);
        const double* ib = IB + 9 * t;
        double c[4][3];
        for (int j = 0; j < 3; ++j) {
            c[1][j] = ib[j];
            c[2][j] = ib[3 + j];
            c[3][j] = ib[6 + j];
            c[0][j] = -c[1][j] - c[2][j] - c[3][j];
        }
        double* o = H + 144 * (size_t)t;
        for (int a = 0; a < 4; ++a)
            for (int i = 0; i < 3; ++i)
                for (int b = 0; b < 4; ++b)
                    for (int r = 0; r < 3; ++r) {
                        double acc = 0;
                        for (int j = 0; j < 3; ++j)
                            for (int s = 0; s < 3; ++s) acc += D9[(j * 3 + i) * 9 + s * 3 + r] * c[a][j] * c[b][s];
                        o[(3 * a + i) * 12 + 3 * b + r] = acc;
                    }
    }
}
