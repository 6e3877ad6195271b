// CPU restatement of the lagged IPC friction of the step (test
// infrastructure: the parity checker and CPU baseline, never the product).
//
//   Compute_Friction_Basis      Library/FEM/FRICTION.h:16-117
//   Compute_Friction_Coef       Library/FEM/FRICTION.h:119-169
//   Compute_Friction_Potential  Library/FEM/FRICTION.h:171-243
//   Compute_Friction_Gradient   Library/FEM/FRICTION.h:245-383
//   Compute_Friction_Hessian    Library/FEM/FRICTION.h:385-660
//   f0 / f1 / f2, closest points, tangent bases, TT maps
//                               Library/FEM/FRICTION_UTILS.h:9-287
//
// A friction constraint is a barrier constraint of the set at the basis
// update that is not mollified (PT, EE, PE, PP in the constraint set's int4
// encoding), kept in the set's order.  Per constraint the basis update
// stores its closest-point parameters (2), tangent basis (3 x 2, by
// columns), and normal force -b'(d - t^2) 2 sqrt(d) (weight 1: no area
// weighting in this barrier).  Vertex order is the constraint's decode
// order: PT (p, t0, t1, t2), EE (a0, a1, b0, b1), PE (p, e0, e1), PP (p, q).
//
// The reference's point-point relative displacement reads the second
// point's position, not its displacement (FRICTION.h:212-214, 318-320,
// 479-481); the restatement keeps that, so its PP friction depends on
// where q is.
#include <cmath>
#include <cstring>
#include <vector>

#include "geom.hpp"
#include "oracle.h"

using namespace orc;

namespace {

V3 node(const double* X, int v) { return {X[3 * v], X[3 * v + 1], X[3 * v + 2]}; }

// Eigen's normalized(): v / sqrt(|v|^2) when |v|^2 > 0, else v
V3 normalized(V3 v)
{
    const double n = sqn(v);
    if (n > 0) {
        const double r = std::sqrt(n);
        return {v.x / r, v.y / r, v.z / r};
    }
    return v;
}

enum { F_PT = 0, F_EE = 1, F_PE = 2, F_PP = 3, F_SKIP = -1 };

// kind of a barrier constraint for friction; vertex ids in decode order;
// multiplicity (PE / PP carry -w)
int fkind(const int* c, int* v, double* mult)
{
    *mult = 1.0;
    if (c[0] >= 0) {
        if (c[2] < 0 || c[3] < 0) return F_SKIP;  // mollified: no friction (FRICTION.h:38-42)
        v[0] = c[0]; v[1] = c[1]; v[2] = c[2]; v[3] = c[3];
        return F_EE;
    }
    v[0] = -c[0] - 1;
    v[1] = c[1];
    if (c[3] >= 0) {
        v[2] = c[2]; v[3] = c[3];
        return F_PT;
    }
    *mult = (double)(-c[3]);
    if (c[2] >= 0) {
        v[2] = c[2]; v[3] = -1;
        return F_PE;
    }
    v[2] = v[3] = -1;
    return F_PP;
}

int fnv(int k) { return k == F_PP ? 2 : (k == F_PE ? 3 : 4); }

// FRICTION_UTILS.h:9-36
double f0_sf(double x2, double epsvh)
{
    if (x2 >= epsvh * epsvh) return std::sqrt(x2);
    return x2 * (-std::sqrt(x2) / 3.0 + epsvh) / (epsvh * epsvh) + epsvh / 3.0;
}
double f1_div(double x2, double epsvh)
{
    if (x2 >= epsvh * epsvh) return 1 / std::sqrt(x2);
    return (-std::sqrt(x2) + 2.0 * epsvh) / (epsvh * epsvh);
}
double f2_term(double, double epsvh) { return -1 / (epsvh * epsvh); }

// TT coefficients: the relative displacement is sum_i c_i dx_i
void tt_coef(int k, const double* cp, double* c)
{
    switch (k) {
    case F_PT:  // FRICTION_UTILS.h:92-102
        c[0] = 1; c[1] = -1 + cp[0] + cp[1]; c[2] = -cp[0]; c[3] = -cp[1];
        break;
    case F_EE:  // FRICTION_UTILS.h:176-186
        c[0] = 1.0 - cp[0]; c[1] = cp[0]; c[2] = cp[1] - 1.0; c[3] = -cp[1];
        break;
    case F_PE:  // FRICTION_UTILS.h:235-244
        c[0] = 1; c[1] = cp[0] - 1.0; c[2] = -cp[0]; c[3] = 0;
        break;
    default:  // FRICTION_UTILS.h:281-287
        c[0] = 1; c[1] = -1; c[2] = c[3] = 0;
    }
}

// relative displacement (FRICTION_UTILS.h:70-79, 152-161, 219-228, 271-277)
V3 rel_dx(int k, const int* v, const double* cp, const double* X, const double* Xn)
{
    auto dx = [&](int i) { return node(X, v[i]) - node(Xn, v[i]); };
    switch (k) {
    case F_PT: {
        V3 d0 = dx(0), d1 = dx(1), d2 = dx(2), d3 = dx(3);
        return d0 - ((d1 + cp[0] * (d2 - d1)) + cp[1] * (d3 - d1));
    }
    case F_EE: {
        V3 d0 = dx(0), d1 = dx(1), d2 = dx(2), d3 = dx(3);
        return (d0 + cp[0] * (d1 - d0)) - (d2 + cp[1] * (d3 - d2));
    }
    case F_PE: {
        V3 d0 = dx(0), d1 = dx(1), d2 = dx(2);
        return d0 - (d1 + cp[0] * (d2 - d1));
    }
    default:
        return dx(0) - node(X, v[1]);  // the reference's PP reads x_q, not its displacement
    }
}

}  // namespace

// Compute_Friction_Basis<double,3,false>: the friction constraints of cs (n
// int4) at X, their closest-point parameters (2 each; PE uses the first),
// tangent bases (6 each: column 0 then column 1) and normal forces.
// Returns the friction constraint count (<= n).
extern "C" long orc_friction_basis(const double* X, long n, const int* cs, double dHat2, double kappa,
                                   double thickness, int* fcs, double* cp, double* basis, double* nforce)
{
    const double th2 = thickness * thickness;
    const double dh = dHat2 + 2 * std::sqrt(dHat2) * thickness;
    long m = 0;
    for (long i = 0; i < n; ++i) {
        int v[4];
        double mult;
        const int k = fkind(cs + 4 * i, v, &mult);
        if (k == F_SKIP) continue;
        double* q = cp + 2 * m;
        double* b = basis + 6 * m;
        V3 c0, c1;
        double d2;
        q[0] = q[1] = 0.0;
        if (k == F_EE) {
            const V3 a0 = node(X, v[0]), a1 = node(X, v[1]), b0 = node(X, v[2]), b1 = node(X, v[3]);
            const V3 e20 = a0 - b0, e01 = a1 - a0, e23 = b1 - b0;
            ldlt2_solve(dot(e01, e01), -dot(e23, e01), dot(e23, e23), -dot(e20, e01), dot(e20, e23), q[0], q[1]);
            c0 = normalized(e01);
            c1 = normalized(cross(cross(e01, b1 - b0), e01));
            d2 = d2_ee(a0, a1, b0, b1);
        } else if (k == F_PT) {
            const V3 p = node(X, v[0]), t0 = node(X, v[1]), t1 = node(X, v[2]), t2 = node(X, v[3]);
            const V3 r0 = t1 - t0, r1 = t2 - t0, pd = p - t0;
            ldlt2_solve(dot(r0, r0), dot(r0, r1), dot(r1, r1), dot(r0, pd), dot(r1, pd), q[0], q[1]);
            c0 = normalized(r0);
            c1 = normalized(cross(cross(r0, t2 - t0), r0));
            d2 = d2_pt(p, t0, t1, t2);
        } else if (k == F_PE) {
            const V3 p = node(X, v[0]), e0 = node(X, v[1]), e1 = node(X, v[2]);
            const V3 e = e1 - e0;
            q[0] = dot(p - e0, e) / sqn(e);
            c0 = normalized(e);
            c1 = normalized(cross(e, p - e0));
            d2 = d2_pe(p, e0, e1);
        } else {
            const V3 p = node(X, v[0]), p1 = node(X, v[1]);
            const V3 w = p1 - p;
            const V3 xc{0.0, -w.z, w.y}, yc{w.z, 0.0, -w.x};  // UnitX x w, UnitY x w
            const V3 r = sqn(xc) > sqn(yc) ? xc : yc;
            c0 = normalized(r);
            c1 = normalized(cross(w, r));
            d2 = d2_pp(p, p1);
        }
        b[0] = c0.x; b[1] = c0.y; b[2] = c0.z;
        b[3] = c1.x; b[4] = c1.y; b[5] = c1.z;
        const double bg = barrier_db(d2 - th2, dh, kappa);
        nforce[m] = -bg * 2 * std::sqrt(d2) * 1.0;
        std::memcpy(fcs + 4 * m, cs + 4 * i, 4 * sizeof(int));
        ++m;
    }
    return m;
}

// Compute_Friction_Coef: normal force *= mu_comp[c0 + c1 * ncomp], c the
// components (first index with v < comp_range[c]) of the constraint's first
// vertex and of its other side (EE: a0 and b0; others: the point and the
// first vertex after it).  Returns -1 when a vertex is in no component.
extern "C" int orc_friction_coef(long nf, const int* fcs, int ncomp, const int* comp_range, const double* mu_comp,
                                 double* nforce)
{
    auto comp = [&](int v) {
        for (int c = 0; c < ncomp; ++c)
            if (v < comp_range[c]) return c;
        return -1;
    };
    for (long i = 0; i < nf; ++i) {
        const int* c = fcs + 4 * i;
        const int c0 = comp(c[0] >= 0 ? c[0] : -c[0] - 1);
        const int c1 = comp(c[0] >= 0 ? c[2] : c[1]);
        if (c0 < 0 || c1 < 0) return -1;
        nforce[i] *= mu_comp[c0 + c1 * ncomp];
    }
    return 0;
}

// Compute_Friction_Potential: mu * sum_i N_i f0(|T_i^T dx_rel|) (times the
// multiplicity when it is above 1), summed in constraint order
extern "C" double orc_friction_energy(const double* X, const double* Xn, long nf, const int* fcs, const double* cp,
                                      const double* basis, const double* nforce, double epsvh2, double mu)
{
    const double epsvh = std::sqrt(epsvh2);
    std::vector<double> EI(nf);
    for (long i = 0; i < nf; ++i) {
        int v[4];
        double mult;
        const int k = fkind(fcs + 4 * i, v, &mult);
        const V3 r = rel_dx(k, v, cp + 2 * i, X, Xn);
        const double* b = basis + 6 * i;
        const double u0 = r.x * b[0] + r.y * b[1] + r.z * b[2];
        const double u1 = r.x * b[3] + r.y * b[4] + r.z * b[5];
        EI[i] = f0_sf(u0 * u0 + u1 * u1, epsvh);
        EI[i] *= nforce[i];
        if (fcs[4 * i + 3] < -1) EI[i] *= -fcs[4 * i + 3];
    }
    double s = 0.0;
    for (double e : EI) s += e;
    return mu * s;
}

// Compute_Friction_Gradient: g (3 nV) += the friction force terms
extern "C" void orc_friction_gradient(const double* X, const double* Xn, long nf, const int* fcs, const double* cp,
                                      const double* basis, const double* nforce, double epsvh2, double mu, double* g)
{
    const double epsvh = std::sqrt(epsvh2);
    for (long i = 0; i < nf; ++i) {
        int v[4];
        double mult;
        const int k = fkind(fcs + 4 * i, v, &mult);
        const V3 r = rel_dx(k, v, cp + 2 * i, X, Xn);
        const double* b = basis + 6 * i;
        double u0 = r.x * b[0] + r.y * b[1] + r.z * b[2];
        double u1 = r.x * b[3] + r.y * b[4] + r.z * b[5];
        const double f1 = f1_div(u0 * u0 + u1 * u1, epsvh);
        const double sc = (k == F_PE || k == F_PP) ? f1 * mult * mu * nforce[i] : f1 * mu * nforce[i];
        u0 *= sc;
        u1 *= sc;
        const V3 t{b[0] * u0 + b[3] * u1, b[1] * u0 + b[4] * u1, b[2] * u0 + b[5] * u1};
        double c[4];
        tt_coef(k, cp + 2 * i, c);
        for (int a = 0; a < fnv(k); ++a) {
            const V3 ta = (k == F_EE || a > 0) ? c[a] * t : t;
            g[3 * v[a]] += ta.x;
            g[3 * v[a] + 1] += ta.y;
            g[3 * v[a] + 2] += ta.z;
        }
    }
}

// Compute_Friction_Hessian: per constraint the (3 nv)^2 matrix
// TT^T inner TT (row-major, into H + 144 i with stride 3 nv) and its vertex
// ids (4 per constraint, -1 padded)
extern "C" void orc_friction_hessian_local(const double* X, const double* Xn, long nf, const int* fcs,
                                           const double* cp, const double* basis, const double* nforce,
                                           double epsvh2, double mu, int projectSPD, double* H, int* vid)
{
    const double epsvh = std::sqrt(epsvh2);
    for (long i = 0; i < nf; ++i) {
        int v[4];
        double mult;
        const int k = fkind(fcs + 4 * i, v, &mult);
        const int nv = fnv(k), n = 3 * nv;
        for (int a = 0; a < 4; ++a) vid[4 * i + a] = a < nv ? v[a] : -1;
        const V3 r = rel_dx(k, v, cp + 2 * i, X, Xn);
        const double* b = basis + 6 * i;
        const double u[2] = {r.x * b[0] + r.y * b[1] + r.z * b[2], r.x * b[3] + r.y * b[4] + r.z * b[5]};
        const double x2 = u[0] * u[0] + u[1] * u[1];
        const double xn = std::sqrt(x2);
        double c[4];
        tt_coef(k, cp + 2 * i, c);
        double TT[2][12];  // FRICTION_UTILS.h:92-102, 176-186, 235-244, 281-287
        for (int a = 0; a < nv; ++a)
            for (int d = 0; d < 3; ++d) {
                const double w = (k == F_EE || a > 0) ? c[a] : 1.0;
                TT[0][3 * a + d] = (k == F_EE || a > 0) ? w * b[d] : b[d];
                TT[1][3 * a + d] = (k == F_EE || a > 0) ? w * b[3 + d] : b[3 + d];
            }
        const double f1 = f1_div(x2, epsvh), f2 = f2_term(x2, epsvh);
        const bool pm = (k == F_PE || k == F_PP);
        double* Hi = H + 144 * i;
        std::memset(Hi, 0, 144 * sizeof(double));
        if (x2 >= epsvh2) {
            const double s = pm ? mult * mu * nforce[i] * f1 / x2 : mu * nforce[i] * f1 / x2;
            const double ub[2] = {-u[1], u[0]};
            double q[12], w[12];
            for (int j = 0; j < n; ++j) {
                q[j] = TT[0][j] * (s * ub[0]) + TT[1][j] * (s * ub[1]);
                w[j] = ub[0] * TT[0][j] + ub[1] * TT[1][j];
            }
            for (int a = 0; a < n; ++a)
                for (int j = 0; j < n; ++j) Hi[a * n + j] = q[a] * w[j];
        } else if (xn == 0) {
            const double s = pm ? mult * mu * nforce[i] * f1 : mu * nforce[i] * f1;
            for (int a = 0; a < n; ++a)
                for (int j = 0; j < n; ++j) Hi[a * n + j] = (s * TT[0][a]) * TT[0][j] + (s * TT[1][a]) * TT[1][j];
        } else {
            double in[4];
            for (int p = 0; p < 2; ++p)
                for (int q2 = 0; q2 < 2; ++q2) in[2 * p + q2] = ((f2 / xn) * u[p]) * u[q2];
            in[0] += f1;
            in[3] += f1;
            if (projectSPD) make_pd(2, in);
            const double s = pm ? mult * mu * nforce[i] : mu * nforce[i];
            for (int e = 0; e < 4; ++e) in[e] *= s;
            for (int a = 0; a < n; ++a) {
                const double l0 = TT[0][a] * in[0] + TT[1][a] * in[2];
                const double l1 = TT[0][a] * in[1] + TT[1][a] * in[3];
                for (int j = 0; j < n; ++j) Hi[a * n + j] = l0 * TT[0][j] + l1 * TT[1][j];
            }
        }
    }
}
