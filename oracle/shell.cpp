// CPU restatement of the discrete-shell pieces of the step (test
// infrastructure: the parity checker and CPU baseline, never the product).
//
//   shell set-up            Library/FEM/Shell/DISCRETE_SHELL.h:219-354
//                           (Initialize_Discrete_Shell: degenerate-triangle
//                           filter, lumped mass, vol = area * thickness)
//   contact primitives      Library/FEM/Shell/IMPLICIT_EULER.h:229-241 and
//                           Library/Utils/MESHIO.h:769-835 (tet surface in
//                           map order, then the shell triangles in Elem order)
//   normal-flow body force  Library/FEM/Shell/DISCRETE_SHELL.h:359-395
//   flow term -h/2 x^T L x  Library/FEM/Shell/INC_POTENTIAL.h:53-73,
//                           191-211, 323-339
#include <array>
#include <cstring>
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "geom.hpp"
#include "oracle.h"

using namespace orc;

namespace {
V3 node(const double* X, int v) { return {X[3 * v], X[3 * v + 1], X[3 * v + 2]}; }
}  // namespace

// Initialize_Discrete_Shell<double,3,false,false> (DISCRETE_SHELL.h:239-254,
// 279-302, 318-331): keeps the triangles whose first fundamental form has a
// nonzero determinant, in order; mass[v] += |cross(X2-X1, X3-X1)| / 2 *
// thickness * rho / 3 per kept triangle in order; vol = area * thickness;
// lam = E nu / (1 - nu^2), mu = E / (2 (1 + nu)).  Returns the kept count.
// mass is accumulated into (the caller zeroes it for a pure shell).
extern "C" int orc_shell_init(int nV, const double* X, int nE, const int* tris, double rho, double E, double nu,
                              double thickness, int* kept, double* mass, double* vol, double* lam, double* mu)
{
    (void)nV;
    int n = 0;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        V3 e01 = node(X, t[1]) - node(X, t[0]), e02 = node(X, t[2]) - node(X, t[0]);
        const double a = sqn(e01), b = dot(e01, e02), d = sqn(e02);
        if (a * d - b * b == 0.0) continue;
        for (int k = 0; k < 3; ++k) kept[3 * n + k] = t[k];
        ++n;
    }
    const double L = E * nu / (1.0 - nu * nu);
    const double M = E / (2.0 * (1.0 + nu));
    for (int e = 0; e < n; ++e) {
        const int* t = kept + 3 * e;
        V3 c = cross(node(X, t[1]) - node(X, t[0]), node(X, t[2]) - node(X, t[0]));
        const double len = std::sqrt(sqn(c));
        const double portion = len / 2 * thickness * rho / 3;
        for (int k = 0; k < 3; ++k) mass[t[k]] += portion;
        vol[e] = len / 2 * thickness;
        lam[e] = L;
        mu[e] = M;
    }
    return n;
}

// Boundary primitives of a mesh with tets and shell triangles: the tets'
// surface triangles (map order, MESHIO.h:486-545) followed by the shell
// triangles in their order, then MESHIO.h:769-835 over that list: triangles
// as listed, edges with the orientation they are first seen in, in
// lexicographic order of that orientation, nodes with nonzero area ascending.
extern "C" int orc_shell_surface(int nV, const double* X, int nT, const int* tets, int nE, const int* tris, int* nb,
                                 int* ne, int* nf, int* bnode, int* bedge, int* btri)
{
    using T3 = std::array<int, 3>;
    std::vector<T3> tri;
    {
        std::map<T3, int> tri2tet;
        for (int t = 0; t < nT; ++t) {
            const int* e = tets + 4 * t;
            tri2tet[{e[0], e[2], e[1]}] = t;
            tri2tet[{e[0], e[3], e[2]}] = t;
            tri2tet[{e[0], e[1], e[3]}] = t;
            tri2tet[{e[1], e[2], e[3]}] = t;
        }
        for (auto& kv : tri2tet) {
            const T3& k = kv.first;
            if (tri2tet.count({k[2], k[1], k[0]}) || tri2tet.count({k[1], k[0], k[2]}) ||
                tri2tet.count({k[0], k[2], k[1]}))
                continue;
            tri.push_back(k);
        }
    }
    for (int e = 0; e < nE; ++e) tri.push_back({tris[3 * e], tris[3 * e + 1], tris[3 * e + 2]});
    std::map<std::array<int, 2>, double> eset;
    std::vector<double> area(nV, 0.0);
    for (auto& t : tri) {
        V3 v0 = node(X, t[0]), v1 = node(X, t[1]), v2 = node(X, t[2]);
        const double A = 0.5 * std::sqrt(sqn(cross(v1 - v0, v2 - v0)));
        const int pairs[3][2] = {{t[0], t[1]}, {t[1], t[2]}, {t[2], t[0]}};
        for (auto& pr : pairs) {
            auto it = eset.find({pr[1], pr[0]});
            if (it == eset.end()) eset[{pr[0], pr[1]}] = A / 3;
            else it->second += A / 3;
        }
        for (int k = 0; k < 3; ++k) area[t[k]] += A / 3;
    }
    *nf = (int)tri.size();
    for (size_t i = 0; i < tri.size(); ++i)
        for (int k = 0; k < 3; ++k) btri[3 * i + k] = tri[i][k];
    *ne = (int)eset.size();
    int i = 0;
    for (auto& kv : eset) {
        bedge[2 * i] = kv.first[0];
        bedge[2 * i + 1] = kv.first[1];
        ++i;
    }
    int c = 0;
    for (int v = 0; v < nV; ++v)
        if (area[v] != 0.0) bnode[c++] = v;
    *nb = c;
    return 0;
}

// Update_Normal_Flow_Neumann (DISCRETE_SHELL.h:359-395): per vertex, the sum
// over its triangles (in order) of cross(X2-X1, X3-X1), scaled to length
// magnitude * m_v.  A vertex in no triangle gets 0/0, as there.
extern "C" void orc_normal_flow_force(int nV, const double* X, int nE, const int* tris, const double* mass,
                                      double magnitude, double* b)
{
    for (long i = 0; i < 3L * nV; ++i) b[i] = 0.0;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        V3 n = cross(node(X, t[1]) - node(X, t[0]), node(X, t[2]) - node(X, t[0]));
        for (int k = 0; k < 3; ++k) {
            b[3 * t[k]] += n.x;
            b[3 * t[k] + 1] += n.y;
            b[3 * t[k] + 2] += n.z;
        }
    }
    for (int v = 0; v < nV; ++v) {
        double sq = 0;
        for (int d = 0; d < 3; ++d) sq += b[3 * v + d] * b[3 * v + d];
        const double w = magnitude * mass[v] / std::sqrt(sq);
        for (int d = 0; d < 3; ++d) b[3 * v + d] *= w;
    }
}

// The flow term alone at X: E = -h/2 x.Lx (INC_POTENTIAL.h:53-73), g = -h Lx
// (:191-211), and the 9 vertex blocks per triangle, (a,b) at 81 e + 9 (3a+b):
// -h vol/6 I off the diagonal, 2 h vol/6 I on it (:323-339).  Any of E, g, H
// may be null.
extern "C" void orc_flow_terms(int nV, const double* X, int nE, const int* tris, const double* vol, double h,
                               double* E, double* g, double* H)
{
    std::vector<double> Lx(3 * (size_t)nV, 0.0);
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        const double w = vol[e] / 6;
        for (int i = 0; i < 3; ++i) {
            const int a = t[i], b = t[(i + 1) % 3], d = t[(i + 2) % 3];
            for (int k = 0; k < 3; ++k) Lx[3 * a + k] += w * ((X[3 * b + k] + X[3 * d + k]) - 2 * X[3 * a + k]);
        }
    }
    if (E) {
        double s = 0;
        for (int v = 0; v < nV; ++v)
            for (int d = 0; d < 3; ++d) s += -h * 0.5 * X[3 * v + d] * Lx[3 * v + d];
        *E = s;
    }
    if (g)
        for (long i = 0; i < 3L * nV; ++i) g[i] = -h * Lx[i];
    if (H)
        for (int e = 0; e < nE; ++e)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double* o = H + 81 * (size_t)e + 9 * (3 * a + b);
                    const double s = a == b ? 2 * h * vol[e] / 6 : -h * vol[e] / 6;
                    for (int k = 0; k < 9; ++k) o[k] = (k % 4 == 0) ? s : 0.0;
                }
}

// ---------------------------------------------------------------------------
// Non-flow shell terms: neo-Hookean membrane (MEMBRANE.h:8-312, useNH) and
// discrete hinge bending (BENDING.h:9-80, 150-214, 440-497 with
// DIHEDRAL_ANGLE.h:10-24).  Derivatives come from second-order jets of the
// energies, not from the reference's derivative formulas.
// ---------------------------------------------------------------------------
namespace {
template <int N>
Jet<N> jlog(const Jet<N>& a)
{
    return compose(a, std::log(a.v), 1.0 / a.v, -1.0 / (a.v * a.v));
}
template <int N>
Jet<N> jatan(const Jet<N>& u)
{
    const double d = 1.0 + u.v * u.v;
    return compose(u, std::atan(u.v), 1.0 / d, -2.0 * u.v / (d * d));
}
// atan2(y, x) up to a constant: derivatives of the angle (the value is taken
// from the reference formula where it is needed)
template <int N>
Jet<N> jangle(const Jet<N>& y, const Jet<N>& x)
{
    if (std::fabs(x.v) >= std::fabs(y.v)) return jatan(y / x);
    return scale(-1.0, jatan(x / y));
}
V3 row(const double* X, int v) { return {X[3 * v], X[3 * v + 1], X[3 * v + 2]}; }
}  // namespace

// rest first fundamental form per triangle, (E01.E01, E01.E02, E02.E02)
// (DISCRETE_SHELL.h:146-158)
extern "C" void orc_shell_rest_fff(const double* X0, int nE, const int* tris, double* IB)
{
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        V3 a = row(X0, t[1]) - row(X0, t[0]), b = row(X0, t[2]) - row(X0, t[0]);
        IB[3 * e] = sqn(a);
        IB[3 * e + 1] = dot(a, b);
        IB[3 * e + 2] = sqn(b);
    }
}

// Membrane of one triangle: value (the reference's expression order) and, in
// J, the jet over (x0, x1, x2).  false: skipped (degenerate rest form)
namespace {
bool membrane_one(const double* X, const int* t, const double* IB, double vol, double lam, double mu, double w,
                  double* E, Jet<9>* J)
{
    const double a = IB[0], b = IB[1], d = IB[2];
    const double det = a * d - b * b;
    if (det == 0.0) return false;
    const double i00 = d / det, i01 = -b / det, i11 = a / det;  // IB^-1
    const double detI = i00 * i11 - i01 * i01;
    V3 x0 = row(X, t[0]), x1 = row(X, t[1]), x2 = row(X, t[2]);
    V3 e01 = x1 - x0, e02 = x2 - x0;
    const double A00 = sqn(e01), A01 = dot(e01, e02), A11 = sqn(e02);
    const double lnJ = std::log((A00 * A11 - A01 * A01) * detI) / 2;
    const double tr = (i00 * A00 + i01 * A01) + (i01 * A01 + i11 * A11);
    *E = w * vol * (mu / 2 * (tr - 2 - 2 * lnJ) + lam / 2 * lnJ * lnJ);
    if (J) {
        auto X0 = jvar<9>(x0, 0), X1 = jvar<9>(x1, 1), X2 = jvar<9>(x2, 2);
        auto E01 = jsub(X1, X0), E02 = jsub(X2, X0);
        auto JA00 = jdot(E01, E01), JA01 = jdot(E01, E02), JA11 = jdot(E02, E02);
        auto jl = scale(0.5, jlog(scale(detI, JA00 * JA11 - JA01 * JA01)));
        auto jtr = scale(i00, JA00) + scale(2 * i01, JA01) + scale(i11, JA11);
        auto psi = scale(mu / 2, jtr - Jet<9>::cst(2.0) - scale(2.0, jl)) + scale(lam / 2, jl * jl);
        *J = scale(w * vol, psi);
    }
    return true;
}

// dihedral angle of the hinge (x0, x1, x2, x3) around x1-x2, the reference's
// value (DIHEDRAL_ANGLE.h:10-24)
double dihedral(V3 v0, V3 v1, V3 v2, V3 v3)
{
    V3 n1 = cross(v1 - v0, v2 - v0), n2 = cross(v2 - v3, v1 - v3);
    double c = dot(n1, n2) / std::sqrt(sqn(n1) * sqn(n2));
    double da = std::acos(std::max(-1.0, std::min(1.0, c)));
    if (dot(cross(n2, n1), v1 - v2) < 0) da = -da;
    return da;
}

// the same angle as a jet over (x0, x1, x2, x3): atan2((n1 x n2).e, n1.n2)
// with e the unit shared edge x2 - x1 (the sign convention of the value)
Jet<12> jdihedral(V3 v0, V3 v1, V3 v2, V3 v3)
{
    auto J0 = jvar<12>(v0, 0), J1 = jvar<12>(v1, 1), J2 = jvar<12>(v2, 2), J3 = jvar<12>(v3, 3);
    auto n1 = jcross(jsub(J1, J0), jsub(J2, J0));
    auto n2 = jcross(jsub(J2, J3), jsub(J1, J3));
    auto ed = jsub(J2, J1);
    auto len = compose(jdot(ed, ed), std::sqrt(sqn(v2 - v1)), 0.5 / std::sqrt(sqn(v2 - v1)),
                       -0.25 / (sqn(v2 - v1) * std::sqrt(sqn(v2 - v1))));
    auto s = jdot(jcross(n1, n2), ed) / len;
    auto c = jdot(n1, n2);
    return jangle(s, c);
}
}  // namespace

// Hinge stencils of a shell (DISCRETE_SHELL.h:169-215): for each edge
// (e0, e1) of the sorted edge list that has a triangle on both sides, the
// stencil (v, e0, e1, w) with v after e1 in the triangle holding e0->e1 and w
// after e0 in the one holding e1->e0 (the last triangle listed wins, as in
// the std::map there); rest (theta, |e|, (|n1| + |n2|) / (6 |e|)).  Returns
// the hinge count.
extern "C" int orc_shell_hinges(const double* X0, int nE, const int* tris, int ne, const int* edges, int* stencil,
                                double* rest)
{
    std::map<std::pair<int, int>, int> e2t;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        e2t[{t[0], t[1]}] = e;
        e2t[{t[1], t[2]}] = e;
        e2t[{t[2], t[0]}] = e;
    }
    int nh = 0;
    for (int i = 0; i < ne; ++i) {
        const int a = edges[2 * i], b = edges[2 * i + 1];
        auto f = e2t.find({a, b});
        if (f == e2t.end()) continue;
        const int* t = tris + 3 * f->second;
        int v = -1;
        for (int j = 0; j < 3; ++j)
            if (t[j] == b) v = t[(j + 1) % 3];
        auto g = e2t.find({b, a});
        if (g == e2t.end()) continue;
        const int* o = tris + 3 * g->second;
        int w = 0;
        for (int j = 0; j < 3; ++j)
            if (o[j] == a) {
                w = o[(j + 1) % 3];
                break;
            }
        int* s = stencil + 4 * nh;
        s[0] = v;
        s[1] = a;
        s[2] = b;
        s[3] = w;
        V3 X0v = row(X0, v), X1 = row(X0, a), X2 = row(X0, b), X3 = row(X0, w);
        double* r = rest + 3 * nh;
        r[0] = dihedral(X0v, X1, X2, X3);
        r[1] = std::sqrt(sqn(X1 - X2));
        V3 n1 = cross(X1 - X0v, X2 - X0v), n2 = cross(X2 - X3, X1 - X3);
        r[2] = (std::sqrt(sqn(n1)) + std::sqrt(sqn(n2))) / (r[1] * 6);
        ++nh;
    }
    return nh;
}

// Membrane + bending at X: energy, gradient (added into g), local Hessians
// (81 per triangle, vertex order of the triangle; 144 per hinge, stencil
// order), each PSD-projected when project.  w = h^2; k = bending stiffness
// (bendingStiffMult * E t^3 / (24 (1 - nu^2))); dbc (nullable): triangles /
// hinges whose vertices are all DBC are skipped.  Any of E, g, Hm, Hb may be
// null.
extern "C" void orc_shell_elastic_terms(int nV, const double* X, int nE, const int* tris, const double* IB,
                                        const double* vol, const double* lam, const double* mu, int nH,
                                        const int* stencil, const double* rest, double k, double w,
                                        const unsigned char* dbc, int project, double* E, double* g, double* Hm,
                                        double* Hb)
{
    (void)nV;
    double Et = 0.0;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        if (Hm) std::memset(Hm + 81 * (size_t)e, 0, 81 * sizeof(double));
        if (dbc && dbc[t[0]] && dbc[t[1]] && dbc[t[2]]) continue;
        double Ee;
        Jet<9> J;
        if (!membrane_one(X, t, IB + 3 * e, vol[e], lam[e], mu[e], w, &Ee, (g || Hm) ? &J : nullptr)) continue;
        Et += Ee;
        if (g)
            for (int a = 0; a < 3; ++a)
                for (int d = 0; d < 3; ++d) g[3 * t[a] + d] += J.g[3 * a + d];
        if (Hm) {
            double* H = Hm + 81 * (size_t)e;
            for (int i = 0; i < 9; ++i)
                for (int j = 0; j < 9; ++j) H[9 * i + j] = 0.5 * (J.H[i][j] + J.H[j][i]);
            if (project) make_pd(9, H);
        }
    }
    for (int hI = 0; hI < nH; ++hI) {
        const int* s = stencil + 4 * hI;
        if (Hb) std::memset(Hb + 144 * (size_t)hI, 0, 144 * sizeof(double));
        if (dbc && dbc[s[0]] && dbc[s[1]] && dbc[s[2]] && dbc[s[3]]) continue;
        const double* r = rest + 3 * hI;
        V3 v0 = row(X, s[0]), v1 = row(X, s[1]), v2 = row(X, s[2]), v3 = row(X, s[3]);
        const double th = dihedral(v0, v1, v2, v3);
        Et += w * k * (th - r[0]) * (th - r[0]) * r[1] / r[2];
        if (!g && !Hb) continue;
        Jet<12> J = jdihedral(v0, v1, v2, v3);
        const double c1 = w * k * 2 * (th - r[0]) * r[1] / r[2], c2 = w * k * 2 * r[1] / r[2];
        if (g)
            for (int a = 0; a < 4; ++a)
                for (int d = 0; d < 3; ++d) g[3 * s[a] + d] += c1 * J.g[3 * a + d];
        if (Hb) {
            double* H = Hb + 144 * (size_t)hI;
            for (int i = 0; i < 12; ++i)
                for (int j = 0; j < 12; ++j)
                    H[12 * i + j] = c1 * 0.5 * (J.H[i][j] + J.H[j][i]) + c2 * J.g[i] * J.g[j];
            if (project) make_pd(12, H);
        }
    }
    if (E) *E = Et;
}

// the dihedral angle and its jet gradient (tests)
extern "C" double orc_dihedral(const double* x, double* grad)
{
    V3 v0 = row(x, 0), v1 = row(x, 1), v2 = row(x, 2), v3 = row(x, 3);
    if (grad) {
        Jet<12> J = jdihedral(v0, v1, v2, v3);
        for (int i = 0; i < 12; ++i) grad[i] = J.g[i];
    }
    return dihedral(v0, v1, v2, v3);
}
