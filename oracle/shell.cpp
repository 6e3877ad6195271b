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
