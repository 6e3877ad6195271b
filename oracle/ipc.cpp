// Oracle (test infrastructure only): restatement of IDP's contact path.
//   surface extraction      Library/Utils/MESHIO.h:450-545, 769-835
//   spatial hash            Library/Grid/SPATIAL_HASH.h:28-213 (static), 432-622 (swept)
//   constraint set          Library/FEM/IPC.h:19-740
//   barrier E / grad / hess Library/FEM/IPC.h:742-1731
//   min distance            Library/FEM/IPC.h:2246-2388
//   ACCD step size          Library/FEM/IPC.h:1879-2244, Library/Math/Distance/CCD.h:280-395
#include "oracle.h"
#include "geom.hpp"

#include <array>
#include <cstdint>
#include <map>
#include <set>
#include <unordered_map>
#include <unordered_set>
#include <vector>

using namespace orc;

namespace {

inline V3 P(const double* X, int i) { return {X[3 * i], X[3 * i + 1], X[3 * i + 2]}; }

// Fixed-order sum: sequential over chunks of 256, then sequential over the
// chunk sums.  The HIP path reduces in the same order.
double chunked_sum(const double* a, long n)
{
    double tot = 0.0;
    for (long c = 0; c < n; c += 256) {
        double s = 0.0;
        long e = std::min(n, c + 256);
        for (long i = c; i < e; ++i) s += a[i];
        tot += s;
    }
    return tot;
}

struct Grid {
    V3 lbc;
    double inv;
    int cnt[3];
    long cxy;
    void locate(V3 p, int out[3]) const
    {
        out[0] = (int)std::floor((p.x - lbc.x) * inv);
        out[1] = (int)std::floor((p.y - lbc.y) * inv);
        out[2] = (int)std::floor((p.z - lbc.z) * inv);
    }
    long id(int x, int y, int z) const { return (long)x + (long)y * cnt[0] + (long)z * cxy; }
};

// SPATIAL_HASH.h:68-84 (voxel count from range, 1e9 cap, overflow guard)
void grid_setup(Grid& g, V3 lo, V3 hi, double voxelSize)
{
    g.lbc = lo;
    double range[3] = {hi.x - lo.x, hi.y - lo.y, hi.z - lo.z};
    g.inv = 1.0 / voxelSize;
    long vcl[3];
    for (int d = 0; d < 3; ++d) vcl[d] = std::max(1L, (long)std::ceil(range[d] * g.inv));
    long amt = vcl[0] * vcl[1] * vcl[2];
    if (amt > 1e9) {
        voxelSize *= std::pow(amt / 1.0e9, 1.0 / 3);
        g.inv = 1.0 / voxelSize;
    }
    bool bad = false;
    for (int d = 0; d < 3; ++d) {
        g.cnt[d] = std::max(1, (int)std::ceil(range[d] * g.inv));
        if (g.cnt[d] <= 0) bad = true;
    }
    if (bad) {
        g.inv = 1.0 / (std::max(range[0], std::max(range[1], range[2])) * 1.01);
        g.cnt[0] = g.cnt[1] = g.cnt[2] = 1;
    }
    g.cxy = (long)g.cnt[0] * g.cnt[1];
}

double mean_edge_len(const double* X, int ne, const int* bedge)
{
    std::vector<double> len(ne);
    for (int i = 0; i < ne; ++i) len[i] = std::sqrt(sqn(P(X, bedge[2 * i]) - P(X, bedge[2 * i + 1])));
    return chunked_sum(len.data(), ne) / ne;
}

struct MeshView {
    int nV;
    const double* X;
    const double* X0;  // rest positions (mollifier threshold)
    const unsigned char* dbc;
    int nb, ne, nf;
    const int* bnode;
    const int* bedge;
    const int* btri;
};

// ---------------------------------------------------------------------------
// Constraint set (3D tet-surface branch of Compute_Constraint_Set)
// ---------------------------------------------------------------------------
using I4 = std::array<int, 4>;

void pt_candidates_hash(const MeshView& m, double r, std::vector<std::vector<int>>& cand,
                        std::vector<std::vector<int>>& candEE)
{
    // SPATIAL_HASH::Build (static) + Query_Point_For_Triangles / Query_Edge_For_Edges
    V3 lo = P(m.X, 0), hi = lo;
    for (int i = 1; i < m.nV; ++i) {
        V3 p = P(m.X, i);
        lo = {std::min(lo.x, p.x), std::min(lo.y, p.y), std::min(lo.z, p.z)};
        hi = {std::max(hi.x, p.x), std::max(hi.y, p.y), std::max(hi.z, p.z)};
    }
    double vs = m.ne ? mean_edge_len(m.X, m.ne, m.bedge) : 1.0;
    Grid g;
    grid_setup(g, lo, hi, vs);
    std::vector<std::array<int, 3>> vai(m.nV);
    for (int i = 0; i < m.nb; ++i) {
        int v = m.bnode[i];
        g.locate(P(m.X, v), vai[v].data());
    }
    std::unordered_map<long, std::vector<int>> vox;  // id -> entries (>=0 edges tagged +ne? use two maps)
    std::unordered_map<long, std::vector<int>> voxE;
    for (int e = 0; e < m.ne; ++e) {
        auto& a = vai[m.bedge[2 * e]];
        auto& b = vai[m.bedge[2 * e + 1]];
        for (int z = std::min(a[2], b[2]); z <= std::max(a[2], b[2]); ++z)
            for (int y = std::min(a[1], b[1]); y <= std::max(a[1], b[1]); ++y)
                for (int x = std::min(a[0], b[0]); x <= std::max(a[0], b[0]); ++x) voxE[g.id(x, y, z)].push_back(e);
    }
    for (int t = 0; t < m.nf; ++t) {
        auto& a = vai[m.btri[3 * t]];
        auto& b = vai[m.btri[3 * t + 1]];
        auto& c = vai[m.btri[3 * t + 2]];
        int mn[3], mx[3];
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::min(a[d], std::min(b[d], c[d]));
            mx[d] = std::max(a[d], std::max(b[d], c[d]));
        }
        for (int z = mn[2]; z <= mx[2]; ++z)
            for (int y = mn[1]; y <= mx[1]; ++y)
                for (int x = mn[0]; x <= mx[0]; ++x) vox[g.id(x, y, z)].push_back(t);
    }
    cand.assign(m.nb, {});
    candEE.assign(m.ne, {});
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < m.nb; ++i) {
        V3 p = P(m.X, m.bnode[i]);
        int mn[3], mx[3];
        g.locate(p - V3{r, r, r}, mn);
        g.locate(p + V3{r, r, r}, mx);
        std::set<int> s;
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::max(mn[d], 0);
            mx[d] = std::min(mx[d], g.cnt[d] - 1);
        }
        for (int z = mn[2]; z <= mx[2]; ++z)
            for (int y = mn[1]; y <= mx[1]; ++y)
                for (int x = mn[0]; x <= mx[0]; ++x) {
                    auto it = vox.find(g.id(x, y, z));
                    if (it != vox.end()) s.insert(it->second.begin(), it->second.end());
                }
        cand[i].assign(s.begin(), s.end());
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int e = 0; e < m.ne; ++e) {
        V3 a = P(m.X, m.bedge[2 * e]), b = P(m.X, m.bedge[2 * e + 1]);
        V3 l = {std::min(a.x, b.x) - r, std::min(a.y, b.y) - r, std::min(a.z, b.z) - r};
        V3 h = {std::max(a.x, b.x) + r, std::max(a.y, b.y) + r, std::max(a.z, b.z) + r};
        int mn[3], mx[3];
        g.locate(l, mn);
        g.locate(h, mx);
        for (int d = 0; d < 3; ++d) {
            mn[d] = std::max(mn[d], 0);
            mx[d] = std::min(mx[d], g.cnt[d] - 1);
        }
        std::set<int> s;
        for (int z = mn[2]; z <= mx[2]; ++z)
            for (int y = mn[1]; y <= mx[1]; ++y)
                for (int x = mn[0]; x <= mx[0]; ++x) {
                    auto it = voxE.find(g.id(x, y, z));
                    if (it != voxE.end())
                        for (int f : it->second)
                            if (f > e) s.insert(f);
                }
        candEE[e].assign(s.begin(), s.end());
    }
}

// CCD.h:150-185
inline bool pt_cd_broadphase(V3 p, V3 t0, V3 t1, V3 t2, double dist)
{
    double mx[3] = {std::max(t0.x, std::max(t1.x, t2.x)), std::max(t0.y, std::max(t1.y, t2.y)),
                    std::max(t0.z, std::max(t1.z, t2.z))};
    double mn[3] = {std::min(t0.x, std::min(t1.x, t2.x)), std::min(t0.y, std::min(t1.y, t2.y)),
                    std::min(t0.z, std::min(t1.z, t2.z))};
    double pp[3] = {p.x, p.y, p.z};
    for (int d = 0; d < 3; ++d)
        if (pp[d] - mx[d] > dist || mn[d] - pp[d] > dist) return false;
    return true;
}
inline bool ee_cd_broadphase(V3 a0, V3 a1, V3 b0, V3 b1, double dist)
{
    double A0[3] = {a0.x, a0.y, a0.z}, A1[3] = {a1.x, a1.y, a1.z}, B0[3] = {b0.x, b0.y, b0.z},
           B1[3] = {b1.x, b1.y, b1.z};
    for (int d = 0; d < 3; ++d) {
        double mxa = std::max(A0[d], A1[d]), mna = std::min(A0[d], A1[d]);
        double mxb = std::max(B0[d], B1[d]), mnb = std::min(B0[d], B1[d]);
        if (mna - mxb > dist || mnb - mxa > dist) return false;
    }
    return true;
}

// PT part, IPC.h:158-268 (tet surfaces: no shell / codim branches)
void pt_pair(const MeshView& m, int vI, int f, double dHat, double dHat2, std::vector<I4>& out)
{
    const int* t = m.btri + 3 * f;
    if (vI == t[0] || vI == t[1] || vI == t[2]) return;
    if (m.dbc[vI] && m.dbc[t[0]] && m.dbc[t[1]] && m.dbc[t[2]]) return;
    V3 p = P(m.X, vI), t0 = P(m.X, t[0]), t1 = P(m.X, t[1]), t2 = P(m.X, t[2]);
    if (!pt_cd_broadphase(p, t0, t1, t2, dHat)) return;
    double d;
    switch (pt_type(p, t0, t1, t2)) {
    case 0: d = d2_pp(p, t0); if (d < dHat2) out.push_back({-vI - 1, t[0], -1, -1}); break;
    case 1: d = d2_pp(p, t1); if (d < dHat2) out.push_back({-vI - 1, t[1], -1, -1}); break;
    case 2: d = d2_pp(p, t2); if (d < dHat2) out.push_back({-vI - 1, t[2], -1, -1}); break;
    case 3: d = d2_pe(p, t0, t1); if (d < dHat2) out.push_back({-vI - 1, t[0], t[1], -1}); break;
    case 4: d = d2_pe(p, t1, t2); if (d < dHat2) out.push_back({-vI - 1, t[1], t[2], -1}); break;
    case 5: d = d2_pe(p, t2, t0); if (d < dHat2) out.push_back({-vI - 1, t[2], t[0], -1}); break;
    default: d = d2_pt(p, t0, t1, t2); if (d < dHat2) out.push_back({-vI - 1, t[0], t[1], t[2]}); break;
    }
}

// EE part, IPC.h:371-568
void ee_pair(const MeshView& m, int eI, int eJ, double dHat, double dHat2, std::vector<I4>& out)
{
    const int* A = m.bedge + 2 * eI;
    const int* B = m.bedge + 2 * eJ;
    if (A[0] == B[0] || A[0] == B[1] || A[1] == B[0] || A[1] == B[1] || eI > eJ) return;
    if (m.dbc[A[0]] && m.dbc[A[1]] && m.dbc[B[0]] && m.dbc[B[1]]) return;
    V3 a0 = P(m.X, A[0]), a1 = P(m.X, A[1]), b0 = P(m.X, B[0]), b1 = P(m.X, B[1]);
    if (!ee_cd_broadphase(a0, a1, b0, b1, dHat)) return;
    double cr = ee_cross_sqn(a0, a1, b0, b1);
    double eps = mollifier_eps(P(m.X0, A[0]), P(m.X0, A[1]), P(m.X0, B[0]), P(m.X0, B[1]));
    bool mol = cr < eps;
    double d;
    switch (ee_type(a0, a1, b0, b1)) {
    case 0:
        d = d2_pp(a0, b0);
        if (d < dHat2) out.push_back(mol ? I4{A[0], B[0], -A[1] - 1, -B[1] - 1} : I4{-A[0] - 1, B[0], -1, -1});
        break;
    case 1:
        d = d2_pp(a0, b1);
        if (d < dHat2) out.push_back(mol ? I4{A[0], B[1], -A[1] - 1, -B[0] - 1} : I4{-A[0] - 1, B[1], -1, -1});
        break;
    case 2:
        d = d2_pe(a0, b0, b1);
        if (d < dHat2) out.push_back(mol ? I4{A[0], B[0], B[1], -A[1] - 1} : I4{-A[0] - 1, B[0], B[1], -1});
        break;
    case 3:
        d = d2_pp(a1, b0);
        if (d < dHat2) out.push_back(mol ? I4{A[1], B[0], -A[0] - 1, -B[1] - 1} : I4{-A[1] - 1, B[0], -1, -1});
        break;
    case 4:
        d = d2_pp(a1, b1);
        if (d < dHat2) out.push_back(mol ? I4{A[1], B[1], -A[0] - 1, -B[0] - 1} : I4{-A[1] - 1, B[1], -1, -1});
        break;
    case 5:
        d = d2_pe(a1, b0, b1);
        if (d < dHat2) out.push_back(mol ? I4{A[1], B[0], B[1], -A[0] - 1} : I4{-A[1] - 1, B[0], B[1], -1});
        break;
    case 6:
        d = d2_pe(b0, a0, a1);
        if (d < dHat2) out.push_back(mol ? I4{B[0], A[0], A[1], -B[1] - 1} : I4{-B[0] - 1, A[0], A[1], -1});
        break;
    case 7:
        d = d2_pe(b1, a0, a1);
        if (d < dHat2) out.push_back(mol ? I4{B[1], A[0], A[1], -B[0] - 1} : I4{-B[1] - 1, A[0], A[1], -1});
        break;
    default:
        d = d2_ee(a0, a1, b0, b1);
        if (d < dHat2) out.push_back(mol ? I4{A[0], A[1], -B[0] - 1, B[1]} : I4{A[0], A[1], B[0], B[1]});
        break;
    }
}

}  // namespace

extern "C" long orc_constraint_set(int nV, const double* X, const double* X0, const unsigned char* dbc, int nb,
                                   const int* bnode, int ne, const int* bedge, int nf, const int* btri,
                                   double dHat2, double thickness, int use_hash, int* out, long cap)
{
    MeshView m{nV, X, X0, dbc, nb, ne, nf, bnode, bedge, btri};
    // IPC.h:51-52: the radius is re-squared
    const double dHat = std::sqrt(dHat2) + thickness;
    const double dHat2r = dHat * dHat;
    std::vector<std::vector<I4>> csPT(nb), csEE(ne);
    if (use_hash) {
        std::vector<std::vector<int>> cand, candEE;
        pt_candidates_hash(m, dHat, cand, candEE);
#pragma omp parallel for schedule(dynamic, 64)
        for (int i = 0; i < nb; ++i)
            for (int f : cand[i]) pt_pair(m, bnode[i], f, dHat, dHat2r, csPT[i]);
#pragma omp parallel for schedule(dynamic, 64)
        for (int e = 0; e < ne; ++e)
            for (int f : candEE[e]) ee_pair(m, e, f, dHat, dHat2r, csEE[e]);
    } else {
#pragma omp parallel for schedule(dynamic, 16)
        for (int i = 0; i < nb; ++i)
            for (int f = 0; f < nf; ++f) pt_pair(m, bnode[i], f, dHat, dHat2r, csPT[i]);
#pragma omp parallel for schedule(dynamic, 16)
        for (int e = 0; e < ne; ++e)
            for (int f = e + 1; f < ne; ++f) ee_pair(m, e, f, dHat, dHat2r, csEE[e]);
    }
    // Merge, IPC.h:587-654: PT-4 and EE/mollified pass through, plain PP/PE
    // are counted and appended in key order with -count in the last slot.
    std::vector<I4> cs;
    std::map<I4, int> counter;
    for (auto& v : csPT)
        for (auto& c : v) {
            if (c[3] < 0) ++counter[c];
            else cs.push_back(c);
        }
    for (auto& v : csEE)
        for (auto& c : v) {
            if (c[0] < 0) ++counter[c];
            else cs.push_back(c);
        }
    for (auto& kv : counter) cs.push_back({kv.first[0], kv.first[1], kv.first[2], -kv.second});
    long n = (long)cs.size();
    if (n > cap) return -n;
    for (long i = 0; i < n; ++i)
        for (int k = 0; k < 4; ++k) out[4 * i + k] = cs[i][k];
    return n;
}

// ---------------------------------------------------------------------------
// Barrier energy / gradient / Hessian per constraint (IPC.h:742-1731).
// ---------------------------------------------------------------------------
namespace {

struct PairEval {
    int nv;          // vertices involved
    int vid[4];      // global vertex ids, in the reference's triplet order
    double E;
    double g[12];
    double H[144];   // nv*3 square, row-major
};

// Builds f = mult * b(d) [* e(q)] as a jet over the 3*nv coordinates.
template <int N>
void finish(PairEval& pe, const Jet<N>& f)
{
    pe.E = f.v;
    for (int i = 0; i < N; ++i) pe.g[i] = f.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) pe.H[i * N + j] = f.H[i][j];
}

template <int N>
Jet<N> barrier_of(const Jet<N>& d, double th2, double dh, double k)
{
    Jet<N> dd = d;
    dd.v -= th2;
    return compose(dd, barrier_b(dd.v, dh, k), barrier_db(dd.v, dh, k), barrier_d2b(dd.v, dh, k));
}

bool eval_pair(const double* X, const double* X0, const int* c, double dHat2, double kappa, double thickness,
               PairEval& pe)
{
    const double th2 = thickness * thickness;
    const double dh = dHat2 + 2 * std::sqrt(dHat2) * thickness;  // IPC.h:758
    if (c[0] >= 0) {
        if (c[2] >= 0 && c[3] >= 0) {  // ++++ EE
            pe.nv = 4;
            for (int k = 0; k < 4; ++k) pe.vid[k] = c[k];
            auto d = jd2_ee(P(X, c[0]), P(X, c[1]), P(X, c[2]), P(X, c[3]));
            if (d.v - th2 <= 0) return false;
            finish(pe, barrier_of(d, th2, dh, kappa));
            return true;
        }
        int ev[4];
        int kind;
        if (c[3] >= 0) {  // ++-+ mollified EE
            ev[0] = c[0]; ev[1] = c[1]; ev[2] = -c[2] - 1; ev[3] = c[3];
            kind = 0;
        } else if (c[2] >= 0) {  // +++- mollified PE
            ev[0] = c[0]; ev[1] = -c[3] - 1; ev[2] = c[1]; ev[3] = c[2];
            kind = 1;
        } else {  // ++-- mollified PP
            ev[0] = c[0]; ev[1] = -c[2] - 1; ev[2] = c[1]; ev[3] = -c[3] - 1;
            kind = 2;
        }
        pe.nv = 4;
        for (int k = 0; k < 4; ++k) pe.vid[k] = ev[k];
        V3 a0 = P(X, ev[0]), a1 = P(X, ev[1]), b0 = P(X, ev[2]), b1 = P(X, ev[3]);
        Jet<12> d;
        if (kind == 0) d = jd2_ee(a0, a1, b0, b1);
        else if (kind == 1) d = jd2_pe_slots<12>(a0, b0, b1, 0, 2, 3);
        else d = jd2_pp_slots<12>(a0, b0, 0, 2);
        if (d.v - th2 <= 0) return false;
        auto b = barrier_of(d, th2, dh, kappa);
        double eps = mollifier_eps(P(X0, ev[0]), P(X0, ev[1]), P(X0, ev[2]), P(X0, ev[3]));
        auto q = jee_cross_sqn(a0, a1, b0, b1);
        Jet<12> e = (q.v < eps) ? compose(q, moll_e(q.v, eps), moll_de(q.v, eps), moll_d2e(q.v, eps))
                                : Jet<12>::cst(1.0);
        finish(pe, b * e);
        return true;
    }
    int p = -c[0] - 1;
    if (c[3] >= 0) {  // -+++ PT
        pe.nv = 4;
        pe.vid[0] = p; pe.vid[1] = c[1]; pe.vid[2] = c[2]; pe.vid[3] = c[3];
        auto d = jd2_pt(P(X, p), P(X, c[1]), P(X, c[2]), P(X, c[3]));
        if (d.v - th2 <= 0) return false;
        finish(pe, barrier_of(d, th2, dh, kappa));
        return true;
    }
    double mult = (double)(-c[3]);
    if (c[2] >= 0) {  // -++- PE with multiplicity
        pe.nv = 3;
        pe.vid[0] = p; pe.vid[1] = c[1]; pe.vid[2] = c[2];
        auto d = jd2_pe_slots<9>(P(X, p), P(X, c[1]), P(X, c[2]), 0, 1, 2);
        if (d.v - th2 <= 0) return false;
        finish(pe, scale(mult, barrier_of(d, th2, dh, kappa)));
        return true;
    }
    pe.nv = 2;  // -+-- PP with multiplicity
    pe.vid[0] = p; pe.vid[1] = c[1];
    auto d = jd2_pp_slots<6>(P(X, p), P(X, c[1]), 0, 1);
    if (d.v - th2 <= 0) return false;
    finish(pe, scale(mult, barrier_of(d, th2, dh, kappa)));
    return true;
}

}  // namespace

extern "C" int orc_barrier_energy(const double* X, const double* X0, long n, const int* cs, double dHat2,
                                  double kappa, double thickness, double* E_out, double* per_pair)
{
    std::vector<double> e(n);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (long i = 0; i < n; ++i) {
        PairEval pe;
        if (!eval_pair(X, X0, cs + 4 * i, dHat2, kappa, thickness, pe)) {
            ++bad;
            e[i] = 0;
            continue;
        }
        e[i] = pe.E;
    }
    if (bad) return ORC_E_NONPOS_DIST;
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += e[i];
    if (per_pair)
        for (long i = 0; i < n; ++i) per_pair[i] = e[i];
    *E_out = s;
    return 0;
}

extern "C" int orc_barrier_gradient(const double* X, const double* X0, long n, const int* cs, double dHat2,
                                    double kappa, double thickness, int nV, double* g /* 3nV, += */)
{
    std::vector<PairEval> pes(n);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 256) reduction(+ : bad)
    for (long i = 0; i < n; ++i)
        if (!eval_pair(X, X0, cs + 4 * i, dHat2, kappa, thickness, pes[i])) ++bad;
    if (bad) return ORC_E_NONPOS_DIST;
    for (long i = 0; i < n; ++i)
        for (int a = 0; a < pes[i].nv; ++a)
            for (int d = 0; d < 3; ++d) g[3 * pes[i].vid[a] + d] += pes[i].g[3 * a + d];
    (void)nV;
    return 0;
}

// Local (per-constraint) PSD-projected Hessians: H_out[i*144 ..] holds the
// nv*3 square matrix row-major, vid_out[i*4 ..] the vertex ids (unused = -1).
extern "C" int orc_barrier_hessian_local(const double* X, const double* X0, long n, const int* cs, double dHat2,
                                         double kappa, double thickness, int projectSPD, double* H_out,
                                         int* vid_out)
{
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : bad)
    for (long i = 0; i < n; ++i) {
        PairEval pe;
        if (!eval_pair(X, X0, cs + 4 * i, dHat2, kappa, thickness, pe)) {
            ++bad;
            continue;
        }
        int N = 3 * pe.nv;
        if (projectSPD) make_pd(N, pe.H);
        for (int k = 0; k < 144; ++k) H_out[144 * i + k] = (k < N * N) ? pe.H[k] : 0.0;
        for (int k = 0; k < 4; ++k) vid_out[4 * i + k] = (k < pe.nv) ? pe.vid[k] : -1;
    }
    return bad ? ORC_E_NONPOS_DIST : 0;
}

// IPC.h:2246-2388: per-pair squared distance by stored type, minimum minus th^2
extern "C" int orc_min_dist2(const double* X, long n, const int* cs, double thickness, double* dist2,
                             double* minDist2)
{
    double mn = INFINITY;
    for (long i = 0; i < n; ++i) {
        const int* c = cs + 4 * i;
        double d;
        if (c[0] >= 0) {
            if (c[2] >= 0 && c[3] >= 0) d = d2_ee(P(X, c[0]), P(X, c[1]), P(X, c[2]), P(X, c[3]));
            else if (c[3] >= 0) d = d2_ee(P(X, c[0]), P(X, c[1]), P(X, -c[2] - 1), P(X, c[3]));
            else if (c[2] >= 0) d = d2_pe(P(X, c[0]), P(X, c[1]), P(X, c[2]));
            else d = d2_pp(P(X, c[0]), P(X, c[1]));
        } else {
            int p = -c[0] - 1;
            if (c[3] >= 0) d = d2_pt(P(X, p), P(X, c[1]), P(X, c[2]), P(X, c[3]));
            else if (c[2] >= 0) d = d2_pe(P(X, p), P(X, c[1]), P(X, c[2]));
            else d = d2_pp(P(X, p), P(X, c[1]));
        }
        if (dist2) dist2[i] = d;
        mn = std::min(mn, d);
    }
    *minDist2 = mn - thickness * thickness;
    return 0;
}

// ---------------------------------------------------------------------------
// Additive CCD step size (Compute_Intersection_Free_StepSize, 3D tet surface).
// ---------------------------------------------------------------------------
namespace {

int g_accd_max_steps = 0;  // orc_set_accd_max_steps; <= 0: no cap

// CCD.h:280-328
bool accd_pt(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double eta, double th, double& toc)
{
    V3 mov = 0.25 * (((dt0 + dt1) + dt2) + dp);
    // the reference divides by 4; multiplying by 0.25 is exact
    dt0 = dt0 - mov;
    dt1 = dt1 - mov;
    dt2 = dt2 - mov;
    dp = dp - mov;
    double m2 = std::max(sqn(dt0), std::max(sqn(dt1), sqn(dt2)));
    double maxDisp = std::sqrt(sqn(dp)) + std::sqrt(m2);
    if (maxDisp == 0) return false;
    double d2 = d2_pt_unclassified(p, t0, t1, t2);
    double dc = std::sqrt(d2);
    double gap = eta * (d2 - th * th) / (dc + th);
    double tprev = toc;
    toc = 0;
    int steps = 0;
    while (true) {
        double lb = (1 - eta) * (d2 - th * th) / ((dc + th) * maxDisp);
        p = p + lb * dp;
        t0 = t0 + lb * dt0;
        t1 = t1 + lb * dt1;
        t2 = t2 + lb * dt2;
        d2 = d2_pt_unclassified(p, t0, t1, t2);
        dc = std::sqrt(d2);
        if (toc != 0.0 && ((d2 - th * th) / (dc + th) < gap)) break;
        toc += lb;
        if (toc > tprev) return false;
        if (++steps >= g_accd_max_steps && g_accd_max_steps > 0) break;
    }
    return true;
}

// CCD.h:331-395
bool accd_ee(V3 a0, V3 a1, V3 b0, V3 b1, V3 da0, V3 da1, V3 db0, V3 db1, double eta, double th, double& toc)
{
    V3 mov = 0.25 * (((da0 + da1) + db0) + db1);
    da0 = da0 - mov;
    da1 = da1 - mov;
    db0 = db0 - mov;
    db1 = db1 - mov;
    double maxDisp = std::sqrt(std::max(sqn(da0), sqn(da1))) + std::sqrt(std::max(sqn(db0), sqn(db1)));
    if (maxDisp == 0) return false;
    auto fallback = [&](double& d2v) {
        double ds[4] = {sqn(a0 - b0), sqn(a0 - b1), sqn(a1 - b0), sqn(a1 - b1)};
        d2v = std::min(std::min(ds[0], ds[1]), std::min(ds[2], ds[3]));
    };
    double d2 = d2_ee_unclassified(a0, a1, b0, b1);
    double f = d2 - th * th;
    if (f <= 0) {
        fallback(d2);
        f = d2 - th * th;
    }
    double dc = std::sqrt(d2);
    double gap = eta * f / (dc + th);
    double tprev = toc;
    toc = 0;
    int steps = 0;
    while (true) {
        double lb = (1 - eta) * f / ((dc + th) * maxDisp);
        a0 = a0 + lb * da0;
        a1 = a1 + lb * da1;
        b0 = b0 + lb * db0;
        b1 = b1 + lb * db1;
        d2 = d2_ee_unclassified(a0, a1, b0, b1);
        f = d2 - th * th;
        if (f <= 0) {
            fallback(d2);
            f = d2 - th * th;
        }
        dc = std::sqrt(d2);
        if (toc != 0.0 && (f / (dc + th) < gap)) break;
        toc += lb;
        if (toc > tprev) return false;
        if (++steps >= g_accd_max_steps && g_accd_max_steps > 0) break;
    }
    return true;
}

bool pt_ccd_broadphase(V3 p, V3 t0, V3 t1, V3 t2, V3 dp, V3 dt0, V3 dt1, V3 dt2, double dist)
{
    V3 pe = p + dp, t0e = t0 + dt0, t1e = t1 + dt1, t2e = t2 + dt2;
    double P0[3] = {p.x, p.y, p.z}, P1[3] = {pe.x, pe.y, pe.z};
    V3 ts[6] = {t0, t1, t2, t0e, t1e, t2e};
    for (int d = 0; d < 3; ++d) {
        double mxp = std::max(P0[d], P1[d]), mnp = std::min(P0[d], P1[d]);
        double mxt = -INFINITY, mnt = INFINITY;
        for (auto& t : ts) {
            double c = d == 0 ? t.x : (d == 1 ? t.y : t.z);
            mxt = std::max(mxt, c);
            mnt = std::min(mnt, c);
        }
        if (mnp - mxt > dist || mnt - mxp > dist) return false;
    }
    return true;
}
bool ee_ccd_broadphase(V3 a0, V3 a1, V3 b0, V3 b1, V3 da0, V3 da1, V3 db0, V3 db1, double dist)
{
    V3 as[4] = {a0, a1, a0 + da0, a1 + da1};
    V3 bs[4] = {b0, b1, b0 + db0, b1 + db1};
    for (int d = 0; d < 3; ++d) {
        double mxa = -INFINITY, mna = INFINITY, mxb = -INFINITY, mnb = INFINITY;
        for (int k = 0; k < 4; ++k) {
            double ca = d == 0 ? as[k].x : (d == 1 ? as[k].y : as[k].z);
            double cb = d == 0 ? bs[k].x : (d == 1 ? bs[k].y : bs[k].z);
            mxa = std::max(mxa, ca);
            mna = std::min(mna, ca);
            mxb = std::max(mxb, cb);
            mnb = std::min(mnb, cb);
        }
        if (mna - mxb > dist || mnb - mxa > dist) return false;
    }
    return true;
}

}  // namespace

extern "C" int orc_ccd_stepsize(int nV, const double* X, const double* pdir, const unsigned char* dbc, int nb,
                                const int* bnode, int ne, const int* bedge, int nf, const int* btri,
                                double thickness, double alpha_in, double* alpha_out, long* n_cand)
{
    (void)nV;
    double stepSize = alpha_in;
    // SPATIAL_HASH::Build (CCD form), SPATIAL_HASH.h:432-622
    double vs = ne ? mean_edge_len(X, ne, bedge) : 1.0;
    double pSize = 0;
    for (int i = 0; i < nb; ++i) {
        int v = bnode[i];
        pSize += std::fabs(pdir[3 * v]);
        pSize += std::fabs(pdir[3 * v + 1]);
        pSize += std::fabs(pdir[3 * v + 2]);
    }
    pSize /= (double)(nb * 3);
    const double span = stepSize * pSize / vs;
    if (span > 1) stepSize /= span;

    std::vector<V3> SV(nb), SVt(nb);
    std::vector<int> v2s(std::max(nV, 1), -1);
    V3 lo{INFINITY, INFINITY, INFINITY}, hi{-INFINITY, -INFINITY, -INFINITY};
    for (int i = 0; i < nb; ++i) {
        int v = bnode[i];
        v2s[v] = i;
        SV[i] = P(X, v);
        SVt[i] = {X[3 * v] + stepSize * pdir[3 * v], X[3 * v + 1] + stepSize * pdir[3 * v + 1],
                  X[3 * v + 2] + stepSize * pdir[3 * v + 2]};
    }
    // colwise min of SV, colwise min of SVt, then componentwise min
    V3 loA{INFINITY, INFINITY, INFINITY}, loB = loA, hiA{-INFINITY, -INFINITY, -INFINITY}, hiB = hiA;
    for (int i = 0; i < nb; ++i) {
        loA = {std::min(loA.x, SV[i].x), std::min(loA.y, SV[i].y), std::min(loA.z, SV[i].z)};
        loB = {std::min(loB.x, SVt[i].x), std::min(loB.y, SVt[i].y), std::min(loB.z, SVt[i].z)};
        hiA = {std::max(hiA.x, SV[i].x), std::max(hiA.y, SV[i].y), std::max(hiA.z, SV[i].z)};
        hiB = {std::max(hiB.x, SVt[i].x), std::max(hiB.y, SVt[i].y), std::max(hiB.z, SVt[i].z)};
    }
    lo = {std::min(loA.x, loB.x) - thickness / 2, std::min(loA.y, loB.y) - thickness / 2,
          std::min(loA.z, loB.z) - thickness / 2};
    hi = {std::max(hiA.x, hiB.x) + thickness / 2, std::max(hiA.y, hiB.y) + thickness / 2,
          std::max(hiA.z, hiB.z) + thickness / 2};
    Grid g;
    grid_setup(g, lo, hi, vs);
    std::vector<std::array<int, 3>> mn(nb), mx(nb);
    for (int i = 0; i < nb; ++i) {
        V3 a = {std::min(SV[i].x, SVt[i].x) - thickness / 2, std::min(SV[i].y, SVt[i].y) - thickness / 2,
                std::min(SV[i].z, SVt[i].z) - thickness / 2};
        V3 b = {std::max(SV[i].x, SVt[i].x) + thickness / 2, std::max(SV[i].y, SVt[i].y) + thickness / 2,
                std::max(SV[i].z, SVt[i].z) + thickness / 2};
        g.locate(a, mn[i].data());
        g.locate(b, mx[i].data());
    }
    // occupancy ids of points, edges (0..nb+ne) and triangles
    std::unordered_map<long, std::vector<int>> voxT, voxE;
    auto boxOf = [&](const int* vs_, int k, int bmn[3], int bmx[3]) {
        for (int d = 0; d < 3; ++d) {
            bmn[d] = INT32_MAX;
            bmx[d] = INT32_MIN;
        }
        for (int j = 0; j < k; ++j) {
            int s = v2s[vs_[j]];
            for (int d = 0; d < 3; ++d) {
                bmn[d] = std::min(bmn[d], mn[s][d]);
                bmx[d] = std::max(bmx[d], mx[s][d]);
            }
        }
    };
    std::vector<std::vector<long>> occE(ne);
    for (int e = 0; e < ne; ++e) {
        int bmn[3], bmx[3];
        boxOf(bedge + 2 * e, 2, bmn, bmx);
        for (int z = bmn[2]; z <= bmx[2]; ++z)
            for (int y = bmn[1]; y <= bmx[1]; ++y)
                for (int x = bmn[0]; x <= bmx[0]; ++x) {
                    long id = g.id(x, y, z);
                    occE[e].push_back(id);
                    voxE[id].push_back(e);
                }
    }
    for (int f = 0; f < nf; ++f) {
        int bmn[3], bmx[3];
        boxOf(btri + 3 * f, 3, bmn, bmx);
        for (int z = bmn[2]; z <= bmx[2]; ++z)
            for (int y = bmn[1]; y <= bmx[1]; ++y)
                for (int x = bmn[0]; x <= bmx[0]; ++x) voxT[g.id(x, y, z)].push_back(f);
    }
    std::vector<double> aPT(nb, stepSize), aEE(ne, stepSize);
    long ncand = 0;
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : ncand)
    for (int i = 0; i < nb; ++i) {
        int v = bnode[i];
        std::set<int> tris;
        for (int z = mn[i][2]; z <= mx[i][2]; ++z)
            for (int y = mn[i][1]; y <= mx[i][1]; ++y)
                for (int x = mn[i][0]; x <= mx[i][0]; ++x) {
                    auto it = voxT.find(g.id(x, y, z));
                    if (it != voxT.end()) tris.insert(it->second.begin(), it->second.end());
                }
        V3 p = P(X, v), dp = P(pdir, v);
        for (int f : tris) {
            const int* t = btri + 3 * f;
            if (v == t[0] || v == t[1] || v == t[2]) continue;
            if (dbc[v] && dbc[t[0]] && dbc[t[1]] && dbc[t[2]]) continue;
            V3 t0 = P(X, t[0]), t1 = P(X, t[1]), t2 = P(X, t[2]);
            V3 dt0 = P(pdir, t[0]), dt1 = P(pdir, t[1]), dt2 = P(pdir, t[2]);
            if (!pt_ccd_broadphase(p, t0, t1, t2, dp, dt0, dt1, dt2, thickness)) continue;
            ++ncand;
            double la = aPT[i];
            if (accd_pt(p, t0, t1, t2, dp, dt0, dt1, dt2, 0.1, thickness, la))
                if (aPT[i] > la) aPT[i] = la;
        }
    }
#pragma omp parallel for schedule(dynamic, 32) reduction(+ : ncand)
    for (int e = 0; e < ne; ++e) {
        std::set<int> edges;
        for (long id : occE[e]) {
            auto it = voxE.find(id);
            if (it != voxE.end())
                for (int f : it->second)
                    if (f > e) edges.insert(f);
        }
        const int* A = bedge + 2 * e;
        V3 a0 = P(X, A[0]), a1 = P(X, A[1]), da0 = P(pdir, A[0]), da1 = P(pdir, A[1]);
        for (int f : edges) {
            const int* B = bedge + 2 * f;
            if (A[0] == B[0] || A[0] == B[1] || A[1] == B[0] || A[1] == B[1]) continue;
            if (dbc[A[0]] && dbc[A[1]] && dbc[B[0]] && dbc[B[1]]) continue;
            V3 b0 = P(X, B[0]), b1 = P(X, B[1]), db0 = P(pdir, B[0]), db1 = P(pdir, B[1]);
            if (!ee_ccd_broadphase(a0, a1, b0, b1, da0, da1, db0, db1, thickness)) continue;
            ++ncand;
            double la = aEE[e];
            if (accd_ee(a0, a1, b0, b1, da0, da1, db0, db1, 0.1, thickness, la))
                if (aEE[e] > la) aEE[e] = la;
        }
    }
    for (double a : aPT) stepSize = std::min(stepSize, a);
    for (double a : aEE) stepSize = std::min(stepSize, a);
    *alpha_out = stepSize;
    if (n_cand) *n_cand = ncand;
    return 0;
}

extern "C" void orc_set_accd_max_steps(int n) { g_accd_max_steps = n; }

// ---------------------------------------------------------------------------
// Surface primitives (MESHIO.h:450-545 Find_Surface_TriMesh<T,false>, and
// MESHIO.h:769-835 Find_Surface_Primitives_And_Compute_Area)
// ---------------------------------------------------------------------------
extern "C" int orc_surface(int nV, const double* X, int nT, const int* tets, int* nb, int* ne, int* nf,
                           int* bnode, int* bedge, int* btri)
{
    using T3 = std::array<int, 3>;
    std::map<T3, int> tri2tet;
    for (int t = 0; t < nT; ++t) {
        const int* e = tets + 4 * t;
        tri2tet[{e[0], e[2], e[1]}] = t;
        tri2tet[{e[0], e[3], e[2]}] = t;
        tri2tet[{e[0], e[1], e[3]}] = t;
        tri2tet[{e[1], e[2], e[3]}] = t;
    }
    std::vector<T3> tri;
    for (auto& kv : tri2tet) {
        const T3& k = kv.first;
        if (tri2tet.count({k[2], k[1], k[0]}) || tri2tet.count({k[1], k[0], k[2]}) ||
            tri2tet.count({k[0], k[2], k[1]}))
            continue;
        tri.push_back(k);
    }
    std::map<std::array<int, 2>, double> eset;
    std::vector<double> area(nV, 0.0);
    for (auto& t : tri) {
        V3 v0 = P(X, t[0]), v1 = P(X, t[1]), v2 = P(X, t[2]);
        double A = 0.5 * std::sqrt(sqn(cross(v1 - v0, v2 - v0)));
        int pairs[3][2] = {{t[0], t[1]}, {t[1], t[2]}, {t[2], t[0]}};
        for (auto& pr : pairs) {
            auto it = eset.find({pr[1], pr[0]});
            if (it == eset.end()) eset[{pr[0], pr[1]}] = A / 3;
            else it->second += A / 3;
        }
        area[t[0]] += A / 3;
        area[t[1]] += A / 3;
        area[t[2]] += A / 3;
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

// Geometry probe for tests: kind 0 = point-triangle (p,t0,t1,t2),
// kind 1 = edge-edge (a0,a1,b0,b1), kind 2 = point-edge (p,e0,e1).
// Writes the distance type and the unclassified squared distance.
extern "C" int orc_geom_probe(int kind, const double* x, int* type, double* d2)
{
    V3 a = P(x, 0), b = P(x, 1), c = P(x, 2), d = P(x, 3);
    if (kind == 0) {
        *type = pt_type(a, b, c, d);
        *d2 = d2_pt_unclassified(a, b, c, d);
    } else if (kind == 1) {
        *type = ee_type(a, b, c, d);
        *d2 = d2_ee_unclassified(a, b, c, d);
    } else {
        *type = pe_type(a, b, c);
        *d2 = (*type == 0) ? d2_pp(a, b) : (*type == 1 ? d2_pp(a, c) : d2_pe(a, b, c));
    }
    return 0;
}
