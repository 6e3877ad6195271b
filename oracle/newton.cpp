// Oracle (test infrastructure only): restatement of the projected-Newton
// time step with IPC barrier for tet meshes,
//   Advance_One_Step_IE_Discrete_Shell  Library/FEM/Shell/IMPLICIT_EULER.h:152-891
//   Line_Search                         Library/FEM/Shell/IMPLICIT_EULER.h:10-149
//   Compute_IncPotential{,_Gradient,_Hessian}  Library/FEM/Shell/INC_POTENTIAL.h:15-395
//   CSR assembly + DBC projection       Library/Math/CSR_MATRIX.h:49-56, 130-141
//   lagged friction                     Library/FEM/Shell/IMPLICIT_EULER.h:413-437,
//                                       463-465, 757-848 (friction.cpp)
// The linear solve restates the direct solve (DIRECT_SOLVER.h:14-89) with a
// block-Jacobi preconditioned CG run to a tight relative residual (1e-12 by
// default); the direction therefore agrees with a direct factorization to
// that residual.
#include "oracle.h"
#include "geom.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <deque>
#include <vector>

using namespace orc;

namespace {

struct Blk {
    long key;  // row*nV + col
    double v[9];
};

struct Sys {
    int nV;
    std::vector<int> rowptr, col;
    std::vector<double> val;  // 9 per block, row-major 3x3
};

void add_local(std::vector<Blk>& out, int nv, const int* vid, const double* H, int nV)
{
    int N = 3 * nv;
    for (int a = 0; a < nv; ++a)
        for (int b = 0; b < nv; ++b) {
            Blk k;
            k.key = (long)vid[a] * nV + vid[b];
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) k.v[3 * i + j] = H[(3 * a + i) * N + 3 * b + j];
            out.push_back(k);
        }
}

void build_sys(std::vector<Blk>& blks, int nV, Sys& S)
{
    std::stable_sort(blks.begin(), blks.end(), [](const Blk& a, const Blk& b) { return a.key < b.key; });
    S.nV = nV;
    S.rowptr.assign(nV + 1, 0);
    S.col.clear();
    S.val.clear();
    for (size_t i = 0; i < blks.size();) {
        size_t j = i;
        double acc[9] = {0};
        while (j < blks.size() && blks[j].key == blks[i].key) {
            for (int k = 0; k < 9; ++k) acc[k] += blks[j].v[k];
            ++j;
        }
        int r = (int)(blks[i].key / nV), c = (int)(blks[i].key % nV);
        S.rowptr[r + 1]++;
        S.col.push_back(c);
        S.val.insert(S.val.end(), acc, acc + 9);
        i = j;
    }
    for (int r = 0; r < nV; ++r) S.rowptr[r + 1] += S.rowptr[r];
}

void spmv(const Sys& S, const double* x, double* y)
{
#pragma omp parallel for schedule(static)
    for (int r = 0; r < S.nV; ++r) {
        double a0 = 0, a1 = 0, a2 = 0;
        for (int k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k) {
            const double* b = S.val.data() + 9 * (size_t)k;
            const double* xc = x + 3 * (size_t)S.col[k];
            a0 += b[0] * xc[0] + b[1] * xc[1] + b[2] * xc[2];
            a1 += b[3] * xc[0] + b[4] * xc[1] + b[5] * xc[2];
            a2 += b[6] * xc[0] + b[7] * xc[1] + b[8] * xc[2];
        }
        y[3 * r] = a0;
        y[3 * r + 1] = a1;
        y[3 * r + 2] = a2;
    }
}

bool inv3(const double* a, double* o)
{
    double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8], c2 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
    if (det == 0) return false;
    double id = 1.0 / det;
    o[0] = c0 * id;
    o[3] = c1 * id;
    o[6] = c2 * id;
    o[1] = (a[2] * a[7] - a[1] * a[8]) * id;
    o[4] = (a[0] * a[8] - a[2] * a[6]) * id;
    o[7] = (a[1] * a[6] - a[0] * a[7]) * id;
    o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return true;
}

// Block-Jacobi PCG; returns iterations (negative on breakdown)
int pcg(const Sys& S, const double* b, double* x, double rtol, int maxit, double* relres)
{
    size_t n = 3 * (size_t)S.nV;
    std::vector<double> Dinv(9 * (size_t)S.nV, 0.0), r(n), z(n), p(n), q(n);
    for (int v = 0; v < S.nV; ++v)
        for (int k = S.rowptr[v]; k < S.rowptr[v + 1]; ++k)
            if (S.col[k] == v)
                if (!inv3(S.val.data() + 9 * (size_t)k, Dinv.data() + 9 * (size_t)v)) return -1;
    auto prec = [&](const double* in, double* out) {
#pragma omp parallel for
        for (int v = 0; v < S.nV; ++v) {
            const double* D = Dinv.data() + 9 * (size_t)v;
            for (int i = 0; i < 3; ++i)
                out[3 * v + i] = D[3 * i] * in[3 * v] + D[3 * i + 1] * in[3 * v + 1] + D[3 * i + 2] * in[3 * v + 2];
        }
    };
    auto dotp = [&](const double* a, const double* c) {
        double s = 0;
#pragma omp parallel for reduction(+ : s)
        for (long i = 0; i < (long)n; ++i) s += a[i] * c[i];
        return s;
    };
    std::fill(x, x + n, 0.0);
    for (size_t i = 0; i < n; ++i) r[i] = b[i];
    double bn = std::sqrt(dotp(b, b));
    if (bn == 0) {
        *relres = 0;
        return 0;
    }
    prec(r.data(), z.data());
    p = z;
    double rz = dotp(r.data(), z.data());
    int it = 0;
    double rn = bn;
    for (; it < maxit; ++it) {
        rn = std::sqrt(dotp(r.data(), r.data()));
        if (rn <= rtol * bn) break;
        spmv(S, p.data(), q.data());
        double pq = dotp(p.data(), q.data());
        if (pq <= 0) return -2;
        double al = rz / pq;
#pragma omp parallel for
        for (long i = 0; i < (long)n; ++i) {
            x[i] += al * p[i];
            r[i] -= al * q[i];
        }
        prec(r.data(), z.data());
        double rz2 = dotp(r.data(), z.data());
        double be = rz2 / rz;
        rz = rz2;
#pragma omp parallel for
        for (long i = 0; i < (long)n; ++i) p[i] = z[i] + be * p[i];
    }
    *relres = rn / bn;
    return it;
}

struct Ctx {
    int nV, nT, nb, ne, nf;
    const double* X0;
    const double* mass;
    const int* tets;
    const double *IB, *vol, *lam, *mu;
    const int *bnode, *bedge, *btri;
    std::vector<unsigned char> dbc;
    // moving DBC (Shell/IMPLICIT_EULER.h:304-401, Shell/DIRICHLET.h): DBC entries in order,
    // their targets, the nodes with zero displacement, and the penalty stiffness
    std::vector<int> dbc_idx;
    std::vector<double> dbc_tgt;
    std::vector<unsigned char> dbc_fixed;
    double dbc_stiff = 0;
    double h, dHat2, thickness;
    double* kappa;
    bool withCollision;
    // Advance_One_Step_IE_Flow (flow = true): the energy is inertia plus the
    // barrier, the volumetric term is skipped (INC_POTENTIAL.h:53-73 takes
    // only the shell flow term, none for a tet mesh)
    bool flow = false;
    // shell triangles (Elem) and their vol = area * thickness: the flow term
    // -h/2 x^T L x of INC_POTENTIAL.h:53-73, 191-211, 323-339
    int nE = 0;
    const int* tris = nullptr;
    const double* tvol = nullptr;
    // non-flow shell terms (membrane, hinge bending); null IB: none
    const double *tIB = nullptr, *tlam = nullptr, *tmu = nullptr;
    int nH = 0;
    const int* hinge = nullptr;
    const double* hrest = nullptr;
    double bend_k = 0.0;
    std::vector<double> Xtilde;
    std::vector<int> cs;
    long ncs = 0;
    // lagged friction (mu > 0 or per-component coefficients): the set, its
    // closest points, bases and normal forces at the last basis update, and
    // the step's start positions Xn
    bool fric = false;
    double fmu = 0.0, epsvh2 = 0.0;
    int ncomp = 0;
    const int* comp_range = nullptr;
    const double* mu_comp = nullptr;
    const double* Xn = nullptr;
    std::vector<int> fcs;
    std::vector<double> fcp, fbasis, fnf;
    long nfc = 0;
    orc_stats* st;
    double t_cs = 0, t_ccd = 0, t_solve = 0, t_hess = 0, t_grad = 0, t_E = 0;
};

double now()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

void compute_cs(Ctx& c, const double* X)
{
    double t0 = now();
    long cap = std::max<long>(1024, c.ncs * 2 + 1024);
    for (;;) {
        c.cs.resize(4 * cap);
        long n = orc_constraint_set(c.nV, X, c.X0, c.dbc.data(), c.nb, c.bnode, c.ne, c.bedge, c.nf, c.btri,
                                    c.dHat2, c.thickness, 1, c.cs.data(), cap);
        if (n >= 0) {
            c.ncs = n;
            break;
        }
        cap = -n;
    }
    c.t_cs += now() - t0;
    if (c.st) c.st->n_constraint_sets++;
}

// Shell/DIRICHLET.h:6-23
double dbc_dist2(const Ctx& c, const double* X)
{
    std::vector<double> pen(c.dbc_idx.size(), 0.0);
    for (size_t i = 0; i < c.dbc_idx.size(); ++i) {
        const int v = c.dbc_idx[i];
        for (int d = 0; d < 3; ++d) pen[i] += (c.dbc_tgt[3 * i + d] - X[3 * v + d]) * (c.dbc_tgt[3 * i + d] - X[3 * v + d]);
    }
    double s = 0.0;
    for (double p : pen) s += p;
    return s;
}

// Shell/DIRICHLET.h:25-48: E += 1/2 k sum m |x_D - x|^2
void dbc_energy(const Ctx& c, const double* X, double& E)
{
    std::vector<double> pen(c.dbc_idx.size(), 0.0);
    for (size_t i = 0; i < c.dbc_idx.size(); ++i) {
        const int v = c.dbc_idx[i];
        for (int d = 0; d < 3; ++d) pen[i] += (c.dbc_tgt[3 * i + d] - X[3 * v + d]) * (c.dbc_tgt[3 * i + d] - X[3 * v + d]);
        pen[i] *= c.mass[v];
    }
    double s = 0.0;
    for (double p : pen) s += p;
    E += 0.5 * c.dbc_stiff * s;
}

extern "C" long orc_friction_basis(const double* X, long n, const int* cs, double dHat2, double kappa,
                                   double thickness, int* fcs, double* cp, double* basis, double* nforce);
extern "C" int orc_friction_coef(long nf, const int* fcs, int ncomp, const int* comp_range, const double* mu_comp,
                                 double* nforce);
extern "C" double orc_friction_energy(const double* X, const double* Xn, long nf, const int* fcs, const double* cp,
                                      const double* basis, const double* nforce, double epsvh2, double mu);
extern "C" void orc_friction_gradient(const double* X, const double* Xn, long nf, const int* fcs, const double* cp,
                                      const double* basis, const double* nforce, double epsvh2, double mu, double* g);
extern "C" void orc_friction_hessian_local(const double* X, const double* Xn, long nf, const int* fcs,
                                           const double* cp, const double* basis, const double* nforce,
                                           double epsvh2, double mu, int projectSPD, double* H, int* vid);

// Compute_Friction_Basis (+ Compute_Friction_Coef with per-component
// coefficients, which sets mu to 1) on the current constraint set at X
// (IMPLICIT_EULER.h:421-426, 763-768); false when a vertex is in no component
bool friction_basis(Ctx& c, const double* X)
{
    const long n = c.ncs;
    c.fcs.resize(4 * (size_t)n + 4);
    c.fcp.resize(2 * (size_t)n + 2);
    c.fbasis.resize(6 * (size_t)n + 6);
    c.fnf.resize((size_t)n + 1);
    c.nfc = orc_friction_basis(X, n, c.cs.data(), c.dHat2, c.kappa[0], c.thickness, c.fcs.data(), c.fcp.data(),
                               c.fbasis.data(), c.fnf.data());
    if (c.ncomp > 0) {
        if (orc_friction_coef(c.nfc, c.fcs.data(), c.ncomp, c.comp_range, c.mu_comp, c.fnf.data())) return false;
        c.fmu = 1.0;
    }
    return true;
}

extern "C" void orc_shell_elastic_terms(int nV, const double* X, int nE, const int* tris, const double* IB,
                                        const double* vol, const double* lam, const double* mu, int nH,
                                        const int* stencil, const double* rest, double k, double w,
                                        const unsigned char* dbc, int project, double* E, double* g, double* Hm,
                                        double* Hb);

// Lx of the flow term (INC_POTENTIAL.h:53-66): per triangle in order, for
// each corner i, Lx_i += vol/6 (x_(i+1) + x_(i+2) - 2 x_i)
void flow_lx(const Ctx& c, const double* X, std::vector<double>& Lx)
{
    Lx.assign(3 * (size_t)c.nV, 0.0);
    for (int e = 0; e < c.nE; ++e) {
        const int* t = c.tris + 3 * e;
        const double w = c.tvol[e] / 6;
        for (int i = 0; i < 3; ++i) {
            const int a = t[i], b = t[(i + 1) % 3], d = t[(i + 2) % 3];
            for (int k = 0; k < 3; ++k) Lx[3 * a + k] += w * ((X[3 * b + k] + X[3 * d + k]) - 2 * X[3 * a + k]);
        }
    }
}

// Compute_IncPotential (tet + inertia + barrier); returns validity
bool energy(Ctx& c, const double* X, bool withBarrier, double& E)
{
    double t0 = now();
    double Ee = 0;
    if (!c.flow && !orc_elastic_energy(X, c.nT, c.tets, c.IB, c.vol, c.lam, c.mu, c.h * c.h, &Ee)) {
        c.t_E += now() - t0;
        return false;
    }
    if (!c.flow && c.nE && c.tIB) {  // membrane and bending, w = h^2 (INC_POTENTIAL.h:80-95)
        double Es = 0.0;
        orc_shell_elastic_terms(c.nV, X, c.nE, c.tris, c.tIB, c.tvol, c.tlam, c.tmu, c.nH, c.hinge, c.hrest,
                                c.bend_k, c.h * c.h, c.dbc.data(), 1, &Es, nullptr, nullptr, nullptr);
        Ee += Es;
    }
    if (c.flow && c.nE) {  // INC_POTENTIAL.h:67-72: value += -h/2 x.Lx, vertex by vertex
        std::vector<double> Lx;
        flow_lx(c, X, Lx);
        for (int v = 0; v < c.nV; ++v)
            for (int d = 0; d < 3; ++d) Ee += -c.h * 0.5 * X[3 * v + d] * Lx[3 * v + d];
    }
    double Ei = 0;
    for (int v = 0; v < c.nV; ++v)
        for (int d = 0; d < 3; ++d) {
            double dx = X[3 * v + d] - c.Xtilde[3 * v + d];
            Ei += (c.mass[v] * dx) * dx;
        }
    E = Ee + 0.5 * Ei;
    if (withBarrier && c.ncs) {
        double Eb = 0;
        if (orc_barrier_energy(X, c.X0, c.ncs, c.cs.data(), c.dHat2, c.kappa[0], c.thickness, &Eb, nullptr))
            return false;
        E += Eb;
    }
    if (withBarrier && c.fric && c.fmu > 0)  // IMPLICIT_EULER.h:137-139: E += mu sum N f0
        E += orc_friction_energy(X, c.Xn, c.nfc, c.fcs.data(), c.fcp.data(), c.fbasis.data(), c.fnf.data(), c.epsvh2,
                                 c.fmu);
    c.t_E += now() - t0;
    return true;
}

// Projected Hessian of the incremental potential at X (INC_POTENTIAL.h:278-395):
// tet blocks + barrier blocks of the current constraint set + mass, summed
// per (row, col) in a stable key order, then Project_DBC.
void assemble_sys(Ctx& c, const double* X, Sys& S)
{
    const int nV = c.nV, nT = c.nT;
    const int* tets = c.tets;
    std::vector<Blk> blks;
    blks.reserve(16 * (size_t)nT + nV + 16 * c.ncs);
    if (!c.flow) {
        std::vector<double> He(144 * (size_t)nT);
        orc_elastic_hessian_local(X, nT, tets, c.IB, c.vol, c.lam, c.mu, c.h * c.h, 1, He.data());
        for (int t = 0; t < nT; ++t) add_local(blks, 4, tets + 4 * t, He.data() + 144 * (size_t)t, nV);
    }
    if (c.withCollision && c.ncs) {
        std::vector<double> Hl(144 * c.ncs);
        std::vector<int> vl(4 * c.ncs);
        orc_barrier_hessian_local(X, c.X0, c.ncs, c.cs.data(), c.dHat2, c.kappa[0], c.thickness, 1, Hl.data(),
                                  vl.data());
        for (long i = 0; i < c.ncs; ++i) {
            int nv = 0;
            while (nv < 4 && vl[4 * i + nv] >= 0) ++nv;
            add_local(blks, nv, vl.data() + 4 * i, Hl.data() + 144 * i, nV);
        }
    }
    if (c.withCollision && c.fric && c.fmu > 0 && c.nfc) {  // INC_POTENTIAL.h:375-377
        std::vector<double> Hl(144 * c.nfc);
        std::vector<int> vl(4 * c.nfc);
        orc_friction_hessian_local(X, c.Xn, c.nfc, c.fcs.data(), c.fcp.data(), c.fbasis.data(), c.fnf.data(),
                                   c.epsvh2, c.fmu, 1, Hl.data(), vl.data());
        for (long i = 0; i < c.nfc; ++i) {
            int nv = 0;
            while (nv < 4 && vl[4 * i + nv] >= 0) ++nv;
            add_local(blks, nv, vl.data() + 4 * i, Hl.data() + 144 * i, nV);
        }
    }
    if (!c.flow && c.nE && c.tIB) {  // projected membrane (9 x 9) and hinge (12 x 12) blocks
        std::vector<double> Hm(81 * (size_t)c.nE), Hb(144 * (size_t)c.nH);
        orc_shell_elastic_terms(nV, X, c.nE, c.tris, c.tIB, c.tvol, c.tlam, c.tmu, c.nH, c.hinge, c.hrest,
                                c.bend_k, c.h * c.h, c.dbc.data(), 1, nullptr, nullptr, Hm.data(), Hb.data());
        for (int e = 0; e < c.nE; ++e) add_local(blks, 3, c.tris + 3 * e, Hm.data() + 81 * (size_t)e, nV);
        for (int k = 0; k < c.nH; ++k) add_local(blks, 4, c.hinge + 4 * k, Hb.data() + 144 * (size_t)k, nV);
    }
    if (c.flow && c.nE) {  // INC_POTENTIAL.h:323-339: -h vol/6 to both neighbours, 2 h vol/6 on the diagonal
        for (int e = 0; e < c.nE; ++e) {
            const int* t = c.tris + 3 * e;
            for (int i = 0; i < 3; ++i) {
                const int a = t[i];
                const int nb3[3] = {t[(i + 1) % 3], t[(i + 2) % 3], a};
                const double sv[3] = {-c.h * c.tvol[e] / 6, -c.h * c.tvol[e] / 6, 2 * c.h * c.tvol[e] / 6};
                for (int q = 0; q < 3; ++q) {
                    Blk k;
                    k.key = (long)a * nV + nb3[q];
                    std::memset(k.v, 0, sizeof(k.v));
                    k.v[0] = k.v[4] = k.v[8] = sv[q];
                    blks.push_back(k);
                }
            }
        }
    }
    for (int v = 0; v < nV; ++v) {
        Blk k;
        k.key = (long)v * nV + v;
        std::memset(k.v, 0, sizeof(k.v));
        k.v[0] = k.v[4] = k.v[8] = c.mass[v];
        blks.push_back(k);
    }
    if (c.dbc_stiff) {  // Shell/DIRICHLET.h:72-93
        for (int v : c.dbc_idx) {
            Blk k;
            k.key = (long)v * nV + v;
            std::memset(k.v, 0, sizeof(k.v));
            k.v[0] = k.v[4] = k.v[8] = c.dbc_stiff * c.mass[v];
            blks.push_back(k);
        }
    }
    build_sys(blks, nV, S);
    // Project_DBC: rows/cols of DBC nodes -> identity; under the penalty only
    // the nodes that do not move (INC_POTENTIAL.h:387-394)
    const std::vector<unsigned char>& pm = c.dbc_stiff ? c.dbc_fixed : c.dbc;
    for (int r = 0; r < nV; ++r)
        for (int k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k)
            if (pm[r] || pm[S.col[k]])
                for (int i = 0; i < 3; ++i)
                    for (int j = 0; j < 3; ++j) S.val[9 * (size_t)k + 3 * i + j] = (r == S.col[k] && i == j);
}

}  // namespace

namespace {
struct ShellElastic {
    const double *IB, *lam, *mu;
    int nH;
    const int* hinge;
    const double* rest;
    double k;
};

// friction arguments of the step: mu, epsv2, fricIterAmt, compNodeRange, muComp
struct FrictionArgs {
    double mu, epsv2;
    int iters;
    int ncomp;
    const int* comp_range;
    const double* mu_comp;
};

int advance_impl(int nV, double* X, const double* X0, double* vel, const double* mass, int nT, const int* tets,
                 const double* IB, const double* vol, const double* lam, const double* mu, int nE, const int* tris,
                 const double* tvol, int nb, const int* bnode, int ne, const int* bedge, int nf, const int* btri,
                 int nDBC, const int* dbc_idx, const double* dbc_target, double h, double NewtonTol,
                 int withCollision, double dHat2, double* kappa, double thickness, double lin_rtol, int max_newton,
                 const double* body_force, int flow, orc_stats* st, const ShellElastic* se = nullptr,
                 const FrictionArgs* fa = nullptr)
{
    Ctx c;
    c.nE = nE;
    c.tris = tris;
    c.tvol = tvol;
    if (se) {
        c.tIB = se->IB;
        c.tlam = se->lam;
        c.tmu = se->mu;
        c.nH = se->nH;
        c.hinge = se->hinge;
        c.hrest = se->rest;
        c.bend_k = se->k;
    }
    c.nV = nV; c.nT = nT; c.nb = nb; c.ne = ne; c.nf = nf;
    c.X0 = X0; c.mass = mass; c.tets = tets; c.IB = IB; c.vol = vol; c.lam = lam; c.mu = mu;
    c.bnode = bnode; c.bedge = bedge; c.btri = btri;
    c.dbc.assign(nV, 0);
    for (int i = 0; i < nDBC; ++i) c.dbc[dbc_idx[i]] = 1;
    c.h = h; c.dHat2 = dHat2; c.thickness = thickness; c.kappa = kappa; c.withCollision = withCollision != 0;
    c.st = st;
    c.flow = flow != 0;
    if (st) std::memset(st, 0, sizeof(*st));
    const size_t n3 = 3 * (size_t)nV;
    std::vector<double> Xn(X, X + n3);
    c.Xn = Xn.data();
    if (fa && c.withCollision && (fa->mu > 0 || fa->ncomp > 0)) {
        c.fric = true;
        c.fmu = fa->mu;
        c.epsvh2 = fa->epsv2 * h * h;
        c.ncomp = fa->ncomp;
        c.comp_range = fa->comp_range;
        c.mu_comp = fa->mu_comp;
    }
    c.Xtilde.resize(n3);
    // IMPLICIT_EULER.h:196-217: a = M^-1 b with the lumped (diagonal) mass,
    // the flow variant zeroes v, x~ = x + (h v + h^2 a)
    for (size_t i = 0; i < n3; ++i) {
        if (c.flow) vel[i] = 0.0;
        const double a = body_force ? body_force[i] / mass[i / 3] : 0.0;
        c.Xtilde[i] = X[i] + (h * vel[i] + h * h * a);
    }

    // Dirichlet nodes toward their targets (IMPLICIT_EULER.h:304-401): the
    // CCD-limited fraction of the move; all of it when that is 1, otherwise
    // none, and the penalty (augmented Lagrangian) pulls them in the loop
    c.dbc_idx.assign(dbc_idx, dbc_idx + nDBC);
    c.dbc_tgt.resize(3 * (size_t)nDBC);
    c.dbc_fixed.assign(nV, 0);
    std::vector<double> disp(n3, 0.0);
    for (int i = 0; i < nDBC; ++i) {
        const int v = dbc_idx[i];
        for (int d = 0; d < 3; ++d) {
            c.dbc_tgt[3 * i + d] = dbc_target ? dbc_target[3 * i + d] : X[3 * v + d];
            disp[3 * v + d] = c.dbc_tgt[3 * i + d] - X[3 * v + d];
        }
        if (!(disp[3 * v] || disp[3 * v + 1] || disp[3 * v + 2])) c.dbc_fixed[v] = 1;
    }
    double DBCAlpha = 1.0;
    if (c.withCollision) {
        std::vector<unsigned char> dbcm = c.dbc;
        orc_ccd_stepsize(nV, X, disp.data(), dbcm.data(), nb, bnode, ne, bedge, nf, btri, thickness, 1.0, &DBCAlpha,
                         nullptr);
    }
    double DBCPenaltyXn = 0.0;
    if (DBCAlpha == 1.0) {
        for (int i = 0; i < nDBC; ++i)
            for (int d = 0; d < 3; ++d) X[3 * dbc_idx[i] + d] = c.dbc_tgt[3 * i + d];
    } else {
        c.dbc_stiff = 1e6;
        DBCPenaltyXn = dbc_dist2(c, Xn.data());
    }
    if (st) st->dbc_alpha = DBCAlpha;

    if (c.withCollision) compute_cs(c, X);
    if (c.fric && !friction_basis(c, X)) return ORC_E_INVALID_START;
    double Eprev;
    if (!energy(c, X, c.withCollision, Eprev)) return ORC_E_INVALID_START;
    if (c.dbc_stiff) dbc_energy(c, X, Eprev);
    std::deque<double> resRecord;
    std::vector<double> g(n3), rhs(n3), sol(n3), Xprev(n3), dist2, dist2_prev;
    std::vector<int> cs_prev;
    long ncs_prev = 0;
    bool useGD = false;
    int PNIter = 0, fricIterI = 0;
    double L2Norm = 0;
    // gradient (INC_POTENTIAL.h:154-275), the friction term, DBC rows zeroed
    // (IMPLICIT_EULER.h:456-491); rhs = -g
    auto gradient_at = [&]() {
        std::fill(g.begin(), g.end(), 0.0);
        if (!c.flow) orc_elastic_gradient(X, nT, tets, IB, vol, lam, mu, h * h, g.data());
        if (!c.flow && c.nE && c.tIB)
            orc_shell_elastic_terms(nV, X, c.nE, c.tris, c.tIB, c.tvol, c.tlam, c.tmu, c.nH, c.hinge, c.hrest,
                                    c.bend_k, h * h, c.dbc.data(), 1, nullptr, g.data(), nullptr, nullptr);
        if (c.flow && c.nE) {  // INC_POTENTIAL.h:191-211: g += -h Lx
            std::vector<double> Lx;
            flow_lx(c, X, Lx);
            for (size_t i = 0; i < n3; ++i) g[i] += -h * Lx[i];
        }
        for (int v = 0; v < nV; ++v)
            for (int d = 0; d < 3; ++d) g[3 * v + d] += mass[v] * (X[3 * v + d] - c.Xtilde[3 * v + d]);
        if (c.withCollision && c.ncs)
            orc_barrier_gradient(X, X0, c.ncs, c.cs.data(), dHat2, kappa[0], thickness, nV, g.data());
        if (c.fric && c.fmu > 0)
            orc_friction_gradient(X, c.Xn, c.nfc, c.fcs.data(), c.fcp.data(), c.fbasis.data(), c.fnf.data(),
                                  c.epsvh2, c.fmu, g.data());
        if (c.dbc_stiff) {  // Shell/DIRICHLET.h:50-70, then the fixed nodes' rows (IMPLICIT_EULER.h:467-474)
            for (int i = 0; i < nDBC; ++i) {
                const int v = dbc_idx[i];
                for (int d = 0; d < 3; ++d) g[3 * v + d] += c.dbc_stiff * mass[v] * (X[3 * v + d] - c.dbc_tgt[3 * i + d]);
            }
            for (int v = 0; v < nV; ++v)
                if (c.dbc_fixed[v]) g[3 * v] = g[3 * v + 1] = g[3 * v + 2] = 0;
        } else {
            for (int v = 0; v < nV; ++v)
                if (c.dbc[v]) g[3 * v] = g[3 * v + 1] = g[3 * v + 2] = 0;
        }
        for (size_t i = 0; i < n3; ++i) rhs[i] = -g[i];
    };
    do {
        double t0 = now();
        gradient_at();
        c.t_grad += now() - t0;

        if (!useGD) {
            t0 = now();
            Sys S;
            assemble_sys(c, X, S);
            c.t_hess += now() - t0;
            t0 = now();
            double relres = 0;
            int it = pcg(S, rhs.data(), sol.data(), lin_rtol, 100000, &relres);
            c.t_solve += now() - t0;
            if (st) {
                st->linear_iters += std::max(it, 0);
                st->last_relres = relres;
            }
            if (it < 0) {
                useGD = true;
                sol = rhs;
            }
        } else {
            sol = rhs;
        }

        // Line_Search (IMPLICIT_EULER.h:10-149)
        std::copy(X, X + n3, Xprev.begin());
        double alpha = 1.0, E = 0;
        bool valid;
        do {
            for (size_t i = 0; i < n3; ++i) X[i] = Xprev[i] + alpha * sol[i];
            valid = energy(c, X, false, E);
            if (!valid) alpha /= 2.0;
        } while (!valid);
        std::copy(Xprev.begin(), Xprev.end(), X);
        if (c.withCollision) {
            t0 = now();
            std::vector<unsigned char> dbcm = c.dbc;
            orc_ccd_stepsize(nV, X, sol.data(), dbcm.data(), nb, bnode, ne, bedge, nf, btri, thickness, alpha,
                             &alpha, nullptr);
            c.t_ccd += now() - t0;
        }
        double feasibleAlpha = alpha;
        if (st && st->n_newton == 0) st->first_ccd_alpha = alpha;
        do {
            for (size_t i = 0; i < n3; ++i) X[i] = Xprev[i] + alpha * sol[i];
            valid = energy(c, X, false, E);
            if (valid) {
                if (c.withCollision) {
                    compute_cs(c, X);
                    if (c.ncs) {
                        double md;
                        orc_min_dist2(X, c.ncs, c.cs.data(), thickness, nullptr, &md);
                        if (md <= 0) {
                            alpha /= 2;
                            E = Eprev + 1;
                            continue;
                        }
                    }
                }
                valid = energy(c, X, c.withCollision, E);
                if (valid && c.dbc_stiff) dbc_energy(c, X, E);
            }
            alpha /= 2.0;
            if (st) st->line_search_trials++;
        } while (E > Eprev || !valid);
        Eprev = E;
        if (st) st->last_alpha = alpha * 2;

        // kappa update (IMPLICIT_EULER.h:566-595)
        if (ncs_prev) {
            std::vector<double> cur(ncs_prev);
            double md;
            orc_min_dist2(X, ncs_prev, cs_prev.data(), thickness, cur.data(), &md);
            bool upd = false;
            for (long i = 0; i < ncs_prev; ++i)
                if (dist2_prev[i] < 1e-18 && cur[i] < dist2_prev[i]) {
                    upd = true;
                    break;
                }
            if (upd && kappa[0] < kappa[1]) {
                kappa[0] *= 2;
                energy(c, X, c.withCollision, Eprev);
                if (c.dbc_stiff) dbc_energy(c, X, Eprev);
            }
        }
        dist2.assign(c.ncs, 0.0);
        if (c.ncs) {
            double md;
            orc_min_dist2(X, c.ncs, c.cs.data(), thickness, dist2.data(), &md);
            if (st) st->min_dist2 = md;
        }
        cs_prev.assign(c.cs.begin(), c.cs.begin() + 4 * c.ncs);
        ncs_prev = c.ncs;
        dist2_prev = dist2;

        // stopping criterion (IMPLICIT_EULER.h:647-676)
        L2Norm = 0;
        for (size_t i = 0; i < n3; ++i) L2Norm += sol[i] * sol[i];
        L2Norm = std::sqrt(L2Norm / (nV - nDBC)) / h;
        ++PNIter;
        resRecord.push_back(L2Norm);
        if (resRecord.size() > 3) resRecord.pop_front();
        L2Norm = *std::max_element(resRecord.begin(), resRecord.end());
        if (useGD) L2Norm = NewtonTol * 10;
        if (alpha * 2 < 1e-8 && feasibleAlpha > 1e-8) useGD = true;
        else useGD = false;
        // moving DBC: raise the penalty near convergence until the nodes are
        // within 1% of their move, then drop it (IMPLICIT_EULER.h:701-755)
        if (c.dbc_stiff) {
            const double progress = 1 - std::sqrt(dbc_dist2(c, X) / DBCPenaltyXn);
            if (progress < 0.99) {
                if (L2Norm < NewtonTol * 10) {
                    if (c.dbc_stiff < 1e8) {
                        c.dbc_stiff *= 2;
                        energy(c, X, c.withCollision, Eprev);
                        dbc_energy(c, X, Eprev);
                    }
                }
                L2Norm = NewtonTol * 10;
            } else {
                c.dbc_stiff = 0;
                energy(c, X, c.withCollision, Eprev);
            }
        }
        // lagged friction update at convergence (IMPLICIT_EULER.h:757-848): a
        // new basis, then one more direction whose size decides the exit
        if (resRecord.size() == 3 && L2Norm <= NewtonTol) {
            ++fricIterI;
            if ((fricIterI < (fa ? fa->iters : 1) || (fa && fa->iters <= 0)) && c.fric) {
                if (!friction_basis(c, X)) return ORC_E_INVALID_START;
                gradient_at();
                Sys S;
                assemble_sys(c, X, S);
                double relres = 0;
                int it = pcg(S, rhs.data(), sol.data(), lin_rtol, 100000, &relres);
                if (it < 0) return ORC_E_DEGENERATE;  // the reference exits when the solve fails
                if (st) st->linear_iters += it;
                L2Norm = 0;
                for (size_t i = 0; i < n3; ++i) L2Norm += sol[i] * sol[i];
                L2Norm = std::sqrt(L2Norm / (nV - nDBC)) / h;
                if (L2Norm > NewtonTol) {
                    energy(c, X, c.withCollision, Eprev);
                    if (c.dbc_stiff) dbc_energy(c, X, Eprev);
                }
                if (st) st->friction_updates++;
            }
        }
        if (st) {
            st->n_newton = PNIter;
            st->n_constraints = c.ncs;
            st->n_friction = c.nfc;
            st->last_res = L2Norm;
            st->E = Eprev;
            st->dbc_stiff = c.dbc_stiff;
        }
        // Advance_One_Step_IE_Flow ends once no contact is left (IMPLICIT_EULER.h:850-854)
        if (c.flow && (!c.withCollision || c.ncs == 0)) break;
        if (max_newton > 0 && PNIter >= max_newton) break;
    } while (resRecord.size() < 3 || L2Norm > NewtonTol);

    for (size_t i = 0; i < n3; ++i) vel[i] = (X[i] - Xn[i]) / h;
    if (st) {
        st->t_constraint_set = c.t_cs;
        st->t_ccd = c.t_ccd;
        st->t_solve = c.t_solve;
        st->t_hessian = c.t_hess;
        st->t_gradient = c.t_grad;
        st->t_energy = c.t_E;
    }
    return PNIter;
}

// The Newton system at X: the projected Hessian (BSR, 3x3 blocks) and the
// gradient of the incremental potential with inertia target Xtilde, for the
// constraint set at X.  Returns nnzb, or -nnzb when cap is too small.
}  // namespace

extern "C" int orc_advance_one_step(int nV, double* X, const double* X0, double* vel, const double* mass,
                                    int nT, const int* tets, const double* IB, const double* vol,
                                    const double* lam, const double* mu, int nb, const int* bnode, int ne,
                                    const int* bedge, int nf, const int* btri, int nDBC, const int* dbc_idx,
                                    const double* dbc_target, double h, double NewtonTol, int withCollision,
                                    double dHat2, double* kappa, double thickness, double lin_rtol, int max_newton,
                                    const double* body_force, int flow, orc_stats* st)
{
    return advance_impl(nV, X, X0, vel, mass, nT, tets, IB, vol, lam, mu, 0, nullptr, nullptr, nb, bnode, ne, bedge,
                        nf, btri, nDBC, dbc_idx, dbc_target, h, NewtonTol, withCollision, dHat2, kappa, thickness,
                        lin_rtol, max_newton, body_force, flow, st);
}

extern "C" int orc_advance_shell_flow(int nV, double* X, const double* X0, double* vel, const double* mass, int nE,
                                      const int* tris, const double* tvol, int nb, const int* bnode, int ne,
                                      const int* bedge, int nf, const int* btri, int nDBC, const int* dbc_idx,
                                      double h, double NewtonTol, int withCollision, double dHat2, double* kappa,
                                      double thickness, double lin_rtol, int max_newton, const double* body_force,
                                      orc_stats* st)
{
    return advance_impl(nV, X, X0, vel, mass, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nE, tris, tvol, nb,
                        bnode, ne, bedge, nf, btri, nDBC, dbc_idx, nullptr, h, NewtonTol, withCollision, dHat2, kappa,
                        thickness, lin_rtol, max_newton, body_force, 1, st);
}

namespace {
long newton_system_impl(int nV, const double* X, const double* Xtilde, const double* X0, const double* mass, int nT,
                        const int* tets, const double* IB, const double* vol, const double* lam, const double* mu,
                        int nE, const int* tris, const double* tvol, bool flow, int nb, const int* bnode, int ne,
                        const int* bedge, int nf, const int* btri, int nDBC, const int* dbc_idx, double h,
                        double dHat2, double kappa, double thickness, int* rowptr, int* col, double* val, double* g,
                        long cap, const ShellElastic* se = nullptr)
{
    Ctx c;
    c.nV = nV; c.nT = nT; c.nb = nb; c.ne = ne; c.nf = nf;
    c.X0 = X0; c.mass = mass; c.tets = tets; c.IB = IB; c.vol = vol; c.lam = lam; c.mu = mu;
    c.nE = nE; c.tris = tris; c.tvol = tvol; c.flow = flow;
    if (se) {
        c.tIB = se->IB; c.tlam = se->lam; c.tmu = se->mu;
        c.nH = se->nH; c.hinge = se->hinge; c.hrest = se->rest; c.bend_k = se->k;
    }
    c.bnode = bnode; c.bedge = bedge; c.btri = btri;
    c.dbc.assign(nV, 0);
    for (int i = 0; i < nDBC; ++i) c.dbc[dbc_idx[i]] = 1;
    double kap[3] = {kappa, kappa, 0};
    c.h = h; c.dHat2 = dHat2; c.thickness = thickness; c.kappa = kap; c.withCollision = true;
    c.st = nullptr;
    c.Xtilde.assign(Xtilde, Xtilde + 3 * (size_t)nV);
    compute_cs(c, X);
    Sys S;
    assemble_sys(c, X, S);
    const long nnzb = (long)S.col.size();
    if (nnzb > cap) return -nnzb;
    std::copy(S.rowptr.begin(), S.rowptr.end(), rowptr);
    std::copy(S.col.begin(), S.col.end(), col);
    std::copy(S.val.begin(), S.val.end(), val);
    std::fill(g, g + 3 * (size_t)nV, 0.0);
    if (!flow) orc_elastic_gradient(X, nT, tets, IB, vol, lam, mu, h * h, g);
    if (!flow && nE && c.tIB)
        orc_shell_elastic_terms(nV, X, nE, tris, c.tIB, tvol, c.tlam, c.tmu, c.nH, c.hinge, c.hrest, c.bend_k, h * h,
                                c.dbc.data(), 1, nullptr, g, nullptr, nullptr);
    if (flow && nE) {
        std::vector<double> Lx;
        flow_lx(c, X, Lx);
        for (size_t i = 0; i < Lx.size(); ++i) g[i] += -h * Lx[i];
    }
    for (int v = 0; v < nV; ++v)
        for (int d = 0; d < 3; ++d) g[3 * v + d] += mass[v] * (X[3 * v + d] - Xtilde[3 * v + d]);
    if (c.ncs) orc_barrier_gradient(X, X0, c.ncs, c.cs.data(), dHat2, kappa, thickness, nV, g);
    for (int v = 0; v < nV; ++v)
        if (c.dbc[v]) g[3 * v] = g[3 * v + 1] = g[3 * v + 2] = 0;
    return nnzb;
}
}  // namespace

// Advance_One_Step_IE_Hinge on a shell: membrane + hinge bending + inertia +
// barrier, body force b, DBC nodes staying
extern "C" int orc_advance_shell(int nV, double* X, const double* X0, double* vel, const double* mass, int nE,
                                 const int* tris, const double* tvol, const double* tIB, const double* tlam,
                                 const double* tmu, int nH, const int* hinge, const double* hrest, double bend_k,
                                 int nb, const int* bnode, int ne, const int* bedge, int nf, const int* btri, int nDBC,
                                 const int* dbc_idx, double h, double NewtonTol, int withCollision, double dHat2,
                                 double* kappa, double thickness, double lin_rtol, int max_newton,
                                 const double* body_force, orc_stats* st)
{
    ShellElastic se{tIB, tlam, tmu, nH, hinge, hrest, bend_k};
    return advance_impl(nV, X, X0, vel, mass, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nE, tris, tvol, nb,
                        bnode, ne, bedge, nf, btri, nDBC, dbc_idx, nullptr, h, NewtonTol, withCollision, dHat2, kappa,
                        thickness, lin_rtol, max_newton, body_force, 0, st, &se);
}

extern "C" long orc_newton_system(int nV, const double* X, const double* Xtilde, const double* X0,
                                  const double* mass, int nT, const int* tets, const double* IB, const double* vol,
                                  const double* lam, const double* mu, int nb, const int* bnode, int ne,
                                  const int* bedge, int nf, const int* btri, int nDBC, const int* dbc_idx, double h,
                                  double dHat2, double kappa, double thickness, int* rowptr, int* col, double* val,
                                  double* g, long cap)
{
    return newton_system_impl(nV, X, Xtilde, X0, mass, nT, tets, IB, vol, lam, mu, 0, nullptr, nullptr, false, nb,
                              bnode, ne, bedge, nf, btri, nDBC, dbc_idx, h, dHat2, kappa, thickness, rowptr, col, val,
                              g, cap);
}

// the Hinge variant's system on a shell (membrane + bending + inertia + barrier)
extern "C" long orc_newton_system_shell(int nV, const double* X, const double* Xtilde, const double* X0,
                                        const double* mass, int nE, const int* tris, const double* tvol,
                                        const double* tIB, const double* tlam, const double* tmu, int nH,
                                        const int* hinge, const double* hrest, double bend_k, int nb,
                                        const int* bnode, int ne, const int* bedge, int nf, const int* btri, int nDBC,
                                        const int* dbc_idx, double h, double dHat2, double kappa, double thickness,
                                        int* rowptr, int* col, double* val, double* g, long cap)
{
    ShellElastic se{tIB, tlam, tmu, nH, hinge, hrest, bend_k};
    return newton_system_impl(nV, X, Xtilde, X0, mass, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nE, tris, tvol,
                              false, nb, bnode, ne, bedge, nf, btri, nDBC, dbc_idx, h, dHat2, kappa, thickness, rowptr,
                              col, val, g, cap, &se);
}

// the flow variant's system on a shell (flow term + inertia + barrier)
extern "C" long orc_newton_system_shell_flow(int nV, const double* X, const double* Xtilde, const double* X0,
                                             const double* mass, int nE, const int* tris, const double* tvol, int nb,
                                             const int* bnode, int ne, const int* bedge, int nf, const int* btri,
                                             int nDBC, const int* dbc_idx, double h, double dHat2, double kappa,
                                             double thickness, int* rowptr, int* col, double* val, double* g,
                                             long cap)
{
    return newton_system_impl(nV, X, Xtilde, X0, mass, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nE, tris, tvol,
                              true, nb, bnode, ne, bedge, nf, btri, nDBC, dbc_idx, h, dHat2, kappa, thickness, rowptr,
                              col, val, g, cap);
}

// Advance_One_Step_IE_Hinge / _Flow on a tet mesh with lagged friction: mu,
// epsv2 (epsv h is the static-friction speed bound), fric_iters
// (fricIterAmt: basis updates before the step may end; <= 0 until the
// direction after an update is below the tolerance), and per-component
// coefficients (ncomp > 0: compNodeRange / muComp, mu then 1)
extern "C" int orc_advance_friction(int nV, double* X, const double* X0, double* vel, const double* mass, int nT,
                                    const int* tets, const double* IB, const double* vol, const double* lam,
                                    const double* mu, int nb, const int* bnode, int ne, const int* bedge, int nf,
                                    const int* btri, int nDBC, const int* dbc_idx, const double* dbc_target, double h,
                                    double NewtonTol, int withCollision, double dHat2, double* kappa, double thickness,
                                    double lin_rtol, int max_newton, const double* body_force, int flow,
                                    double fmu, double epsv2, int fric_iters, int ncomp, const int* comp_range,
                                    const double* mu_comp, orc_stats* st)
{
    FrictionArgs fa{fmu, epsv2, fric_iters, ncomp, comp_range, mu_comp};
    return advance_impl(nV, X, X0, vel, mass, nT, tets, IB, vol, lam, mu, 0, nullptr, nullptr, nb, bnode, ne, bedge,
                        nf, btri, nDBC, dbc_idx, dbc_target, h, NewtonTol, withCollision, dHat2, kappa, thickness,
                        lin_rtol, max_newton, body_force, flow, st, nullptr, &fa);
}

// Advance_One_Step_IE_Hinge on a shell with lagged friction (the arguments of
// orc_advance_shell, then those of orc_advance_friction)
extern "C" int orc_advance_shell_friction(int nV, double* X, const double* X0, double* vel, const double* mass,
                                          int nE, const int* tris, const double* tvol, const double* tIB,
                                          const double* tlam, const double* tmu, int nH, const int* hinge,
                                          const double* hrest, double bend_k, int nb, const int* bnode, int ne,
                                          const int* bedge, int nf, const int* btri, int nDBC, const int* dbc_idx,
                                          double h, double NewtonTol, double dHat2, double* kappa, double thickness,
                                          double lin_rtol, int max_newton, const double* body_force, double fmu,
                                          double epsv2, int fric_iters, orc_stats* st)
{
    ShellElastic se{tIB, tlam, tmu, nH, hinge, hrest, bend_k};
    FrictionArgs fa{fmu, epsv2, fric_iters, 0, nullptr, nullptr};
    return advance_impl(nV, X, X0, vel, mass, 0, nullptr, nullptr, nullptr, nullptr, nullptr, nE, tris, tvol, nb,
                        bnode, ne, bedge, nf, btri, nDBC, dbc_idx, nullptr, h, NewtonTol, 1, dHat2, kappa, thickness,
                        lin_rtol, max_newton, body_force, 0, st, &se, &fa);
}
