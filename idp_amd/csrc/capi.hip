// C ABI and host orchestration of the Newton-with-barrier time step.
//   Advance_One_Step_IE_Discrete_Shell   Library/FEM/Shell/IMPLICIT_EULER.h:152-891
//   Line_Search                          Library/FEM/Shell/IMPLICIT_EULER.h:10-149
// State stays on the device for the whole step; the host drives control flow
// (line-search halving, kappa update, stopping rule) from scalars read back
// once per stage.
#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>
#include <set>
#include <map>

#include "common.hpp"
#include "kernels.hpp"
#include "linalg.hpp"

using namespace idp;

struct BatchEngine;

// Lagged friction of a context (idp_set_friction) and the friction set of its
// last basis update
struct FricState {
    double mu = 0.0, epsv2 = 0.0;
    int iters = 1, ncomp = 0;
    DBuf<int> comp_range;
    DBuf<double> mu_comp;
    DBuf<int4> fcs;              // friction constraints (int4 entries of the barrier set)
    DBuf<double> fdat;           // 10 doubles each: closest points, tangent basis, normal force
    DBuf<double> xn, part;       // x_n of the stage entry points; energy partials
    DBuf<unsigned> flag, pos;
    DBuf<int> err;
    long n = 0;
    bool have = false;           // a basis exists (stage entry points)
    // while a step or a stage evaluation uses it: mu (1 with per-component
    // coefficients), epsv h and x_n
    bool on = false;
    double mu_eff = 0.0, epsvh = 0.0;
    const double* Xn = nullptr;
    bool enabled() const { return mu > 0 || ncomp > 0; }
};

struct idp_ctx {
    int device = 0;
    hipStream_t s = nullptr;
    std::string err;
    MeshDev m;
    ContactWork cw;
    LinWork L;
    bool have_mesh = false;
    int ndbc = 0;
    // moving DBC targets (idp_set_dbc_targets): entries in the caller's order
    std::vector<int> dbc_list;
    std::vector<double> dbc_tgt;
    DBuf<int> dbc_idx_d;
    DBuf<double> dbc_tgt_d, dbc_x_d;
    DBuf<unsigned char> dbc_fixed_d;
    // state and work vectors (3 nV)
    DBuf<double> x, v, xt, xn, xprev, sol, rhs, g, tmpv, snap_x, snap_v, hx, hp;
    DBuf<double> gl, gb, Hl, Hb;  // per-tet / per-constraint local terms
    DBuf<double> part_e, part_i, part_b, part_f;
    DBuf<double> sform;                 // hinge Hessian records (two stages)
    DBuf<unsigned char> sform_ok;
    DBuf<int> flags;
    DBuf<unsigned char> one;  // a device byte holding 1 (warm-start mask of a one-mesh solve)
    DBuf<unsigned long long> dmin;
    DBuf<double> dist2, dist2_prev;
    DBuf<int4> cs_prev;
    long ncs_prev = 0;
    double snap_kappa = 0;
    bool snap_valid = false;
    Profiler prof;
    DBuf<int> bcounts;
    DBuf<int> hperm;          // barrier Hessian: constraint indices grouped by kind
    DBuf<unsigned> hkcnt;     // ... and the per-kind counts / cursors
    DBuf<double> hform;       // ... packed K and H_y per constraint between its two stages
    DBuf<unsigned char> hform_ok;
    int hess_team = 3;        // lanes per order-9 constraint in the projection stage
    SegDev cseg;             // constraint blocks per instance (energy)
    DBuf<double> bpar;       // per instance: dh[B], kappa[B], th2[B]
    DBuf<double> esum;       // per-instance partial-sum results
    BatchEngine* eng = nullptr;  // batched / served solves
    // Advance_One_Step_IE_Flow: the volumetric (tet) term is left out of the
    // energy, gradient and Hessian, and the shell triangles carry the flow
    // term (INC_POTENTIAL.h:53-73, 191-211, 323-339)
    bool no_elastic = false;
    DBuf<double> bforce;  // body force of the current step (3 nV)
    // magnitude of the normal-flow force recomputed from x before each step
    // that is given no body force (0: none), as the shell driver calls
    // Update_Normal_Flow_Neumann before each step (FEMDiscreteShellBase.py:269-273)
    double normal_flow = 0.0;
    SurfWork sw;          // device surface extraction
    double* pinned = nullptr;  // page-locked staging of target positions (3 nV)
    size_t pinned_n = 0;
    DBuf<double> target;
    // this device's address of the fixed scratch array of the allocation-free
    // entry points (resolved once per device when the first context is created)
    unsigned char* scratch = nullptr;
    FricState fr;
};

namespace {

double now_ms()
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ACCD step cap of a solve: 0 selects the library default, < 0 none
// target free vertices per coarse aggregate of a solve (0: no coarse level)
int coarse_from(const idp_solver_opts* o)
{
    const int v = (!o || o->coarse_verts == 0) ? IDP_COARSE_VERTS_DEFAULT : o->coarse_verts;
    return v > 0 ? v : 0;
}

int accd_cap_from(const idp_solver_opts* o)
{
    if (!o || o->accd_max_steps == 0) return IDP_ACCD_MAX_STEPS_DEFAULT;
    return o->accd_max_steps < 0 ? 0 : o->accd_max_steps;
}

template <class F>
int guarded(idp_ctx* c, F&& f)
{
    if (!c) return IDP_E_INVALID;
    try {
        return f();
    } catch (const HipError& e) {
        c->err = e.what();
        return e.code;
    } catch (const std::exception& e) {
        c->err = e.what();
        return IDP_E_INVALID;
    }
}

// ---------------------------------------------------------------------------
// host set-up helpers
// ---------------------------------------------------------------------------
inline double det3h(const double* m)
{
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// surface primitives (MESHIO.h:486-545, 769-835): triangles are the tet faces
// (0,2,1),(0,3,2),(0,1,3),(1,2,3) without an oppositely oriented twin, in
// lexicographic order of the stored vertex triple; edges take the orientation
// of their first boundary triangle and are listed lexicographically; nodes
// are the vertices with nonzero accumulated boundary area, ascending.
// This part: the triangles.
void tet_surface_host(int nT, const int* tets, std::vector<std::array<int, 3>>& bt)
{
    struct Face {
        int s[3];  // sorted vertex set
        int o[3];  // stored orientation
        int par;   // parity of o relative to s (0 even)
    };
    std::vector<Face> f;
    f.reserve(4 * (size_t)nT);
    const int loc[4][3] = {{0, 2, 1}, {0, 3, 2}, {0, 1, 3}, {1, 2, 3}};
    for (int t = 0; t < nT; ++t)
        for (auto& l : loc) {
            Face F;
            for (int k = 0; k < 3; ++k) F.o[k] = tets[4 * t + l[k]];
            int a = F.o[0], b = F.o[1], c = F.o[2];
            int s0 = std::min(a, std::min(b, c)), s2 = std::max(a, std::max(b, c)), s1 = a + b + c - s0 - s2;
            F.s[0] = s0;
            F.s[1] = s1;
            F.s[2] = s2;
            // even iff (a,b,c) is a rotation of (s0,s1,s2)
            bool even = (a == s0 && b == s1) || (b == s0 && c == s1) || (c == s0 && a == s1);
            F.par = even ? 0 : 1;
            f.push_back(F);
        }
    std::sort(f.begin(), f.end(), [](const Face& x, const Face& y) {
        if (x.s[0] != y.s[0]) return x.s[0] < y.s[0];
        if (x.s[1] != y.s[1]) return x.s[1] < y.s[1];
        return x.s[2] < y.s[2];
    });
    bt.clear();
    for (size_t i = 0; i < f.size();) {
        size_t j = i;
        bool has[2] = {false, false};
        while (j < f.size() && f[j].s[0] == f[i].s[0] && f[j].s[1] == f[i].s[1] && f[j].s[2] == f[i].s[2]) {
            has[f[j].par] = true;
            ++j;
        }
        for (size_t k = i; k < j; ++k)
            if (!has[f[k].par ^ 1]) bt.push_back({f[k].o[0], f[k].o[1], f[k].o[2]});
        i = j;
    }
    std::sort(bt.begin(), bt.end());
    bt.erase(std::unique(bt.begin(), bt.end()), bt.end());
}

// edges and nodes of a boundary triangle list, taken in its order
// (MESHIO.h:769-835)
void surface_prims_host(int nV, const double* X, const std::vector<std::array<int, 3>>& bt, std::vector<int>& bn,
                        std::vector<std::array<int, 2>>& be)
{
    // edges: first-seen orientation
    std::unordered_map<long long, std::array<int, 2>> em;
    em.reserve(bt.size() * 2);
    std::vector<double> area(nV, 0.0);
    for (auto& t : bt) {
        const int pr[3][2] = {{t[0], t[1]}, {t[1], t[2]}, {t[2], t[0]}};
        for (auto& p : pr) {
            long long key = (long long)std::min(p[0], p[1]) * nV + std::max(p[0], p[1]);
            if (!em.count(key)) em[key] = {p[0], p[1]};
        }
        double e1[3], e2[3];
        for (int d = 0; d < 3; ++d) {
            e1[d] = X[3 * t[1] + d] - X[3 * t[0] + d];
            e2[d] = X[3 * t[2] + d] - X[3 * t[0] + d];
        }
        double c0 = e1[1] * e2[2] - e1[2] * e2[1], c1 = e1[2] * e2[0] - e1[0] * e2[2], c2 = e1[0] * e2[1] - e1[1] * e2[0];
        double A = 0.5 * std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        for (int k = 0; k < 3; ++k) area[t[k]] += A / 3;
    }
    be.clear();
    be.reserve(em.size());
    for (auto& kv : em) be.push_back(kv.second);
    std::sort(be.begin(), be.end());
    bn.clear();
    for (int v = 0; v < nV; ++v)
        if (area[v] != 0.0) bn.push_back(v);
}

// surface primitives of a tet mesh with shell triangles appended after its
// surface triangles, as Advance_One_Step_IE_Discrete_Shell lists them
// (Shell/IMPLICIT_EULER.h:233-241); nE = 0 is a plain tet mesh
void surface_host(int nV, const double* X, int nT, const int* tets, int nE, const int* tris, std::vector<int>& bn,
                  std::vector<std::array<int, 2>>& be, std::vector<std::array<int, 3>>& bt)
{
    tet_surface_host(nT, tets, bt);
    for (int e = 0; e < nE; ++e) bt.push_back({tris[3 * e], tris[3 * e + 1], tris[3 * e + 2]});
    surface_prims_host(nV, X, bt, bn, be);
}

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
struct EnergyOut {
    bool valid;
    bool dist_ok;
    double E;
};

// per-instance sums of block partials: one block per instance, the partials of
// instance i at [pbeg[i], pbeg[i+1]) added in the fixed order of
// sum_partials_block (with one instance: the plain sum over all partials)
__global__ void __launch_bounds__(RB) k_seg_sums(const int* __restrict__ pbeg, const double* __restrict__ part,
                                                 double* __restrict__ out)
{
    __shared__ double sh[RB];
    const int i = blockIdx.x;
    const double s = sum_partials_block(part + pbeg[i], pbeg[i + 1] - pbeg[i], sh);
    if (threadIdx.x == 0) out[i] = s;
}

BarrierParams bparams(idp_ctx* c)
{
    const int B = c->m.inst.B;
    return BarrierParams{c->bpar.p, c->bpar.p + B, c->bpar.p + 2 * B, B > 1 ? c->m.inst.vinst.p : nullptr};
}

// per-instance barrier parameters from dHat2, kappa, thickness (host, B each)
void set_bparams(idp_ctx* c, const double* dHat2, const double* kappa, const double* thickness)
{
    const int B = c->m.inst.B;
    std::vector<double> h(3 * B);
    for (int i = 0; i < B; ++i) {
        h[i] = dHat2[i] + 2 * std::sqrt(dHat2[i]) * thickness[i];
        h[B + i] = kappa[i];
        h[2 * B + i] = thickness[i] * thickness[i];
    }
    c->bpar.upload(h.data(), h.size(), c->s);
}

void set_bparams1(idp_ctx* c, double dHat2, double kappa, double thickness)
{
    set_bparams(c, &dHat2, &kappa, &thickness);
}

// constraint block table from the per-instance constraint ranges
// Compute_Friction_Basis and, with per-component coefficients,
// Compute_Friction_Coef (Library/FEM/FRICTION.h:16-169) on the context's
// current constraint set at X (device)
void friction_basis(idp_ctx* c, const double* X, double dHat2, double kappa, double thickness)
{
    IDP_TIMER_FLAG("Compute_Friction_Basis");
    auto s = c->s;
    FricState& f = c->fr;
    int* err = f.err.ensure(1, s);
    IDP_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    const int tok = c->prof.begin(IDP_K_FRICTION_BASIS);
    f.n = fric_basis_update(s, c->L.tmp, c->cw.ncs, (const int*)c->cw.cs.p, X, dHat2 + 2 * std::sqrt(dHat2) * thickness,
                            thickness * thickness, kappa, f.ncomp, f.comp_range.p, f.mu_comp.p, f.flag, f.pos, f.fcs,
                            f.fdat, err);
    // bytes: the constraint entries, flags and positions, the vertices read, the records written
    c->prof.end(IDP_K_FRICTION_BASIS, tok, 24.0 * c->cw.ncs + (96.0 + 96.0) * f.n);
    int he = 0;
    IDP_HIP(hipMemcpyAsync(&he, err, sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    if (he) throw std::invalid_argument("friction: a contact vertex lies in no component of comp_range");
    f.have = true;
}

// the contact BSR structure of the current constraint set, followed by the
// friction set while friction is on
void build_contact_structure(idp_ctx* c)
{
    const bool fo = c->fr.on && c->fr.n > 0;
    contact_structure(c->L, c->s, c->m.nV, c->cw.ncs, (const int*)c->cw.cs.p, fo ? c->fr.n : 0,
                      fo ? (const int*)c->fr.fcs.p : nullptr);
}

void update_cseg(idp_ctx* c)
{
    if (c->cw.h_cbeg.size() != (size_t)c->m.inst.B + 1) c->cw.h_cbeg = {0, c->cw.ncs};
    c->cseg.build(c->cw.h_cbeg, RB, c->s);
}

// Shell triangles carry the flow term in the flow variant and the membrane +
// hinge bending otherwise; a non-flow evaluation of a mesh with shell
// triangles whose elasticity was never set (idp_set_shell_elasticity) is
// refused rather than run without those terms.
void shell_terms_check(const idp_ctx* c)
{
    if (c->m.nE > 0 && !c->no_elastic && !c->m.shell_el)
        throw std::invalid_argument("shell triangles: the non-flow step needs idp_set_shell_elasticity");
}

// Compute_IncPotential per instance: elastic (w = h^2) + inertia + barrier
// over the current constraint set (barrier parameters from c->bpar)
void energies(idp_ctx* c, const double* X, bool with_barrier, double h, std::vector<EnergyOut>& out)
{
    IDP_TIMER_FLAG("Compute_IncPotential");
    auto s = c->s;
    auto& m = c->m;
    const InstDev& I = m.inst;
    const int B = I.B;
    int* fl = c->flags.ensure(2 * (size_t)B + 4, s);
    IDP_HIP(hipMemsetAsync(fl, 0, 2 * B * sizeof(int), s));
    const int tok = c->prof.begin(IDP_K_ENERGY);
    double* pe = c->part_e.ensure(I.tblk.nblk + 1, s);
    double* pi = c->part_i.ensure(I.vblk.nblk + 1, s);
    shell_terms_check(c);
    if (c->no_elastic) IDP_HIP(hipMemsetAsync(pe, 0, (I.tblk.nblk + 1) * sizeof(double), s));
    else launch_tet_energy(s, m, I.tblk, X, h * h, pe, fl);
    launch_inertia_energy(s, I.vblk, X, c->xt.p, m.mass.p, pi);
    const bool bar = with_barrier && c->cw.ncs > 0;
    const bool flow = c->no_elastic && m.nE > 0;
    const bool shell = !c->no_elastic && m.nE > 0;  // membrane + bending
    double* es = c->esum.ensure(6 * (size_t)B, s);
    const bool fric = with_barrier && c->fr.on && c->fr.n > 0;  // single-mesh contexts only
    IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.tblk.pbeg.p, pe, es);
    IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.vblk.pbeg.p, pi, es + B);
    if (flow) {
        double* pf = c->part_f.ensure(I.sblk.nblk + 1, s);
        launch_flow_energy(s, m, X, h, pf);
        IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.sblk.pbeg.p, pf, es + 3 * B);
    }
    if (shell) {
        double* pm = c->part_f.ensure(I.sblk.nblk + I.hblk.nblk + 2, s);
        double* ph = pm + I.sblk.nblk + 1;
        launch_shell_energy(s, m, X, h * h, pm, ph);
        IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.sblk.pbeg.p, pm, es + 3 * B);
        IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.hblk.pbeg.p, ph, es + 4 * B);
    }
    if (bar) {
        double* pb = c->part_b.ensure(c->cseg.nblk + 1, s);
        launch_barrier_energy(s, c->cseg.nblk, c->cseg.lo.p, c->cseg.hi.p, (const int*)c->cw.cs.p, X, m.X0.p,
                              bparams(c), pb, fl + B);
        IDP_LAUNCH(k_seg_sums, B, RB, 0, s, c->cseg.pbeg.p, pb, es + 2 * B);
    }
    if (fric) {  // Compute_Friction_Potential (FRICTION.h:171-243)
        const int nb = grid_for(c->fr.n, RB);
        double* pf = c->fr.part.ensure(nb + 1, s);
        launch_fric_energy(s, c->fr.n, c->fr.fcs.p, c->fr.fdat.p, X, c->fr.Xn, c->fr.epsvh, pf);
        launch_sum_partials(s, nb, pf, es + 5 * B);
    }
    c->prof.end(IDP_K_ENERGY, tok);
    std::vector<int> hf(2 * B);
    std::vector<double> he(6 * B, 0.0);
    IDP_HIP(hipMemcpyAsync(hf.data(), fl, 2 * B * sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipMemcpyAsync(he.data(), es, (bar ? 3 : 2) * B * sizeof(double), hipMemcpyDeviceToHost, s));
    if (flow || shell)
        IDP_HIP(hipMemcpyAsync(he.data() + 3 * B, es + 3 * B, (shell ? 2 : 1) * B * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    if (fric) IDP_HIP(hipMemcpyAsync(he.data() + 5 * B, es + 5 * B, sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    out.resize(B);
    for (int i = 0; i < B; ++i) {
        EnergyOut& o = out[i];
        o.valid = hf[i] == 0;
        o.dist_ok = hf[B + i] == 0;
        o.E = o.valid ? ((he[i] + he[3 * B + i]) + he[4 * B + i]) + he[B + i] + (bar ? he[2 * B + i] : 0.0) : 0.0;
        if (o.valid && fric) o.E += c->fr.mu_eff * he[5 * B + i];  // IMPLICIT_EULER.h:137-139
    }
}

// one instance (plain mesh)
EnergyOut energy(idp_ctx* c, const double* X, bool with_barrier, double h)
{
    std::vector<EnergyOut> o;
    energies(c, X, with_barrier, h, o);
    return o[0];
}

double min_dist2_dev(idp_ctx* c, const double* X, long n, const int* cs, double* dist2)
{
    IDP_TIMER_FLAG("Compute_Min_Dist");
    auto s = c->s;
    unsigned long long* dm = c->dmin.ensure(1, s);
    unsigned long long init = 0x7ff0000000000000ull;  // +inf
    IDP_HIP(hipMemcpyAsync(dm, &init, sizeof(init), hipMemcpyHostToDevice, s));
    launch_min_dist2(s, n, cs, X, nullptr, dist2, dm);
    unsigned long long r;
    IDP_HIP(hipMemcpyAsync(&r, dm, sizeof(r), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    double d;
    std::memcpy(&d, &r, sizeof(d));
    return d;
}

// gradient at X with the current constraint set (structure must be built);
// act (device, one byte per instance; null = all): the instances it is formed for
void gradient(idp_ctx* c, const double* X, double h, bool with_barrier, double* g, double* rhs,
              const unsigned char* act = nullptr)
{
    IDP_TIMER_FLAG("Compute_IncPotential_Gradient");
    auto s = c->s;
    auto& m = c->m;
    shell_terms_check(c);
    double* gl = c->gl.ensure(12 * ((size_t)m.nT + m.nE + m.nH), s);
    const int tok = c->prof.begin(IDP_K_GRADIENT);
    if (c->no_elastic) IDP_HIP(hipMemsetAsync(gl, 0, 12 * (size_t)m.nT * sizeof(double), s));
    else launch_tet_grad(s, m, X, h * h, gl, act);
    if (c->no_elastic) {  // shell triangles: the flow term; hinges: nothing
        launch_flow_grad(s, m, X, h, gl);
        if (m.nH)
            IDP_HIP(hipMemsetAsync(gl + 12 * ((size_t)m.nT + m.nE), 0, 12 * (size_t)m.nH * sizeof(double), s));
    } else if (m.nE) {
        launch_shell_grad(s, m, X, h * h, gl);  // membrane and hinge bending
    }
    const double* gb = nullptr;
    long ncs = c->cw.ncs;
    const long nfr = (with_barrier && c->fr.on) ? c->fr.n : 0;
    if (with_barrier && (ncs || nfr)) {
        double* b = c->gb.ensure(12 * (size_t)(ncs + nfr), s);
        if (ncs) {
            const int B = m.inst.B;
            int* fl = c->flags.ensure(4 * (size_t)B + 4, s) + 2 * B;
            IDP_HIP(hipMemsetAsync(fl, 0, B * sizeof(int), s));
            BarrierParams bp = bparams(c);
            bp.act = act;
            launch_barrier_grad(s, ncs, (const int*)c->cw.cs.p, X, m.X0.p, bp, b, fl);
        }
        if (nfr)  // Compute_Friction_Gradient (FRICTION.h:245-383), after the barrier slots
            launch_fric_grad(s, nfr, c->fr.fcs.p, c->fr.fdat.p, X, c->fr.Xn, c->fr.epsvh, c->fr.mu_eff, b + 12 * ncs);
        gb = b;
    }
    contact_gradient_gather(c->L, s, m, X, c->xt.p, gl, gb, g, rhs, act);
    c->prof.end(IDP_K_GRADIENT, tok);
}

// Hessian and assembly at X; act as for gradient(), nT_act: tets of those instances
void hessian(idp_ctx* c, const double* X, double h, bool with_barrier, const unsigned char* act = nullptr,
             long nT_act = -1)
{
    IDP_TIMER_FLAG("Compute_IncPotential_Hessian");
    auto s = c->s;
    auto& m = c->m;
    shell_terms_check(c);
    double* Hl = c->Hl.ensure(144 * ((size_t)m.nT + m.nE + m.nH), s);
    int tok = c->prof.begin(IDP_K_TET_HESS);
    if (c->no_elastic) {
        IDP_HIP(hipMemsetAsync(Hl, 0, 144 * (size_t)m.nT * sizeof(double), s));
        launch_flow_hess(s, m, h, Hl);  // shell triangles: the flow term; hinges: nothing
        if (m.nH)
            IDP_HIP(hipMemsetAsync(Hl + 144 * ((size_t)m.nT + m.nE), 0, 144 * (size_t)m.nH * sizeof(double), s));
    } else {
        IDP_TIMER_FLAG("Compute_Volumetric_Elasticity_Hessian");
        launch_tet_hess(s, m, X, h * h, 1, Hl, act);
        if (m.nE)
            launch_shell_hess(s, m, X, h * h, 1, c->sform.ensure(90 * (size_t)std::max(m.nH, 1), s),
                              c->sform_ok.ensure(std::max(m.nH, 1), s), Hl);
    }
    c->prof.end(IDP_K_TET_HESS, tok, 6500.0 * (nT_act < 0 ? m.nT : nT_act));
    const double* Hb = nullptr;
    long ncs = c->cw.ncs;
    const long nfr = (with_barrier && c->fr.on) ? c->fr.n : 0;
    if (with_barrier && nfr) {  // Compute_Friction_Hessian (FRICTION.h:385-660), after the barrier slots
        double* b = c->Hb.ensure(144 * (size_t)(ncs + nfr), s);
        tok = c->prof.begin(IDP_K_FRICTION);
        launch_fric_hess(s, nfr, c->fr.fcs.p, c->fr.fdat.p, X, c->fr.Xn, c->fr.epsvh, c->fr.mu_eff, 1, 1,
                         b + 144 * ncs);
        // bytes: entry + record, x and x_n of 4 vertices, 16 blocks (PT / EE bound)
        c->prof.end(IDP_K_FRICTION, tok, (96.0 + 192.0 + 1152.0) * nfr);
        Hb = b;
    }
    if (with_barrier && ncs) {
        const int B = m.inst.B;
        int* fl = c->flags.ensure(4 * (size_t)B + 4, s) + 3 * B;
        double* b = c->Hb.ensure(144 * (size_t)(ncs + nfr), s);
        IDP_TIMER_FLAG("Compute_Barrier_Hessian");
        tok = c->prof.begin(IDP_K_BARRIER_HESS);
        BarrierParams bp = bparams(c);
        bp.act = act;
        launch_barrier_hess(s, ncs, (const int*)c->cw.cs.p, X, m.X0.p, bp, 1, 1, b, nullptr, fl,
                            c->hform.ensure(90 * (size_t)ncs, s), c->hform_ok.ensure(ncs, s),
                            c->prof.full() ? c->bcounts.p : nullptr, c->hperm.ensure(ncs, s),
                            c->hkcnt.ensure(barrier_kind_scratch(ncs), s), c->hess_team);
        c->prof.end(IDP_K_BARRIER_HESS, tok);
        Hb = b;
    }
    assemble(c->L, s, m, Hl, Hb, c->ndbc > 0, act);
}

void set_constraints_from_host(idp_ctx* c, long n, const int* cs)
{
    c->cw.cs.upload((const int4*)cs, n, c->s);
    c->cw.ncs = n;
    c->cw.h_cbeg = {0, n};
    update_cseg(c);
}

const double* stage_x(idp_ctx* c, const double* X)
{
    if (X) {
        c->hx.upload(X, 3 * (size_t)c->m.nV, c->s);
        return c->hx.p;
    }
    return c->x.p;
}

// ---------------------------------------------------------------------------
// per-instance vector helpers of the batched Newton loop (mask: instances
// that take part; the arithmetic per element is that of the plain loop)
// ---------------------------------------------------------------------------
// v = (x_target - x) / h per coordinate (Load_Velocity_X0, MESHIO.h:987-1005)
__global__ void k_load_velocity(long n3, const double* __restrict__ x, const double* __restrict__ tgt, double h,
                                double* __restrict__ v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n3) v[i] = (tgt[i] - x[i]) / h;
}

// ---- moving Dirichlet nodes (Shell/DIRICHLET.h) --------------------------
// x of the DBC entries, in entry order
// x~ = x + (h v + h^2 a), a = M^-1 b with the lumped mass; the flow variant
// zeroes v first (IMPLICIT_EULER.h:196-217)
__global__ void k_xtilde(long n3, const double* __restrict__ x, double* __restrict__ v,
                         const double* __restrict__ b, const double* __restrict__ mass, double h, int flow,
                         double* __restrict__ xt)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n3) return;
    if (flow) v[i] = 0.0;
    const double a = b ? b[i] / mass[i / 3] : 0.0;
    xt[i] = x[i] + (h * v[i] + h * h * a);
}

__global__ void k_dbc_gather(int n, const int* __restrict__ idx, const double* __restrict__ X, double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    out[3 * i] = X[3 * v];
    out[3 * i + 1] = X[3 * v + 1];
    out[3 * i + 2] = X[3 * v + 2];
}

// g += k m (x - x_D) for the DBC entries that move (DIRICHLET.h:50-70; the
// rows of the nodes that stay are zero), rhs = -g
__global__ void k_dbc_grad(int n, const int* __restrict__ idx, const double* __restrict__ tgt,
                           const double* __restrict__ mass, const unsigned char* __restrict__ fixed,
                           const double* __restrict__ X, double k, double* __restrict__ g, double* __restrict__ rhs)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    if (fixed[v]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        const double gv = g[3 * v + d] + k * mass[v] * (X[3 * v + d] - tgt[3 * i + d]);
        g[3 * v + d] = gv;
        rhs[3 * v + d] = -gv;
    }
}

// x_D of every DBC entry (the whole move fits, IMPLICIT_EULER.h:385-396)
__global__ void k_dbc_set(int n, const int* __restrict__ idx, const double* __restrict__ tgt, double* __restrict__ X)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    X[3 * v] = tgt[3 * i];
    X[3 * v + 1] = tgt[3 * i + 1];
    X[3 * v + 2] = tgt[3 * i + 2];
}

// Hessian diagonal += k m for the DBC entries that move (DIRICHLET.h:72-93)
__global__ void k_dbc_hess(int n, const int* __restrict__ idx, const double* __restrict__ mass,
                           const unsigned char* __restrict__ fixed, const int* __restrict__ ediag, long eld, double k,
                           double* __restrict__ ev)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = idx[i];
    if (fixed[v]) return;
    const long b = ediag[v];
    const double a = k * mass[v];
    ev[0 * eld + b] += a;
    ev[4 * eld + b] += a;
    ev[8 * eld + b] += a;
}

__global__ void k_axpy_inst(int nV, const int* __restrict__ vinst, const unsigned char* __restrict__ mask,
                            const double* __restrict__ a, const double* __restrict__ alpha,
                            const double* __restrict__ b, double* __restrict__ x)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    const int i = vinst[v];
    if (mask && !mask[i]) return;
    const double al = alpha[i];
#pragma unroll
    for (int d = 0; d < 3; ++d) x[3 * v + d] = a[3 * v + d] + al * b[3 * v + d];
}

__global__ void k_copy_inst(int nV, const int* __restrict__ vinst, const unsigned char* __restrict__ mask,
                            const double* __restrict__ src, double* __restrict__ dst)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    if (mask && !mask[vinst[v]]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) dst[3 * v + d] = src[3 * v + d];
}

__global__ void k_velocity_inst(int nV, const int* __restrict__ vinst, const unsigned char* __restrict__ mask,
                                const double* __restrict__ x, const double* __restrict__ xn, double h,
                                double* __restrict__ v)
{
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nV) return;
    if (mask && !mask[vinst[u]]) return;
#pragma unroll
    for (int d = 0; d < 3; ++d) v[3 * u + d] = (x[3 * u + d] - xn[3 * u + d]) / h;
}

// partial sums of a.b over blocks [lo, hi) (one instance per block)
__global__ void __launch_bounds__(RB) k_dot_parts_seg(const int* __restrict__ lo, const int* __restrict__ hi,
                                                      const double* __restrict__ a, const double* __restrict__ b,
                                                      double* __restrict__ part)
{
    __shared__ double sh[RB];
    const int i = lo[blockIdx.x] + threadIdx.x;
    const double d = (i < hi[blockIdx.x]) ? a[i] * b[i] : 0.0;
    const double s = block_sum<RB>(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// kappa test (IMPLICIT_EULER.h:566-595) per instance: a pair of the previous
// constraint set whose squared distance was below 1e-18 and shrank
__global__ void k_kappa_test(long n, const int4* __restrict__ cs, const int* __restrict__ vinst,
                             const double* __restrict__ cur, const double* __restrict__ prev, int* __restrict__ flag)
{
    const long j = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    if (prev[j] < 1e-18 && cur[j] < prev[j]) {
        const int4 c = cs[j];
        const int v = c.x >= 0 ? c.x : -c.x - 1;  // every encoding has a vertex of its instance in slot 0
        flag[vinst[v]] = 1;
    }
}

}  // namespace

// ---------------------------------------------------------------------------
// Batched Newton engine: every stage runs over all active instances at once;
// each instance keeps the state of its own Newton loop (IMPLICIT_EULER.h:
// 152-891) and takes part in a stage only when its own loop would run it.
// start() begins solves in some instances while others are mid-solve, and
// iterate() advances every active instance by one Newton iteration, so
// instances can join and leave between iterations (continuous batching).
// ---------------------------------------------------------------------------
// ACCD steps a pair may take in one engine iteration before it is handed over
// to the next (about 55 ms of one lane's time at 0.85 us per step)
constexpr int ACCD_STEP_BUDGET = 1 << 16;

struct BatchEngine {
    bool configured = false;
    double h = 0, thickness = 0;
    bool coll = true;
    double rtol = 1e-8;
    int maxit = 10000, max_newton = 0;
    int B = 0, n_active = 0;
    std::vector<double> kappa, kmax, tol, dHat2, th, Eprev, alpha, feasible, E;
    std::vector<std::deque<double>> res;
    std::vector<char> useGD, active, finished_now;
    std::vector<char> pend;   // a CCD pair of this instance was handed over to the next call
    std::vector<char> ppend;  // its PCG solve was handed over to the next call
    std::vector<int> PN, ndbc;
    std::vector<idp_step_stats> S;
    std::vector<unsigned char> mask;
    DBuf<unsigned char> dmask, dmask2, dmask3;
    bool cold = false;  // every PCG solve from zero
    int accd_budget = ACCD_STEP_BUDGET;  // 0: no hand-over
    bool pcg_handover = true;
    DBuf<double> dalpha;
    std::vector<EnergyOut> eo;
    // stage wall times (IDP_STAGE_LOG=1: the stream is drained at each stage
    // boundary, so each stage is charged its own GPU work)
    static constexpr int NSTAGE = 7;
    bool stage_log = false;
    double stage_ms[NSTAGE] = {};
    double stage_t = 0;
    void stage(int k)
    {
        if (!stage_log) return;
        IDP_HIP(hipStreamSynchronize(stream));
        const double t = now_ms();
        if (k >= 0) stage_ms[k] += t - stage_t;
        stage_t = t;
    }

    int configure(idp_ctx* c, double h_, bool coll_, double th_, const idp_solver_opts* opts)
    {
        B = c->m.inst.B;
        stream = c->s;
        h = h_;
        coll = coll_;
        thickness = th_;
        rtol = (opts && opts->pcg_rtol > 0) ? opts->pcg_rtol : 1e-8;
        maxit = (opts && opts->pcg_maxit > 0) ? opts->pcg_maxit : 10000;
        max_newton = opts ? opts->max_newton : 0;
        c->L.precond = opts ? opts->precond : 0;
        cold = opts && opts->pcg_cold_start;
        accd_budget = !opts || opts->accd_budget == 0 ? ACCD_STEP_BUDGET : (opts->accd_budget < 0 ? 0 : opts->accd_budget);
        pcg_handover = !opts || opts->pcg_handover >= 0;
        c->cw.accd_max_steps = accd_cap_from(opts);
        const char* sl = std::getenv("IDP_STAGE_LOG");
        stage_log = sl && std::atoi(sl) > 0;
        std::fill(stage_ms, stage_ms + NSTAGE, 0.0);
        kappa.assign(B, 0);
        kmax.assign(B, 0);
        tol.assign(B, 0);
        dHat2.assign(B, 1e-6);
        th.assign(B, thickness);
        Eprev.assign(B, 0);
        alpha.assign(B, 1);
        feasible.assign(B, 1);
        E.assign(B, 0);
        res.assign(B, {});
        useGD.assign(B, 0);
        pend.assign(B, 0);
        ppend.assign(B, 0);
        active.assign(B, 0);
        finished_now.assign(B, 0);
        PN.assign(B, 0);
        S.assign(B, idp_step_stats{});
        mask.assign(B, 0);
        n_active = 0;
        const long n3 = 3L * c->m.nV;
        auto s = c->s;
        c->xn.ensure(n3, s);
        c->xt.ensure(n3, s);
        c->xprev.ensure(n3, s);
        c->sol.ensure(n3, s);
        c->rhs.ensure(n3, s);
        c->g.ensure(n3, s);
        ndbc.assign(B, 0);
        std::vector<unsigned char> hd(c->m.nV);
        c->m.dbc.download(hd.data(), c->m.nV, s);
        IDP_HIP(hipStreamSynchronize(s));
        const InstDev& I = c->m.inst;
        for (int i = 0; i < B; ++i)
            for (long v = I.vbeg[i]; v < I.vbeg[i + 1]; ++v) ndbc[i] += hd[v];
        c->ncs_prev = 0;
        c->cw.h_cbeg.assign(B + 1, 0);
        c->cw.ncs = 0;
        update_cseg(c);
        configured = true;
        return IDP_OK;
    }

    hipStream_t stream = nullptr;
    const unsigned char* up_mask() { return dmask.upload_get(mask.data(), B, stream); }

    // Begin solves in the listed instances (their x and v are set): x_n = x,
    // x~ = x + h v, constraint set and energy at x.  Arrays are indexed by
    // instance (B entries); kappa3 holds 3 per instance.
    int start(idp_ctx* c, const std::vector<int>& list, const double* tol_, const double* dHat2_,
              const double* kappa3, bool by_instance = true)
    {
        auto s = c->s;
        MeshDev& m = c->m;
        const InstDev& I = m.inst;
        (void)by_instance;
        std::fill(mask.begin(), mask.end(), 0);
        for (int i : list) {
            mask[i] = 1;
            active[i] = 1;
            ++n_active;
            tol[i] = tol_[i];
            dHat2[i] = dHat2_[i];
            kappa[i] = kappa3[3 * i];
            kmax[i] = kappa3[3 * i + 1];
            res[i].clear();
            useGD[i] = 0;
            pend[i] = 0;
            ppend[i] = 0;
            PN[i] = 0;
            std::memset(&S[i], 0, sizeof(idp_step_stats));
        }
        set_bparams(c, dHat2.data(), kappa.data(), th.data());
        const unsigned char* dm = dmask.upload_get(mask.data(), B, s);
        const int vgrid = grid_for(m.nV, 256);
        // Xn = X; Xtilde = X + h v (zero body force, IMPLICIT_EULER.h:196-217)
        IDP_LAUNCH(k_copy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, dm, c->x.p, c->xn.p);
        std::vector<double> hv(B, h);
        IDP_LAUNCH(k_axpy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, dm, c->x.p, dalpha.upload_get(hv.data(), B, s),
                   c->v.p, c->xt.p);
        if (coll) cset(c, dm);
        energies(c, c->x.p, coll, h, eo);
        for (int i : list) {
            if (!eo[i].valid) return IDP_E_INVALID_START;
            if (!eo[i].dist_ok) return IDP_E_NONPOS_DIST;
            Eprev[i] = eo[i].E;
        }
        return IDP_OK;
    }

    void cset(idp_ctx* c, const unsigned char* dm)
    {
        double ta = now_ms();
        const int tk = c->prof.begin(IDP_K_CONSTRAINT_SET);
        contact_constraint_set(c->cw, c->s, c->m, c->x.p, dHat2.data(), thickness, dm);
        update_cseg(c);
        c->prof.end(IDP_K_CONSTRAINT_SET, tk);
        const double dt = now_ms() - ta;
        for (int i = 0; i < B; ++i)
            if (mask[i]) {
                S[i].constraint_sets++;
                S[i].ms_constraint_set += dt;
            }
    }

    // per-instance min dist^2 over the current constraint set
    void min_dist(idp_ctx* c, double* d2out, std::vector<double>& md)
    {
        auto s = c->s;
        unsigned long long* dm = c->dmin.ensure(B, s);
        std::vector<unsigned long long> init(B, 0x7ff0000000000000ull);
        IDP_HIP(hipMemcpyAsync(dm, init.data(), B * 8, hipMemcpyHostToDevice, s));
        if (c->cw.ncs)
            launch_min_dist2(s, c->cw.ncs, (const int*)c->cw.cs.p, c->x.p, c->m.inst.vinst.p, d2out, dm);
        std::vector<unsigned long long> r(B);
        IDP_HIP(hipMemcpyAsync(r.data(), dm, B * 8, hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        md.resize(B);
        for (int i = 0; i < B; ++i) std::memcpy(&md[i], &r[i], 8);
    }

    bool has_constraints(idp_ctx* c, int i) const
    {
        return c->cw.h_cbeg.size() == (size_t)B + 1 && c->cw.h_cbeg[i + 1] > c->cw.h_cbeg[i];
    }

    // One Newton iteration of every active instance.  An instance whose CCD
    // handed a pair over (ACCD step budget) waits: it continues that pair in
    // the next call and takes its line search once the step size is final,
    // while the other instances go on.  Handing over changes no arithmetic.
    int iterate(idp_ctx* c)
    {
        auto s = c->s;
        MeshDev& m = c->m;
        const InstDev& I = m.inst;
        const int vgrid = grid_for(m.nV, 256);
        double* x = c->x.p;
        double* xprev = c->xprev.p;
        double* sol = c->sol.p;
        double* rhs = c->rhs.p;
        double* g = c->g.p;
        std::fill(finished_now.begin(), finished_now.end(), 0);
        // fresh: instances that begin a Newton iteration now
        // (instances with a handed-over PCG solve or CCD pair continue that first)
        std::vector<char> fresh(B, 0), was_ppend(B, 0), was_pend(B, 0);
        int nfresh = 0;
        bool anyres = false, anycont = false;
        for (int i = 0; i < B; ++i) {
            fresh[i] = active[i] && !pend[i] && !ppend[i];
            was_ppend[i] = active[i] && ppend[i];
            was_pend[i] = active[i] && pend[i];
            nfresh += fresh[i];
            anyres |= was_pend[i] != 0;
            anycont |= was_ppend[i] != 0;
        }
        // ---- gradient and Hessian at x with the current constraint sets, formed
        // for the instances that iterate now (a continued PCG solve's system is
        // formed again: its constraint and contact-block positions may move)
        for (int i = 0; i < B; ++i) mask[i] = fresh[i] || was_ppend[i];
        const unsigned char* dact = up_mask();
        stage(-1);
        double t0 = now_ms();
        if (coll) contact_structure(c->L, s, m.nV, c->cw.ncs, (const int*)c->cw.cs.p);
        else c->L.ncb = 0;
        if (nfresh) gradient(c, x, h, coll, g, rhs, dact);
        bool anyN = anycont;
        for (int i = 0; i < B; ++i) anyN |= fresh[i] && !useGD[i];
        std::vector<int> its(B, 0);
        std::vector<double> rel(B, 0.0);
        if (anyN) {
            std::vector<unsigned char> sm(B);
            std::vector<unsigned char> wm(B);
            long nT_act = 0;
            for (int i = 0; i < B; ++i) {
                sm[i] = fresh[i] && !useGD[i] ? 1 : (was_ppend[i] ? 2 : 0);
                wm[i] = sm[i] == 1 && PN[i] > 0 && !cold;  // sol holds this step's previous direction
                if (sm[i]) nT_act += I.tbeg[i + 1] - I.tbeg[i];
            }
            hessian(c, x, h, coll, dmask2.upload_get(sm.data(), B, s), nT_act);
            stage(0);
            pcg_solve_batch(c->L, s, m, rhs, sol, rtol, maxit, sm.data(), its.data(), rel.data(),
                            dmask3.upload_get(wm.data(), B, s), pcg_handover);
            for (int i = 0; i < B; ++i) {
                if (!sm[i]) continue;
                ppend[i] = its[i] == PCG_HANDED_OVER;
                if (ppend[i]) continue;
                if (its[i] < 0) useGD[i] = 1;  // DIRECT_SOLVER.h:42-44 -> gradient descent
                else {
                    S[i].pcg_iters += its[i];
                    S[i].last_relres = rel[i];
                    S[i].pcg_unconverged += c->L.h_maxit[i];
                }
            }
        }
        stage(1);
        // ls: instances that go on to the line search in this call
        std::vector<char> ls(B, 0);
        int nls = 0;
        for (int i = 0; i < B; ++i) {
            ls[i] = (fresh[i] || was_ppend[i]) && !ppend[i];
            nls += ls[i];
        }
        for (int i = 0; i < B; ++i) mask[i] = ls[i] && useGD[i];
        IDP_LAUNCH(k_copy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, up_mask(), rhs, sol);
        const double tsolve = now_ms() - t0;
        for (int i = 0; i < B; ++i) {
            if (fresh[i] || was_ppend[i]) S[i].ms_solve += tsolve;
            if (ls[i] && useGD[i]) S[i].gd_iters++;
        }
        // ---- Line_Search, part 1: halve until no tet inverts
        for (int i = 0; i < B; ++i) mask[i] = ls[i];
        dact = up_mask();
        IDP_LAUNCH(k_copy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, dact, x, xprev);
        for (int i = 0; i < B; ++i)
            if (ls[i]) alpha[i] = 1.0;
        std::vector<unsigned char> srch(ls.begin(), ls.end());
        for (int guard = 0;; ++guard) {
            if (guard > 4000) return IDP_E_INVALID;
            bool any = false;
            for (int i = 0; i < B; ++i) any |= srch[i] != 0;
            if (!any) break;
            mask = srch;
            IDP_LAUNCH(k_axpy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, up_mask(), xprev,
                       dalpha.upload_get(alpha.data(), B, s), sol, x);
            energies(c, x, false, h, eo);
            for (int i = 0; i < B; ++i)
                if (srch[i]) {
                    if (eo[i].valid) srch[i] = 0;
                    else alpha[i] /= 2.0;
                }
        }
        for (int i = 0; i < B; ++i) mask[i] = ls[i];
        dact = up_mask();
        IDP_LAUNCH(k_copy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, dact, xprev, x);
        stage(2);
        // go: instances whose step size is final in this call
        std::vector<char> go(ls);
        if (coll) {
            t0 = now_ms();
            long nc = 0;
            std::vector<double> a2(alpha);
            std::vector<unsigned char> pnow(B, 0);
            // with nothing else in flight a handed-over pair just runs to its end
            const int budget = nls > 0 ? accd_budget : 0;
            contact_ccd_batch(c->cw, s, m, x, sol, thickness, a2.data(), dact, &nc, budget, anyres, pnow.data());
            const double tc = now_ms() - t0;
            for (int i = 0; i < B; ++i)
                if (ls[i] || was_pend[i]) {  // took part in this CCD
                    S[i].accd_capped += c->cw.h_capped[i];
                    go[i] = !pnow[i];
                    if (go[i]) alpha[i] = a2[i];
                    pend[i] = pnow[i];
                    S[i].ms_ccd += tc;
                    if (n_active == 1) S[i].ccd_candidates += nc;
                }
        }
        stage(3);
        for (int i = 0; i < B; ++i)
            if (go[i]) {
                feasible[i] = alpha[i];
                if (PN[i] == 0) S[i].first_ccd_alpha = alpha[i];
            }
        // ---- Line_Search, part 2: backtracking on energy with the
        // minimum-distance safeguard (IMPLICIT_EULER.h:102-146)
        srch.assign(go.begin(), go.end());
        for (int guard = 0;; ++guard) {
            bool any = false;
            for (int i = 0; i < B; ++i) any |= srch[i] != 0;
            if (!any) break;
            if (guard > 4000) return IDP_E_INVALID;  // alpha underflow: x no longer moves
            mask = srch;
            IDP_LAUNCH(k_axpy_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, up_mask(), xprev,
                       dalpha.upload_get(alpha.data(), B, s), sol, x);
            energies(c, x, false, h, eo);
            std::vector<unsigned char> valid(B, 0), retry(B, 0);
            bool anyv = false;
            for (int i = 0; i < B; ++i)
                if (srch[i]) {
                    valid[i] = eo[i].valid;
                    E[i] = eo[i].E;
                    anyv |= valid[i] != 0;
                }
            if (anyv && coll) {
                mask = valid;
                cset(c, up_mask());
                std::vector<double> md;
                min_dist(c, nullptr, md);
                for (int i = 0; i < B; ++i)
                    if (valid[i] && has_constraints(c, i) && md[i] - thickness * thickness <= 0) {
                        alpha[i] /= 2;
                        E[i] = Eprev[i] + 1;
                        S[i].line_search_trials++;
                        retry[i] = 1;
                    }
            }
            bool anyb = false;
            for (int i = 0; i < B; ++i) anyb |= valid[i] && !retry[i];
            if (anyb) {
                energies(c, x, coll, h, eo);
                for (int i = 0; i < B; ++i)
                    if (valid[i] && !retry[i]) {
                        valid[i] = eo[i].valid;
                        E[i] = eo[i].E;
                        if (!eo[i].dist_ok) return IDP_E_NONPOS_DIST;
                    }
            }
            for (int i = 0; i < B; ++i) {
                if (!srch[i] || retry[i]) continue;
                alpha[i] /= 2.0;
                S[i].line_search_trials++;
                if (!(E[i] > Eprev[i] || !valid[i])) srch[i] = 0;  // accepted
            }
        }
        for (int i = 0; i < B; ++i)
            if (go[i]) {
                Eprev[i] = E[i];
                S[i].last_alpha = alpha[i] * 2;
            }
        stage(4);
        // ---- kappa update (IMPLICIT_EULER.h:566-595), instances past their
        // first iteration (the previous constraint set is theirs)
        if (coll && c->ncs_prev) {
            const long np = c->ncs_prev;
            double* cur = c->tmpv.ensure(np, s);
            unsigned long long* dm = c->dmin.ensure(B, s);
            int* kflag = c->flags.ensure(4 * (size_t)B + 4, s);
            launch_min_dist2(s, np, (const int*)c->cs_prev.p, x, I.vinst.p, cur, dm);
            IDP_HIP(hipMemsetAsync(kflag, 0, B * sizeof(int), s));
            IDP_LAUNCH(k_kappa_test, grid_for(np, 256), 256, 0, s, np, c->cs_prev.p, I.vinst.p, cur,
                       c->dist2_prev.p, kflag);
            std::vector<int> hk(B);
            IDP_HIP(hipMemcpyAsync(hk.data(), kflag, B * sizeof(int), hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
            bool anyk = false;
            for (int i = 0; i < B; ++i)
                if (go[i] && PN[i] > 0 && hk[i] && kappa[i] < kmax[i]) {
                    kappa[i] *= 2;
                    anyk = true;
                } else {
                    hk[i] = 0;
                }
            if (anyk) {
                set_bparams(c, dHat2.data(), kappa.data(), th.data());
                energies(c, x, coll, h, eo);
                for (int i = 0; i < B; ++i)
                    if (hk[i]) Eprev[i] = eo[i].E;
            }
        }
        if (coll) {
            const long n = c->cw.ncs;
            double* d2 = c->dist2_prev.ensure(n + 1, s);
            std::vector<double> md;
            min_dist(c, d2, md);
            for (int i = 0; i < B; ++i)
                if (go[i] && has_constraints(c, i)) S[i].min_dist2 = md[i] - thickness * thickness;
            c->cs_prev.ensure(n + 1, s);
            if (n) IDP_HIP(hipMemcpyAsync(c->cs_prev.p, c->cw.cs.p, n * sizeof(int4), hipMemcpyDeviceToDevice, s));
            c->ncs_prev = n;
        }
        // ---- stopping rule (IMPLICIT_EULER.h:647-676, 759, 855)
        double* pd = c->part_i.ensure(I.dblk.nblk + 1, s);
        double* es = c->esum.ensure(3 * (size_t)B, s);
        IDP_LAUNCH(k_dot_parts_seg, I.dblk.nblk, RB, 0, s, I.dblk.lo.p, I.dblk.hi.p, sol, sol, pd);
        IDP_LAUNCH(k_seg_sums, B, RB, 0, s, I.dblk.pbeg.p, pd, es);
        std::vector<double> ss(B);
        IDP_HIP(hipMemcpyAsync(ss.data(), es, B * sizeof(double), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        std::fill(mask.begin(), mask.end(), 0);
        bool anyf = false;
        for (int i = 0; i < B; ++i) {
            if (!go[i]) continue;
            const long nv = I.vbeg[i + 1] - I.vbeg[i];
            double L2 = std::sqrt(ss[i] / (nv - ndbc[i])) / h;
            ++PN[i];
            res[i].push_back(L2);
            if (res[i].size() > 3) res[i].pop_front();
            L2 = *std::max_element(res[i].begin(), res[i].end());
            if (useGD[i]) L2 = tol[i] * 10;
            useGD[i] = (alpha[i] * 2 < 1e-8 && feasible[i] > 1e-8);
            S[i].newton_iters = PN[i];
            S[i].last_res = L2;
            S[i].E = Eprev[i];
            S[i].kappa = kappa[i];
            S[i].n_constraints = has_constraints(c, i) ? c->cw.h_cbeg[i + 1] - c->cw.h_cbeg[i] : 0;
            const bool stop = (res[i].size() >= 3 && L2 <= tol[i]) || (max_newton > 0 && PN[i] >= max_newton);
            if (stop) {
                active[i] = 0;
                finished_now[i] = 1;
                mask[i] = 1;
                anyf = true;
                --n_active;
            }
        }
        // v = (x - xn) / h for the instances that finished
        if (anyf) {
            std::vector<double> ih(B, 1.0 / h);
            IDP_LAUNCH(k_velocity_inst, vgrid, 256, 0, s, m.nV, I.vinst.p, up_mask(), x, c->xn.p, h, c->v.p);
            IDP_HIP(hipStreamSynchronize(s));
        }
        stage(5);
        c->prof.checkpoint();
        return IDP_OK;
    }
};

namespace {

// Uploads a mesh of one or more instances (vertex ranges vbeg, tet ranges
// tbeg; tets in global vertex ids, instance-major).
int setup_mesh(idp_ctx* c, int nV, const double* X_rest, int nT, const int* tets, const double* IB,
               const double* vol, const double* lambda, const double* mu, const double* mass,
               const std::vector<long>& vbeg, const std::vector<long>& tbeg, int nE = 0, const int* tris = nullptr,
               const double* tvol = nullptr, int nH = 0, const int* hinges = nullptr)
{
        if (nV <= 0 || nT < 0 || nE < 0 || nT + (long)nE <= 0 || !X_rest || !mass) return IDP_E_INVALID;
        if (nT > 0 && (!tets || !IB || !vol || !lambda || !mu)) return IDP_E_INVALID;
        if (nE > 0 && (!tris || !tvol || vbeg.size() != 2)) return IDP_E_INVALID;  // shells: one instance
        if ((long)nT + nE + nH > INT32_MAX / 16 || nH < 0 || (nH > 0 && (!hinges || nE <= 0))) return IDP_E_INVALID;
        for (long i = 0; i < 4L * nH; ++i)
            if (hinges[i] < 0 || hinges[i] >= nV) return IDP_E_INVALID;
        for (long i = 0; i < 4L * nT; ++i)
            if (tets[i] < 0 || tets[i] >= nV) return IDP_E_INVALID;
        for (long i = 0; i < 3L * nE; ++i)
            if (tris[i] < 0 || tris[i] >= nV) return IDP_E_INVALID;
        IDP_HIP(hipSetDevice(c->device));
        auto s = c->s;
        MeshDev& m = c->m;
        m.nV = nV;
        m.nT = nT;
        m.X0.upload(X_rest, 3 * (size_t)nV, s);
        m.mass.upload(mass, nV, s);
        m.tets.upload(tets, 4 * (size_t)nT, s);
        m.IB.upload(IB, 9 * (size_t)nT, s);
        m.vol.upload(vol, nT, s);
        m.lam.upload(lambda, nT, s);
        m.mu.upload(mu, nT, s);
        m.nE = nE;
        m.tris.upload(tris, 3 * (size_t)nE, s);
        m.tvol.upload(tvol, nE, s);
        m.nH = nH;
        m.hinge.upload(hinges, 4 * (size_t)nH, s);
        m.shell_el = false;  // membrane / bending data come with idp_set_shell_elasticity
        // vertex -> triangles, ascending (the order of the normal-flow sums)
        std::vector<int> vtp(nV + 1, 0), vtl(3 * (size_t)nE);
        for (long i = 0; i < 3L * nE; ++i) ++vtp[tris[i] + 1];
        for (int v = 0; v < nV; ++v) vtp[v + 1] += vtp[v];
        {
            std::vector<int> f(vtp.begin(), vtp.end() - 1);
            for (int e = 0; e < nE; ++e)
                for (int k = 0; k < 3; ++k) vtl[f[tris[3 * e + k]]++] = e;
        }
        m.vt_ptr.upload(vtp.data(), vtp.size(), s);
        m.vt.upload(vtl.data(), vtl.size(), s);
        std::vector<unsigned char> dbc(nV, 0);
        m.dbc.upload(dbc.data(), nV, s);
        c->ndbc = 0;
        // surface
        std::vector<int> bn;
        std::vector<std::array<int, 2>> be;
        std::vector<std::array<int, 3>> bt;
        // on the device (host restatement only beyond the device keys' widths)
        IDP_HIP(hipStreamSynchronize(s));
        if (nE > 0 || !surface_device(s, c->sw, nV, m.X0.p, nT, m.tets.p, bn, be, bt))
            surface_host(nV, X_rest, nT, tets, nE, tris, bn, be, bt);
        m.nb = (int)bn.size();
        m.ne = (int)be.size();
        m.nf = (int)bt.size();
        m.bn.upload(bn.data(), bn.size(), s);
        m.be.upload((const int*)be.data(), 2 * be.size(), s);
        m.bt.upload((const int*)bt.data(), 3 * bt.size(), s);
        std::vector<int> v2b(nV, -1);
        for (int i = 0; i < m.nb; ++i) v2b[bn[i]] = i;
        m.v2b.upload(v2b.data(), nV, s);
        // static elastic pattern: bucket records by row, sort each row by (col, src)
        // (shell triangle e is element nT + e with its 3 x 3 leading blocks)
        std::vector<int> rcount(nV + 1, 0);
        for (int t = 0; t < nT; ++t)
            for (int a = 0; a < 4; ++a) rcount[tets[4 * t + a] + 1] += 4;
        for (int e = 0; e < nE; ++e)
            for (int a = 0; a < 3; ++a) rcount[tris[3 * e + a] + 1] += 3;
        for (int k = 0; k < nH; ++k)  // hinges: element nT + nE + k, four vertices
            for (int a = 0; a < 4; ++a) rcount[hinges[4 * k + a] + 1] += 4;
        for (int v = 0; v < nV; ++v) rcount[v + 1] += rcount[v];
        std::vector<std::pair<int, int>> rec(16 * (size_t)nT + 9 * (size_t)nE + 16 * (size_t)nH);  // (col, src)
        std::vector<int> fill(rcount.begin(), rcount.end() - 1);
        for (int t = 0; t < nT; ++t)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    rec[fill[tets[4 * t + a]]++] = {tets[4 * t + b], t * 16 + a * 4 + b};
        for (int e = 0; e < nE; ++e)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b)
                    rec[fill[tris[3 * e + a]]++] = {tris[3 * e + b], (nT + e) * 16 + a * 4 + b};
        for (int k = 0; k < nH; ++k)
            for (int a = 0; a < 4; ++a)
                for (int b = 0; b < 4; ++b)
                    rec[fill[hinges[4 * k + a]]++] = {hinges[4 * k + b], (nT + nE + k) * 16 + a * 4 + b};
        std::vector<int> rowptr(nV + 1, 0), col, sptr, src, diag(nV, -1), rowof;
        col.reserve(27 * (size_t)nV);
        src.reserve(16 * (size_t)nT);
        sptr.push_back(0);
        for (int v = 0; v < nV; ++v) {
            std::sort(rec.begin() + rcount[v], rec.begin() + rcount[v + 1]);
            for (int k = rcount[v]; k < rcount[v + 1];) {
                int j = k;
                while (j < rcount[v + 1] && rec[j].first == rec[k].first) src.push_back(rec[j++].second);
                if (rec[k].first == v) diag[v] = (int)col.size();
                col.push_back(rec[k].first);
                rowof.push_back(v);
                sptr.push_back((int)src.size());
                k = j;
            }
            rowptr[v + 1] = (int)col.size();
            if (diag[v] < 0) return IDP_E_INVALID;  // vertex in no element
        }
        m.nnzb_e = (long)col.size();
        m.e_rowptr.upload(rowptr.data(), rowptr.size(), s);
        m.e_col.upload(col.data(), col.size(), s);
        m.e_src_ptr.upload(sptr.data(), sptr.size(), s);
        m.e_src.upload(src.data(), src.size(), s);
        m.e_diag.upload(diag.data(), diag.size(), s);
        init_erow(c->L, s, m, rowof);
        // instance ranges of the boundary primitives (vertex ids are instance-major,
        // so each instance's primitives are contiguous in the surface order)
        const int B = (int)vbeg.size() - 1;
        auto inst_of = [&](int v) { return (int)(std::upper_bound(vbeg.begin(), vbeg.end(), (long)v) - vbeg.begin()) - 1; };
        auto ranges = [&](int n, auto first_vertex, std::vector<long>& beg) -> bool {
            beg.assign(B + 1, 0);
            int prev = 0;
            for (int k = 0; k < n; ++k) {
                const int in = inst_of(first_vertex(k));
                if (in < prev) return false;
                for (int q = prev + 1; q <= in; ++q) beg[q] = k;
                prev = in;
            }
            for (int q = prev + 1; q <= B; ++q) beg[q] = n;
            return true;
        };
        std::vector<long> nbeg, ebeg, fbeg;
        if (!ranges(m.nb, [&](int k) { return bn[k]; }, nbeg) ||
            !ranges(m.ne, [&](int k) { return std::min(be[k][0], be[k][1]); }, ebeg) ||
            !ranges(m.nf, [&](int k) { return std::min(bt[k][0], std::min(bt[k][1], bt[k][2])); }, fbeg))
            return IDP_E_INVALID;
        build_inst_layout(m, vbeg, tbeg, nbeg, ebeg, fbeg, s);
        {
            std::vector<long> sbeg(B + 1, 0), hbeg(B + 1, 0);
            sbeg[B] = nE;  // nE > 0 only with one instance
            hbeg[B] = nH;
            m.inst.sblk.build(sbeg, RB, s);
            m.inst.hblk.build(hbeg, RB, s);
        }
        c->cw.h_cbeg.clear();
        // state
        c->x.upload(X_rest, 3 * (size_t)nV, s);
        std::vector<double> zero(3 * (size_t)nV, 0.0);
        c->v.upload(zero.data(), zero.size(), s);
        c->cw.ncs = 0;
        c->ncs_prev = 0;
        IDP_HIP(hipStreamSynchronize(s));
        c->have_mesh = true;
        c->L.coarse.ready = false;
        c->snap_valid = false;
        return IDP_OK;
}

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
// ---------------------------------------------------------------------------
// allocation-free barrier energy: a fixed scratch array per device
// ---------------------------------------------------------------------------
// Layout (byte offsets): block partials, the three barrier parameters, the
// result (E and the d2 <= 0 flag, adjacent so that one copy returns both),
// then the constraint block bounds.  The block bounds are those of the
// one-instance constraint table (blocks of RB constraints), so the energy is
// summed in the same order as by idp_barrier_energy.
namespace {
constexpr long NA_MAX_N = IDP_NOALLOC_MAX_CONSTRAINTS;
constexpr int NA_MAX_BLK = (int)((NA_MAX_N + RB - 1) / RB);
constexpr size_t NA_PART = 0;
constexpr size_t NA_PAR = NA_PART + 8 * (size_t)NA_MAX_BLK;
constexpr size_t NA_OUT = NA_PAR + 8 * 4;
constexpr size_t NA_LO = NA_OUT + 16;
constexpr size_t NA_HI = NA_LO + 4 * (size_t)NA_MAX_BLK;
constexpr size_t NA_BYTES = NA_HI + 4 * (size_t)NA_MAX_BLK;
struct NaOut {
    double E;
    int flag;
    int pad;
};
static_assert(NA_OUT % 8 == 0 && NA_LO % 4 == 0, "scratch layout alignment");
}  // namespace

// Each GPU the library runs on holds one copy until the process ends.  Host
// code never passes this name to a kernel or copy: it uses the address that
// hipGetSymbolAddress returned for the device (idp_ctx::scratch).
static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {
constexpr int NA_MAX_DEVICES = 64;
std::mutex na_mu;
unsigned char* na_addr[NA_MAX_DEVICES] = {};

// this device's scratch address; `device` must be the current device
int scratch_address(int device, unsigned char** out)
{
    if (device < 0 || device >= NA_MAX_DEVICES) return IDP_E_INVALID;
    std::lock_guard<std::mutex> lk(na_mu);
    if (!na_addr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return IDP_E_HIP;
        na_addr[device] = (unsigned char*)p;
    }
    *out = na_addr[device];
    return IDP_OK;
}

// block bounds of n constraints in blocks of RB, the barrier parameters, and
// a cleared result
__global__ void k_noalloc_setup(long n, int nblk, int* __restrict__ lo, int* __restrict__ hi,
                                double* __restrict__ par, double dh, double kappa, double th2,
                                NaOut* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nblk) {
        lo[i] = (int)(i * RB);
        hi[i] = (int)std::min<long>((i + 1) * RB, n);
    }
    if (i == 0) {
        par[0] = dh;
        par[1] = kappa;
        par[2] = th2;
        out->E = 0.0;
        out->flag = 0;
        out->pad = 0;
    }
}
}  // namespace

extern "C" {

int idp_device_count(int* n)
{
    int c = 0;
    if (hipGetDeviceCount(&c) != hipSuccess) c = 0;
    if (n) *n = c;
    return IDP_OK;
}

int idp_ctx_create(int device, idp_ctx** out)
{
    if (!out) return IDP_E_INVALID;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return IDP_E_NO_GPU;
    if (device < 0 || device >= n) return IDP_E_INVALID;
    idp_ctx* c = new idp_ctx();
    c->eng = new BatchEngine();
    c->device = device;
    try {
        IDP_HIP(hipSetDevice(device));
        IDP_HIP(hipStreamCreateWithFlags(&c->s, hipStreamNonBlocking));
        barrier_init_constants();
        if (scratch_address(device, &c->scratch) != IDP_OK)
            throw HipError(IDP_E_HIP, "hipGetSymbolAddress of the scratch array failed");
        c->prof.s = c->s;
        c->cw.prof = &c->prof;
        c->L.prof = &c->prof;
        // SpMV launch shape override (tuning runs): an index into the shape table
        if (const char* e = std::getenv("IDP_SPMV_SHAPE")) {
            const int k = std::atoi(e);
            if (k >= 0 && k < 8) c->L.spmv_shape = k;
        }
    } catch (const HipError& e) {
        delete c->eng;
        delete c;
        return e.code;
    }
    *out = c;
    return IDP_OK;
}

int idp_ctx_destroy(idp_ctx* c)
{
    if (!c) return IDP_E_INVALID;
    (void)hipSetDevice(c->device);
    if (c->s) {
        (void)hipStreamSynchronize(c->s);
        (void)hipStreamDestroy(c->s);
    }
    if (c->pinned) (void)hipHostFree(c->pinned);
    delete c->eng;
    delete c;
    return IDP_OK;
}

const char* idp_last_error(const idp_ctx* c) { return c ? c->err.c_str() : "null context"; }

int idp_device_memory(long long* in_use, long long* peak)
{
    if (in_use) *in_use = dev_mem().in_use.load();
    if (peak) *peak = dev_mem().peak.load();
    return IDP_OK;
}

int idp_set_launch_config(idp_ctx* c, int spmv_shape, int upd_waves)
{
    return guarded(c, [&]() -> int {
        if (spmv_shape < 0 || spmv_shape > 7 || upd_waves < 1 || upd_waves > 8) return IDP_E_INVALID;
        c->L.spmv_shape = spmv_shape;
        c->L.upd_waves = upd_waves;
        return IDP_OK;
    });
}

int idp_set_barrier_team(idp_ctx* c, int lanes)
{
    if (!c || !(lanes == 3 || lanes == 8 || lanes == 16 || lanes == 32 || lanes == 64)) return IDP_E_INVALID;
    c->hess_team = lanes;
    return IDP_OK;
}

int idp_get_barrier_team(idp_ctx* c, int* lanes)
{
    if (!c || !lanes) return IDP_E_INVALID;
    *lanes = c->hess_team;
    return IDP_OK;
}

int idp_get_launch_config(idp_ctx* c, int* spmv_shape, int* upd_waves)
{
    return guarded(c, [&]() -> int {
        if (!spmv_shape || !upd_waves) return IDP_E_INVALID;
        *spmv_shape = c->L.spmv_shape;
        *upd_waves = c->L.upd_waves;
        return IDP_OK;
    });
}

int idp_profile_enable(idp_ctx* c, int on)
{
    return guarded(c, [&]() -> int {
        if (on < 0 || on > 2) return IDP_E_INVALID;
        c->prof.reset();
        c->prof.level = on;
        c->prof.on = on != 0;
        int* b = c->bcounts.ensure(16, c->s);
        IDP_HIP(hipMemsetAsync(b, 0, 16 * sizeof(int), c->s));
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

// algorithmic flops of one barrier Hessian evaluation: derivative terms, the
// n^2 outer products, a 9n^3 symmetric eigendecomposition and, when clamped,
// a 2n^3 reconstruction (n = 12 for PT/EE and mollified kinds, 9 PE, 6 PP)
static double barrier_hess_flops(int kind, bool clamped)
{
    static const double fdh[7] = {1500, 1200, 450, 60, 1200, 1200, 1200};
    static const int nn[7] = {12, 12, 9, 6, 12, 12, 12};
    const double n = nn[kind];
    return fdh[kind] + 3 * n * n + 9 * n * n * n + (clamped ? 2 * n * n * n : 0.0);
}

int idp_profile_read(idp_ctx* c, idp_profile* out)
{
    return guarded(c, [&]() -> int {
        if (!out) return IDP_E_INVALID;
        c->prof.collect();
        int cnt[16] = {0};
        if (c->bcounts.p) {
            IDP_HIP(hipMemcpyAsync(cnt, c->bcounts.p, sizeof(cnt), hipMemcpyDeviceToHost, c->s));
            IDP_HIP(hipStreamSynchronize(c->s));
        }
        std::memset(out, 0, sizeof(*out));
        for (int i = 0; i < IDP_K_COUNT; ++i) {
            out->ms[i] = c->prof.ms[i];
            out->launches[i] = c->prof.launches[i];
            out->work[i] = c->prof.work[i];
        }
        double fl = 0;
        for (int k = 0; k < 7; ++k) {
            out->barrier_pairs[k] = cnt[2 * k];
            out->barrier_clamped[k] = cnt[2 * k + 1];
            fl += (cnt[2 * k] - cnt[2 * k + 1]) * barrier_hess_flops(k, false) + cnt[2 * k + 1] * barrier_hess_flops(k, true);
        }
        out->work[IDP_K_BARRIER_HESS] = fl;
        out->ccd_steps = c->prof.ccd_steps;
        out->ccd_max_steps = c->prof.ccd_max_steps;
        return IDP_OK;
    });
}

int idp_fem_init(int nV, const double* X, int nT, const int* tets, double density, double E, double nu,
                 double* vol, double* IB, double* lambda, double* mu, double* mass)
{
    if (nV <= 0 || nT <= 0 || !X || !tets) return IDP_E_INVALID;
    const double L = E * nu / ((1 + nu) * (1 - 2 * nu));
    const double M = E / (2 * (1 + nu));
    for (int v = 0; v < nV; ++v) mass[v] = 0.0;
    for (int t = 0; t < nT; ++t) {
        const int* e = tets + 4 * t;
        double D[9];
        for (int k = 0; k < 3; ++k)
            for (int i = 0; i < 3; ++i) D[3 * i + k] = X[3 * e[k + 1] + i] - X[3 * e[0] + i];
        double det = det3h(D);
        if (det == 0.0) return IDP_E_INVALID;  // MATRIX::invert exits on a singular basis
        vol[t] = det / 6.0;
        double c[9];
        c[0] = D[4] * D[8] - D[5] * D[7];
        c[1] = D[5] * D[6] - D[3] * D[8];
        c[2] = D[3] * D[7] - D[4] * D[6];
        c[3] = D[2] * D[7] - D[1] * D[8];
        c[4] = D[0] * D[8] - D[2] * D[6];
        c[5] = D[1] * D[6] - D[0] * D[7];
        c[6] = D[1] * D[5] - D[2] * D[4];
        c[7] = D[2] * D[3] - D[0] * D[5];
        c[8] = D[0] * D[4] - D[1] * D[3];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) IB[9 * t + 3 * i + j] = c[3 * j + i] / det;
        lambda[t] = L;
        mu[t] = M;
        double me = vol[t] * density / 4;
        for (int k = 0; k < 4; ++k) mass[e[k]] += me;
    }
    return IDP_OK;
}

int idp_surface(int nV, const double* X, int nT, const int* tets, int* nb, int* ne, int* nf, int* bnode, int* bedge,
                int* btri)
{
    if (nV <= 0 || nT <= 0) return IDP_E_INVALID;
    std::vector<int> bn;
    std::vector<std::array<int, 2>> be;
    std::vector<std::array<int, 3>> bt;
    surface_host(nV, X, nT, tets, 0, nullptr, bn, be, bt);
    *nb = (int)bn.size();
    *ne = (int)be.size();
    *nf = (int)bt.size();
    if (bnode) std::memcpy(bnode, bn.data(), bn.size() * sizeof(int));
    if (bedge)
        for (size_t i = 0; i < be.size(); ++i) {
            bedge[2 * i] = be[i][0];
            bedge[2 * i + 1] = be[i][1];
        }
    if (btri)
        for (size_t i = 0; i < bt.size(); ++i)
            for (int k = 0; k < 3; ++k) btri[3 * i + k] = bt[i][k];
    return IDP_OK;
}

int idp_shell_init(int nV, const double* X, int nE, const int* tris, double density, double E, double nu,
                   double thickness, int* nE_kept, int* kept, double* mass, double* vol, double* lambda, double* mu)
{
    if (nV <= 0 || nE < 0 || !X || (nE && !tris) || !nE_kept || !kept || !mass || !vol || !lambda || !mu)
        return IDP_E_INVALID;
    for (long i = 0; i < 3L * nE; ++i)
        if (tris[i] < 0 || tris[i] >= nV) return IDP_E_INVALID;
    auto P = [&](int v, int d) { return X[3 * v + d]; };
    int n = 0;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        double u[3], w[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = P(t[1], d) - P(t[0], d);
            w[d] = P(t[2], d) - P(t[0], d);
        }
        const double a = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
        const double b = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
        const double dd = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        if (a * dd - b * b == 0.0) continue;  // degenerate first fundamental form
        kept[3 * n] = t[0];
        kept[3 * n + 1] = t[1];
        kept[3 * n + 2] = t[2];
        ++n;
    }
    const double L = E * nu / (1.0 - nu * nu);
    const double M = E / (2.0 * (1.0 + nu));
    for (int e = 0; e < n; ++e) {
        const int* t = kept + 3 * e;
        double u[3], w[3];
        for (int d = 0; d < 3; ++d) {
            u[d] = P(t[1], d) - P(t[0], d);
            w[d] = P(t[2], d) - P(t[0], d);
        }
        const double c0 = u[1] * w[2] - u[2] * w[1], c1 = u[2] * w[0] - u[0] * w[2], c2 = u[0] * w[1] - u[1] * w[0];
        const double len = std::sqrt(c0 * c0 + c1 * c1 + c2 * c2);
        const double portion = len / 2 * thickness * density / 3;
        for (int k = 0; k < 3; ++k) mass[t[k]] += portion;
        vol[e] = len / 2 * thickness;
        lambda[e] = L;
        mu[e] = M;
    }
    *nE_kept = n;
    return IDP_OK;
}

int idp_shell_surface(int nV, const double* X, int nT, const int* tets, int nE, const int* tris, int* nb, int* ne,
                      int* nf, int* bnode, int* bedge, int* btri)
{
    if (nV <= 0 || nT < 0 || nE < 0 || !X || (nT && !tets) || (nE && !tris) || !nb || !ne || !nf) return IDP_E_INVALID;
    for (long i = 0; i < 4L * nT; ++i)
        if (tets[i] < 0 || tets[i] >= nV) return IDP_E_INVALID;
    for (long i = 0; i < 3L * nE; ++i)
        if (tris[i] < 0 || tris[i] >= nV) return IDP_E_INVALID;
    std::vector<int> bn;
    std::vector<std::array<int, 2>> be;
    std::vector<std::array<int, 3>> bt;
    surface_host(nV, X, nT, tets, nE, tris, bn, be, bt);
    *nb = (int)bn.size();
    *ne = (int)be.size();
    *nf = (int)bt.size();
    if (bnode) std::memcpy(bnode, bn.data(), bn.size() * sizeof(int));
    if (bedge) std::memcpy(bedge, be.data(), be.size() * sizeof(be[0]));
    if (btri) std::memcpy(btri, bt.data(), bt.size() * sizeof(bt[0]));
    return IDP_OK;
}

namespace {
double host_dihedral(const double* X, int v0, int v1, int v2, int v3)
{
    auto sub = [&](int a, int b, double* o) {
        for (int d = 0; d < 3; ++d) o[d] = X[3 * a + d] - X[3 * b + d];
    };
    auto crs = [](const double* a, const double* b, double* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    };
    auto dt = [](const double* a, const double* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; };
    double a[3], b[3], n1[3], n2[3], c[3], e[3];
    sub(v1, v0, a);
    sub(v2, v0, b);
    crs(a, b, n1);
    sub(v2, v3, a);
    sub(v1, v3, b);
    crs(a, b, n2);
    const double cc = dt(n1, n2) / std::sqrt(dt(n1, n1) * dt(n2, n2));
    double da = std::acos(std::max(-1.0, std::min(1.0, cc)));
    crs(n2, n1, c);
    sub(v1, v2, e);
    if (dt(c, e) < 0) da = -da;
    return da;
}
}  // namespace

int idp_shell_hinges(int nV, const double* X, int nE, const int* tris, int* nH, int* hinges, double* hinge_rest,
                     double* tri_IB)
{
    if (nV <= 0 || nE <= 0 || !X || !tris || !nH || !hinges || !hinge_rest) return IDP_E_INVALID;
    for (long i = 0; i < 3L * nE; ++i)
        if (tris[i] < 0 || tris[i] >= nV) return IDP_E_INVALID;
    // edges of the triangles, first-seen orientation, sorted (MESHIO.h:728-766)
    std::set<std::array<int, 2>> es;
    std::map<std::pair<int, int>, int> e2t;
    for (int e = 0; e < nE; ++e) {
        const int* t = tris + 3 * e;
        const int pr[3][2] = {{t[0], t[1]}, {t[1], t[2]}, {t[2], t[0]}};
        for (auto& q : pr)
            if (!es.count({q[1], q[0]})) es.insert({q[0], q[1]});
        e2t[{t[0], t[1]}] = e;
        e2t[{t[1], t[2]}] = e;
        e2t[{t[2], t[0]}] = e;
    }
    // hinge stencils and rest data (DISCRETE_SHELL.h:169-215)
    int n = 0;
    for (auto& ed : es) {
        const int a = ed[0], b = ed[1];
        auto f = e2t.find({a, b});
        auto g = e2t.find({b, a});
        if (f == e2t.end() || g == e2t.end()) continue;
        const int* t = tris + 3 * f->second;
        const int* o = tris + 3 * g->second;
        int v = -1, w = -1;
        for (int j = 0; j < 3; ++j) {
            if (t[j] == b) v = t[(j + 1) % 3];
            if (o[j] == a && w < 0) w = o[(j + 1) % 3];
        }
        int* s = hinges + 4 * n;
        s[0] = v;
        s[1] = a;
        s[2] = b;
        s[3] = w;
        double* r = hinge_rest + 3 * n;
        r[0] = host_dihedral(X, v, a, b, w);
        double d[3], p[3], q[3], n1[3], n2[3];
        for (int k = 0; k < 3; ++k) d[k] = X[3 * a + k] - X[3 * b + k];
        r[1] = std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
        for (int k = 0; k < 3; ++k) {
            p[k] = X[3 * a + k] - X[3 * v + k];
            q[k] = X[3 * b + k] - X[3 * v + k];
        }
        n1[0] = p[1] * q[2] - p[2] * q[1];
        n1[1] = p[2] * q[0] - p[0] * q[2];
        n1[2] = p[0] * q[1] - p[1] * q[0];
        for (int k = 0; k < 3; ++k) {
            p[k] = X[3 * b + k] - X[3 * w + k];
            q[k] = X[3 * a + k] - X[3 * w + k];
        }
        n2[0] = p[1] * q[2] - p[2] * q[1];
        n2[1] = p[2] * q[0] - p[0] * q[2];
        n2[2] = p[0] * q[1] - p[1] * q[0];
        r[2] = (std::sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]) +
                std::sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2])) / (r[1] * 6);
        ++n;
    }
    *nH = n;
    if (tri_IB)  // rest first fundamental form (DISCRETE_SHELL.h:146-158)
        for (int e = 0; e < nE; ++e) {
            const int* t = tris + 3 * e;
            double u[3], w[3];
            for (int k = 0; k < 3; ++k) {
                u[k] = X[3 * t[1] + k] - X[3 * t[0] + k];
                w[k] = X[3 * t[2] + k] - X[3 * t[0] + k];
            }
            tri_IB[3 * e] = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
            tri_IB[3 * e + 1] = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
            tri_IB[3 * e + 2] = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        }
    return IDP_OK;
}

int idp_initialize_oipc(double dhat, int nV, const double* mass, double stiff_mult, double* kappa3, double* dHat2)
{
    if (nV <= 0 || !mass || !kappa3) return IDP_E_INVALID;
    const double d2 = dhat * dhat;
    const double dd = 1.0e-16, t = dd - d2;
    const double Hb = (-2.0 * std::log(dd / d2) - 4.0 * t / dd + t * t / (dd * dd));
    double msum = 0;
    for (int v = 0; v < nV; ++v) msum += mass[v];
    const double k0 = stiff_mult * 1.0e11 * (msum / nV) * 3 / (4.0e-16 * Hb);
    kappa3[0] = k0;
    kappa3[1] = 100 * k0;
    kappa3[2] = 0.0;
    if (dHat2) *dHat2 = d2;
    return IDP_OK;
}

int idp_set_mesh(idp_ctx* c, int nV, const double* X_rest, int nT, const int* tets, const double* IB,
                 const double* vol, const double* lambda, const double* mu, const double* mass)
{
    return guarded(c, [&]() -> int {
        return setup_mesh(c, nV, X_rest, nT, tets, IB, vol, lambda, mu, mass, {0, (long)nV}, {0, (long)nT});
    });
}

int idp_set_mesh_shell(idp_ctx* c, int nV, const double* X_rest, const double* mass, int nT, const int* tets,
                       const double* IB, const double* vol, const double* lambda, const double* mu, int nE,
                       const int* tris, const double* tri_vol, int nH, const int* hinges)
{
    return guarded(c, [&]() -> int {
        if (nE <= 0) return IDP_E_INVALID;
        return setup_mesh(c, nV, X_rest, nT, tets, IB, vol, lambda, mu, mass, {0, (long)nV}, {0, (long)nT}, nE,
                          tris, tri_vol, nH, hinges);
    });
}

int idp_set_shell_elasticity(idp_ctx* c, const double* tri_IB, const double* tri_lambda, const double* tri_mu,
                             const double* hinge_rest, double bend_k)
{
    return guarded(c, [&]() -> int {
        MeshDev& m = c->m;
        if (!c->have_mesh || m.nE <= 0 || !tri_IB || !tri_lambda || !tri_mu || (m.nH > 0 && !hinge_rest) ||
            !std::isfinite(bend_k))
            return IDP_E_INVALID;
        m.tIB.upload(tri_IB, 3 * (size_t)m.nE, c->s);
        m.tlam.upload(tri_lambda, m.nE, c->s);
        m.tmu.upload(tri_mu, m.nE, c->s);
        m.hrest.upload(hinge_rest, 3 * (size_t)m.nH, c->s);
        m.bend_k = bend_k;
        m.shell_el = true;
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_normal_flow_force(idp_ctx* c, const double* X, double magnitude, double* b)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || c->m.nE <= 0 || !b) return IDP_E_INVALID;
        IDP_HIP(hipSetDevice(c->device));
        const size_t n3 = 3 * (size_t)c->m.nV;
        double* o = c->bforce.ensure(n3, c->s);
        launch_normal_flow(c->s, c->m, stage_x(c, X), magnitude, o);
        c->bforce.download(b, n3, c->s);
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_set_friction(idp_ctx* c, double mu, double epsv2, int fric_iters, int ncomp, const int* comp_range,
                     const double* mu_comp)
{
    return guarded(c, [&]() -> int {
        if (!std::isfinite(mu) || mu < 0 || !std::isfinite(epsv2) || ncomp < 0 ||
            (ncomp > 0 && (!comp_range || !mu_comp)))
            return IDP_E_INVALID;
        if ((mu > 0 || ncomp > 0) && !(epsv2 > 0)) return IDP_E_INVALID;
        if ((mu > 0 || ncomp > 0) && c->have_mesh && c->m.inst.B != 1) return IDP_E_INVALID;
        for (int i = 1; i < ncomp; ++i)
            if (comp_range[i] < comp_range[i - 1]) return IDP_E_INVALID;
        FricState& f = c->fr;
        f.mu = mu;
        f.epsv2 = epsv2;
        f.iters = fric_iters;
        f.ncomp = ncomp;
        if (ncomp > 0) {
            f.comp_range.upload(comp_range, ncomp, c->s);
            f.mu_comp.upload(mu_comp, (size_t)ncomp * ncomp, c->s);
            IDP_HIP(hipStreamSynchronize(c->s));
        }
        f.have = false;
        return IDP_OK;
    });
}

int idp_set_normal_flow(idp_ctx* c, double magnitude)
{
    return guarded(c, [&]() -> int {
        if (!std::isfinite(magnitude)) return IDP_E_INVALID;
        c->normal_flow = magnitude;
        return IDP_OK;
    });
}

int idp_set_dbc(idp_ctx* c, int n, const int* idx)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || n < 0) return IDP_E_INVALID;
        std::vector<unsigned char> d(c->m.nV, 0);
        for (int i = 0; i < n; ++i) {
            if (idx[i] < 0 || idx[i] >= c->m.nV) return IDP_E_INVALID;
            d[idx[i]] = 1;
        }
        c->ndbc = 0;
        for (auto b : d) c->ndbc += b;
        c->m.dbc.upload(d.data(), d.size(), c->s);
        c->L.coarse.ready = false;  // its aggregates leave the Dirichlet nodes out
        c->dbc_list.assign(idx, idx + n);
        c->dbc_tgt.clear();  // the nodes stay where each step finds them
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_set_dbc_targets(idp_ctx* c, int n, const int* idx, const double* targets)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || n < 0 || (n && (!idx || !targets))) return IDP_E_INVALID;
        std::vector<unsigned char> d(c->m.nV, 0);
        for (int i = 0; i < n; ++i) {
            if (idx[i] < 0 || idx[i] >= c->m.nV || d[idx[i]]) return IDP_E_INVALID;  // no repeated node
            d[idx[i]] = 1;
        }
        c->ndbc = n;
        c->m.dbc.upload(d.data(), d.size(), c->s);
        c->L.coarse.ready = false;
        c->dbc_list.assign(idx, idx + n);
        c->dbc_tgt.assign(targets, targets + 3 * (size_t)n);
        if (n) {
            c->dbc_idx_d.upload(idx, n, c->s);
            c->dbc_tgt_d.upload(targets, 3 * (size_t)n, c->s);
        }
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_surface_sizes(idp_ctx* c, int* nb, int* ne, int* nf)
{
    if (!c || !c->have_mesh) return IDP_E_INVALID;
    *nb = c->m.nb;
    *ne = c->m.ne;
    *nf = c->m.nf;
    return IDP_OK;
}

int idp_get_surface(idp_ctx* c, int* bnode, int* bedge, int* btri)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const MeshDev& m = c->m;
        if (bnode && m.nb) IDP_HIP(hipMemcpyAsync(bnode, m.bn.p, m.nb * sizeof(int), hipMemcpyDeviceToHost, s));
        if (bedge && m.ne) IDP_HIP(hipMemcpyAsync(bedge, m.be.p, 2 * (size_t)m.ne * sizeof(int), hipMemcpyDeviceToHost, s));
        if (btri && m.nf) IDP_HIP(hipMemcpyAsync(btri, m.bt.p, 3 * (size_t)m.nf * sizeof(int), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        return IDP_OK;
    });
}

int idp_set_state(idp_ctx* c, const double* X, const double* v)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        size_t n = 3 * (size_t)c->m.nV;
        if (X) c->x.upload(X, n, c->s);
        if (v) c->v.upload(v, n, c->s);
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_load_velocity(idp_ctx* c, const double* X_target, double h)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !X_target || !(h > 0)) return IDP_E_INVALID;
        auto s = c->s;
        const size_t n3 = 3 * (size_t)c->m.nV;
        IDP_HIP(hipSetDevice(c->device));
        if (c->pinned_n < n3) {
            IDP_HIP(hipStreamSynchronize(s));
            if (c->pinned) IDP_HIP(hipHostFree(c->pinned));
            c->pinned = nullptr;
            c->pinned_n = 0;
            IDP_HIP(hipHostMalloc((void**)&c->pinned, n3 * sizeof(double), hipHostMallocDefault));
            c->pinned_n = n3;
        } else {
            IDP_HIP(hipStreamSynchronize(s));  // the previous target's copy has left the staging buffer
        }
        std::memcpy(c->pinned, X_target, n3 * sizeof(double));
        double* t = c->target.ensure(n3, s);
        IDP_HIP(hipMemcpyAsync(t, c->pinned, n3 * sizeof(double), hipMemcpyHostToDevice, s));
        IDP_LAUNCH(k_load_velocity, grid_for((long)n3, 256), 256, 0, s, (long)n3, c->x.p, t, h, c->v.p);
        return IDP_OK;
    });
}

int idp_get_state(idp_ctx* c, double* X, double* v)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        size_t n = 3 * (size_t)c->m.nV;
        if (X) c->x.download(X, n, c->s);
        if (v) c->v.download(v, n, c->s);
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_snapshot_state(idp_ctx* c)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        size_t n = 3 * (size_t)c->m.nV;
        c->snap_x.ensure(n, c->s);
        c->snap_v.ensure(n, c->s);
        IDP_HIP(hipMemcpyAsync(c->snap_x.p, c->x.p, n * 8, hipMemcpyDeviceToDevice, c->s));
        IDP_HIP(hipMemcpyAsync(c->snap_v.p, c->v.p, n * 8, hipMemcpyDeviceToDevice, c->s));
        IDP_HIP(hipStreamSynchronize(c->s));
        c->snap_valid = true;
        return IDP_OK;
    });
}

int idp_restore_state(idp_ctx* c)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !c->snap_valid) return IDP_E_INVALID;
        size_t n = 3 * (size_t)c->m.nV;
        IDP_HIP(hipMemcpyAsync(c->x.p, c->snap_x.p, n * 8, hipMemcpyDeviceToDevice, c->s));
        IDP_HIP(hipMemcpyAsync(c->v.p, c->snap_v.p, n * 8, hipMemcpyDeviceToDevice, c->s));
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

// ---------------------------------------------------------------------------
// the time step
// ---------------------------------------------------------------------------
int idp_advance_one_step(idp_ctx* c, double h, double newton_tol, int with_collision, double dHat2, double* kappa3,
                         double thickness, const double* b, int flow, const idp_solver_opts* opts,
                         idp_step_stats* st)
{
    struct ElasticOn {  // the flags never outlive the call, whatever path leaves it
        idp_ctx* c;
        ~ElasticOn()
        {
            c->no_elastic = false;
            c->fr.on = false;
        }
    } restore_elastic{c};
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !kappa3 || !(h > 0)) return IDP_E_INVALID;
        if (c->m.inst.B != 1) return IDP_E_INVALID;  // batched contexts step with idp_advance_batch
        c->no_elastic = flow != 0;
        IDP_TIMER_FLAG("implicitEuler");
        IDP_HIP(hipSetDevice(c->device));
        const double t_start = now_ms();
        idp_step_stats S;
        std::memset(&S, 0, sizeof(S));
        const double rtol = (opts && opts->pcg_rtol > 0) ? opts->pcg_rtol : 1e-8;
        const int maxit = (opts && opts->pcg_maxit > 0) ? opts->pcg_maxit : 10000;
        const int max_newton = opts ? opts->max_newton : 0;
        const int verbose = opts ? opts->verbose : 0;
        c->L.precond = opts ? opts->precond : 0;
        const bool cold = opts && opts->pcg_cold_start;
        c->cw.accd_max_steps = accd_cap_from(opts);
        c->L.coarse.verts_per_agg = coarse_from(opts);
        c->L.coarse.ready = false;  // built anew in each step (its Dirichlet set may differ)
        auto s = c->s;
        MeshDev& m = c->m;
        const long n3 = 3L * m.nV;
        const bool coll = with_collision != 0;
        double* x = c->x.p;
        struct XposOff {  // the coarse level reads the positions of this step's systems only
            LinWork& L;
            ~XposOff() { L.xpos = nullptr; }
        } xpos_off{c->L};
        c->L.xpos = x;
        double* xn = c->xn.ensure(n3, s);
        double* xt = c->xt.ensure(n3, s);
        double* xprev = c->xprev.ensure(n3, s);
        double* sol = c->sol.ensure(n3, s);
        double* rhs = c->rhs.ensure(n3, s);
        double* g = c->g.ensure(n3, s);
        // Xn = X; Xtilde = X + h v + h^2 M^-1 b, v = 0 first for the flow variant (IMPLICIT_EULER.h:196-217)
        IDP_HIP(hipMemcpyAsync(xn, x, n3 * 8, hipMemcpyDeviceToDevice, s));
        const double* bdev = nullptr;
        if (b) {
            bdev = c->bforce.upload_get(b, n3, s);
        } else if (c->normal_flow != 0.0 && m.nE > 0) {
            bdev = c->bforce.ensure(n3, s);
            launch_normal_flow(s, m, x, c->normal_flow, c->bforce.p);
        }
        if (bdev || flow)
            IDP_LAUNCH(k_xtilde, grid_for(n3, 256), 256, 0, s, n3, x, c->v.p, bdev, m.mass.p, h, flow ? 1 : 0, xt);
        else
            dev_axpy(s, n3, x, h, c->v.p, xt);
        double kappa = kappa3[0];
        set_bparams1(c, dHat2, kappa, thickness);
        // lagged friction (IMPLICIT_EULER.h:413-437): mu becomes 1 with
        // per-component coefficients; epsv h = sqrt(epsv2 h h); x_n = the start
        FricState& fr = c->fr;
        const bool fric = coll && fr.enabled();
        if (fric) {
            fr.mu_eff = fr.ncomp > 0 ? 1.0 : fr.mu;
            fr.epsvh = std::sqrt(fr.epsv2 * h * h);
            fr.Xn = xn;
        }
        int fricIterI = 0;

        // ---- Dirichlet nodes toward their targets (IMPLICIT_EULER.h:304-401): the
        // CCD-limited fraction of the move; all of it when that is 1, otherwise
        // none, and a penalty (augmented Lagrangian, Shell/DIRICHLET.h) pulls them
        const int nd = (int)(c->dbc_tgt.size() / 3);  // moving-DBC entries (0: the DBC nodes stay)
        double DBCStiff = 0.0, DBCPenaltyXn = 0.0, DBCAlpha = 1.0;
        m.proj = nullptr;
        std::vector<double> hxd(3 * (size_t)nd);
        auto dbc_x = [&](const double* X) {  // x of the DBC entries on the host, entry order
            double* o = c->dbc_x_d.ensure(3 * (size_t)nd, s);
            IDP_LAUNCH(k_dbc_gather, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, X, o);
            IDP_HIP(hipMemcpyAsync(hxd.data(), o, 3 * (size_t)nd * 8, hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
        };
        auto dbc_dist2 = [&](const double* X) {  // DIRICHLET.h:6-23
            dbc_x(X);
            double sum = 0.0;
            for (int i = 0; i < nd; ++i) {
                double pen = 0.0;
                for (int d = 0; d < 3; ++d)
                    pen += (c->dbc_tgt[3 * i + d] - hxd[3 * i + d]) * (c->dbc_tgt[3 * i + d] - hxd[3 * i + d]);
                sum += pen;
            }
            return sum;
        };
        std::vector<double> hmass;
        auto dbc_energy = [&](const double* X, double& E) {  // DIRICHLET.h:25-48
            dbc_x(X);
            double sum = 0.0;
            for (int i = 0; i < nd; ++i) {
                double pen = 0.0;
                for (int d = 0; d < 3; ++d)
                    pen += (c->dbc_tgt[3 * i + d] - hxd[3 * i + d]) * (c->dbc_tgt[3 * i + d] - hxd[3 * i + d]);
                pen *= hmass[i];
                sum += pen;
            }
            E += 0.5 * DBCStiff * sum;
        };
        if (nd) {
            hmass.resize(nd);
            {
                std::vector<double> mall(m.nV);
                m.mass.download(mall.data(), m.nV, s);
                for (int i = 0; i < nd; ++i) hmass[i] = mall[c->dbc_list[i]];
            }
            dbc_x(x);
            std::vector<double> hp(n3, 0.0);
            std::vector<unsigned char> fixed(m.nV, 0);
            for (int i = 0; i < nd; ++i) {
                const int v = c->dbc_list[i];
                for (int d = 0; d < 3; ++d) hp[3 * v + d] = c->dbc_tgt[3 * i + d] - hxd[3 * i + d];
                if (!(hp[3 * v] || hp[3 * v + 1] || hp[3 * v + 2])) fixed[v] = 1;
            }
            if (coll) {
                long nc = 0;
                double* P = c->hp.upload_get(hp.data(), n3, s);
                DBCAlpha = contact_ccd(c->cw, s, m, x, P, thickness, 1.0, &nc);
                if (DBCAlpha == 0.0) {  // IPC.h:2014-2031 exits on a zero time of impact
                    c->err = "zero CCD step size for the Dirichlet move";
                    return IDP_E_CCD_ZERO;
                }
            }
            if (DBCAlpha == 1.0) {
                IDP_LAUNCH(k_dbc_set, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, c->dbc_tgt_d.p, x);
            } else {
                DBCStiff = 1e6;
                DBCPenaltyXn = dbc_dist2(xn);
                m.proj = c->dbc_fixed_d.upload_get(fixed.data(), m.nV, s);
            }
        }
        S.dbc_alpha = DBCAlpha;
        auto t0 = now_ms();
        auto cset = [&](const double* X) {
            IDP_TIMER_FLAG("Compute_Constraint_Set");
            double ta = now_ms();
            const int tk = c->prof.begin(IDP_K_CONSTRAINT_SET);
            contact_constraint_set(c->cw, s, m, X, &dHat2, thickness);
            update_cseg(c);
            c->prof.end(IDP_K_CONSTRAINT_SET, tk);
            S.constraint_sets++;
            S.ms_constraint_set += now_ms() - ta;
        };
        if (coll) cset(x);
        if (fric) {
            friction_basis(c, x, dHat2, kappa, thickness);
            fr.on = true;
        }
        t0 = now_ms();
        EnergyOut e0 = energy(c, x, coll, h);
        S.ms_energy += now_ms() - t0;
        if (!e0.valid) return IDP_E_INVALID_START;
        if (!e0.dist_ok) return IDP_E_NONPOS_DIST;
        double Eprev = e0.E;
        if (DBCStiff) dbc_energy(x, Eprev);
        std::deque<double> res;
        bool useGD = false;
        int PNIter = 0;
        double L2 = 0;
        DBuf<int4>& csp = c->cs_prev;
        c->ncs_prev = 0;
        do {
            // ---- gradient and Hessian at x with the current constraint set
            t0 = now_ms();
            if (coll) build_contact_structure(c);
            else c->L.ncb = 0;
            gradient(c, x, h, coll, g, rhs);
            if (DBCStiff)
                IDP_LAUNCH(k_dbc_grad, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, c->dbc_tgt_d.p, m.mass.p,
                           m.proj, x, DBCStiff, g, rhs);
            IDP_HIP(hipStreamSynchronize(s));
            S.ms_gradient += now_ms() - t0;
            if (!useGD) {
                t0 = now_ms();
                hessian(c, x, h, coll);
                if (DBCStiff)
                    IDP_LAUNCH(k_dbc_hess, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, m.mass.p, m.proj,
                               m.e_diag.p, m.nnzb_e, DBCStiff, c->L.evals.p);
                IDP_HIP(hipStreamSynchronize(s));
                S.ms_hessian += now_ms() - t0;
                t0 = now_ms();
                double rel = 0;
                // from the second Newton iteration on, sol holds this step's previous direction
                const unsigned char on = 1;
                const unsigned char* warm = PNIter > 0 && !cold ? c->one.upload_get(&on, 1, s) : nullptr;
                int it;
                {
                    IDP_TIMER_FLAG("linearSolve");
                    it = pcg_solve_persistent(c->L, s, m, rhs, sol, rtol, maxit, &rel, warm);
                }
                S.ms_solve += now_ms() - t0;
                if (it < 0) {
                    useGD = true;  // DIRECT_SOLVER.h:42-44 -> gradient descent (IMPLICIT_EULER.h:531-551)
                    IDP_HIP(hipMemcpyAsync(sol, rhs, n3 * 8, hipMemcpyDeviceToDevice, s));
                } else {
                    S.pcg_iters += it;
                    S.last_relres = rel;
                    S.pcg_unconverged += c->L.h_maxit[0];
                }
            } else {
                IDP_HIP(hipMemcpyAsync(sol, rhs, n3 * 8, hipMemcpyDeviceToDevice, s));
            }
            if (useGD) S.gd_iters++;

            // ---- Line_Search
            IDP_HIP(hipMemcpyAsync(xprev, x, n3 * 8, hipMemcpyDeviceToDevice, s));
            double alpha = 1.0, E = 0.0;
            bool valid;
            int guard = 0;
            do {
                if (++guard > 4000) return IDP_E_INVALID;
                dev_axpy(s, n3, xprev, alpha, sol, x);
                t0 = now_ms();
                EnergyOut eo = energy(c, x, false, h);
                S.ms_energy += now_ms() - t0;
                valid = eo.valid;
                if (!valid) alpha /= 2.0;
            } while (!valid);
            IDP_HIP(hipMemcpyAsync(x, xprev, n3 * 8, hipMemcpyDeviceToDevice, s));
            if (coll) {
                t0 = now_ms();
                long nc = 0;
                {
                    IDP_TIMER_FLAG("Compute_Intersection_Free_StepSize");
                    alpha = contact_ccd(c->cw, s, m, x, sol, thickness, alpha, &nc);
                }
                S.ms_ccd += now_ms() - t0;
                S.ccd_candidates += nc;
                S.accd_capped += c->cw.last_capped;
                S.accd_certified += c->cw.last_certified;
                if (alpha == 0.0) {  // IPC.h:2014-2031 exits on a zero time of impact
                    c->err = "zero CCD step size";
                    if (st) *st = S;
                    return IDP_E_CCD_ZERO;
                }
            }
            const double feasibleAlpha = alpha;
            if (PNIter == 0) S.first_ccd_alpha = alpha;
            guard = 0;
            do {
                if (++guard > 4000) return IDP_E_INVALID;  // alpha underflow: x no longer moves
                dev_axpy(s, n3, xprev, alpha, sol, x);
                t0 = now_ms();
                EnergyOut eo = energy(c, x, false, h);
                S.ms_energy += now_ms() - t0;
                valid = eo.valid;
                E = eo.E;
                if (valid) {
                    if (coll) {
                        cset(x);
                        if (c->cw.ncs) {
                            double md = min_dist2_dev(c, x, c->cw.ncs, (const int*)c->cw.cs.p, nullptr) -
                                        thickness * thickness;
                            if (md <= 0) {
                                alpha /= 2;
                                E = Eprev + 1;
                                S.line_search_trials++;
                                continue;
                            }
                        }
                    }
                    t0 = now_ms();
                    eo = energy(c, x, coll, h);
                    S.ms_energy += now_ms() - t0;
                    valid = eo.valid;
                    E = eo.E;
                    if (!eo.dist_ok) return IDP_E_NONPOS_DIST;
                    if (valid && DBCStiff) dbc_energy(x, E);
                }
                alpha /= 2.0;
                S.line_search_trials++;
            } while (E > Eprev || !valid);
            Eprev = E;
            S.last_alpha = alpha * 2;

            // ---- kappa update (IMPLICIT_EULER.h:566-595)
            if (coll && c->ncs_prev) {
                long np = c->ncs_prev;
                double* cur = c->tmpv.ensure(np, s);
                min_dist2_dev(c, x, np, (const int*)csp.p, cur);
                std::vector<double> hc(np), hp(np);
                IDP_HIP(hipMemcpyAsync(hc.data(), cur, np * 8, hipMemcpyDeviceToHost, s));
                IDP_HIP(hipMemcpyAsync(hp.data(), c->dist2_prev.p, np * 8, hipMemcpyDeviceToHost, s));
                IDP_HIP(hipStreamSynchronize(s));
                bool upd = false;
                for (long i = 0; i < np; ++i)
                    if (hp[i] < 1e-18 && hc[i] < hp[i]) {
                        upd = true;
                        break;
                    }
                if (upd && kappa < kappa3[1]) {
                    kappa *= 2;
                    kappa3[0] = kappa;
                    set_bparams1(c, dHat2, kappa, thickness);
                    EnergyOut eo = energy(c, x, coll, h);
                    Eprev = eo.E;
                    if (DBCStiff) dbc_energy(x, Eprev);
                }
            }
            if (coll) {
                long n = c->cw.ncs;
                double* d2 = c->dist2_prev.ensure(n + 1, s);
                if (n) S.min_dist2 = min_dist2_dev(c, x, n, (const int*)c->cw.cs.p, d2) - thickness * thickness;
                csp.ensure(n + 1, s);
                if (n) IDP_HIP(hipMemcpyAsync(csp.p, c->cw.cs.p, n * sizeof(int4), hipMemcpyDeviceToDevice, s));
                c->ncs_prev = n;
            }

            // ---- stopping rule (IMPLICIT_EULER.h:647-676, 759, 855)
            double ss = dev_dot(c->L, s, n3, sol, sol);
            L2 = std::sqrt(ss / (m.nV - c->ndbc)) / h;
            ++PNIter;
            res.push_back(L2);
            if (res.size() > 3) res.pop_front();
            L2 = *std::max_element(res.begin(), res.end());
            if (useGD) L2 = newton_tol * 10;
            useGD = (alpha * 2 < 1e-8 && feasibleAlpha > 1e-8);
            // moving DBC: raise the penalty near convergence until the nodes are
            // within 1% of their move, then drop it (IMPLICIT_EULER.h:701-755)
            if (DBCStiff) {
                const double progress = 1 - std::sqrt(dbc_dist2(x) / DBCPenaltyXn);
                if (progress < 0.99) {
                    if (L2 < newton_tol * 10 && DBCStiff < 1e8) {
                        DBCStiff *= 2;
                        Eprev = energy(c, x, coll, h).E;
                        dbc_energy(x, Eprev);
                    }
                    L2 = newton_tol * 10;
                } else {
                    DBCStiff = 0.0;
                    m.proj = nullptr;
                    Eprev = energy(c, x, coll, h).E;
                }
            }
            // ---- lagged friction update at convergence (IMPLICIT_EULER.h:757-848):
            // a new basis, then one more direction whose size decides the exit
            if (res.size() == 3 && L2 <= newton_tol) {
                ++fricIterI;
                if ((fricIterI < fr.iters || fr.iters <= 0) && fric) {
                    friction_basis(c, x, dHat2, kappa, thickness);
                    build_contact_structure(c);
                    gradient(c, x, h, coll, g, rhs);
                    if (DBCStiff)
                        IDP_LAUNCH(k_dbc_grad, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, c->dbc_tgt_d.p,
                                   m.mass.p, m.proj, x, DBCStiff, g, rhs);
                    hessian(c, x, h, coll);
                    if (DBCStiff)
                        IDP_LAUNCH(k_dbc_hess, grid_for(nd, 256), 256, 0, s, nd, c->dbc_idx_d.p, m.mass.p, m.proj,
                                   m.e_diag.p, m.nnzb_e, DBCStiff, c->L.evals.p);
                    double rel = 0;
                    const unsigned char on = 1;
                    int it;
                    {
                        IDP_TIMER_FLAG("linearSolve");
                        it = pcg_solve_persistent(c->L, s, m, rhs, sol, rtol, maxit, &rel,
                                                  cold ? nullptr : c->one.upload_get(&on, 1, s));
                    }
                    if (it < 0) {  // the reference exits when this solve fails (IMPLICIT_EULER.h:818-825)
                        c->err = "linear solve failed after a friction basis update";
                        if (st) *st = S;
                        return IDP_E_NOT_SPD;
                    }
                    S.pcg_iters += it;
                    S.pcg_unconverged += c->L.h_maxit[0];
                    L2 = std::sqrt(dev_dot(c->L, s, n3, sol, sol) / (m.nV - c->ndbc)) / h;
                    if (L2 > newton_tol) {
                        Eprev = energy(c, x, coll, h).E;
                        if (DBCStiff) dbc_energy(x, Eprev);
                    }
                    S.friction_updates++;
                }
            }
            S.dbc_stiff = DBCStiff;
            S.newton_iters = PNIter;
            if (verbose > 0 && PNIter % verbose == 0)
                std::fprintf(stderr, "[idp] newton %d res %.3e tol %.3e alpha %.3e feasible %.3e ncs %ld E %.6e kappa %.3e t %.1f ms\n",
                             PNIter, L2, newton_tol, alpha * 2, feasibleAlpha, c->cw.ncs, Eprev, kappa,
                             now_ms() - t_start);
            S.last_res = L2;
            S.E = Eprev;
            // Advance_One_Step_IE_Flow ends once no contact is left (IMPLICIT_EULER.h:850-854)
            if (flow && (!coll || c->cw.ncs == 0)) break;
            if (max_newton > 0 && PNIter >= max_newton) break;
        } while (res.size() < 3 || L2 > newton_tol);

        m.proj = nullptr;
        // v = (x - xn) / h
        dev_velocity(s, n3, x, xn, h, c->v.p);
        IDP_HIP(hipStreamSynchronize(s));
        S.n_constraints = c->cw.ncs;
        S.n_friction = fric ? fr.n : 0;
        S.kappa = kappa;
        c->prof.collect();
        S.ms_total = now_ms() - t_start;
        if (st) *st = S;
        return PNIter;
    });
}

// ---------------------------------------------------------------------------
// stage entry points
// ---------------------------------------------------------------------------
int idp_constraint_set(idp_ctx* c, const double* X, double dHat2, double thickness, long* n)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        const double* xd = stage_x(c, X);
        long k = contact_constraint_set(c->cw, c->s, c->m, xd, &dHat2, thickness);
        update_cseg(c);
        IDP_HIP(hipStreamSynchronize(c->s));
        if (n) *n = k;
        return IDP_OK;
    });
}

int idp_get_constraints(idp_ctx* c, int* cs, long cap)
{
    return guarded(c, [&]() -> int {
        if (cap < c->cw.ncs) return IDP_E_CAPACITY;
        c->cw.cs.download((int4*)cs, c->cw.ncs, c->s);
        IDP_HIP(hipStreamSynchronize(c->s));
        return IDP_OK;
    });
}

int idp_barrier_energy(idp_ctx* c, const double* X, long n, const int* cs, double dHat2, double kappa, double thickness,
                       double* E)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        const double* xd = stage_x(c, X);
        set_constraints_from_host(c, n, cs);
        set_bparams1(c, dHat2, kappa, thickness);
        int* fl = c->flags.ensure(4, c->s);
        IDP_HIP(hipMemsetAsync(fl, 0, sizeof(int), c->s));
        const int nb = c->cseg.nblk;
        double* pb = c->part_b.ensure(nb + 1, c->s);
        if (n) launch_barrier_energy(c->s, nb, c->cseg.lo.p, c->cseg.hi.p, (const int*)c->cw.cs.p, xd, c->m.X0.p,
                                     bparams(c), pb, fl);
        int hf = 0;
        IDP_HIP(hipMemcpyAsync(&hf, fl, sizeof(int), hipMemcpyDeviceToHost, c->s));
        IDP_HIP(hipStreamSynchronize(c->s));
        if (hf) return IDP_E_NONPOS_DIST;
        *E = n ? dev_sum_partials(c->L, c->s, nb, pb) : 0.0;
        return IDP_OK;
    });
}

int idp_barrier_energy_noalloc(idp_ctx* c, const double* X_dev, long n, const int* cs_dev, double dHat2,
                               double kappa, double thickness, double* E)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || c->m.inst.B != 1 || !E || n < 0 || (n > 0 && (!X_dev || !cs_dev)) || !c->scratch)
            return IDP_E_INVALID;
        if (n > NA_MAX_N) {
            c->err = "idp_barrier_energy_noalloc: more constraints than IDP_NOALLOC_MAX_CONSTRAINTS";
            return IDP_E_CAPACITY;
        }
        auto s = c->s;
        IDP_HIP(hipSetDevice(c->device));
        unsigned char* base = c->scratch;
        double* part = (double*)(base + NA_PART);
        double* par = (double*)(base + NA_PAR);
        NaOut* out = (NaOut*)(base + NA_OUT);
        int* lo = (int*)(base + NA_LO);
        int* hi = (int*)(base + NA_HI);
        const int nblk = (int)((n + RB - 1) / RB);
        // the parameters exactly as set_bparams forms them
        const double dh = dHat2 + 2 * std::sqrt(dHat2) * thickness;
        const double th2 = thickness * thickness;
        IDP_LAUNCH(k_noalloc_setup, grid_for(std::max(nblk, 1), 256), 256, 0, s, n, nblk, lo, hi, par, dh, kappa,
                   th2, out);
        if (n) {
            launch_barrier_energy(s, nblk, lo, hi, cs_dev, X_dev, c->m.X0.p,
                                  BarrierParams{par, par + 1, par + 2, nullptr}, part, &out->flag);
            launch_sum_partials(s, nblk, part, &out->E);
        }
        NaOut h{};
        IDP_HIP(hipMemcpyAsync(&h, out, sizeof(NaOut), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        if (h.flag) return IDP_E_NONPOS_DIST;
        *E = h.E;
        return IDP_OK;
    });
}

int idp_barrier_gradient(idp_ctx* c, const double* X, long n, const int* cs, double dHat2, double kappa,
                         double thickness, double* g)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        set_constraints_from_host(c, n, cs);
        contact_structure(c->L, s, c->m.nV, n, (const int*)c->cw.cs.p);
        set_bparams1(c, dHat2, kappa, thickness);
        int* fl = c->flags.ensure(4, s);
        IDP_HIP(hipMemsetAsync(fl, 0, sizeof(int), s));
        double* b = c->gb.ensure(12 * (size_t)std::max(n, 1L), s);
        if (n) launch_barrier_grad(s, n, (const int*)c->cw.cs.p, xd, c->m.X0.p, bparams(c), b, fl);
        double* gd = c->g.ensure(3 * (size_t)c->m.nV, s);
        contact_gradient_gather(c->L, s, c->m, xd, nullptr, nullptr, n ? b : nullptr, gd, nullptr);
        int hf = 0;
        IDP_HIP(hipMemcpyAsync(&hf, fl, sizeof(int), hipMemcpyDeviceToHost, s));
        c->g.download(g, 3 * (size_t)c->m.nV, s);
        IDP_HIP(hipStreamSynchronize(s));
        return hf ? IDP_E_NONPOS_DIST : IDP_OK;
    });
}

int idp_barrier_hessian_local(idp_ctx* c, const double* X, long n, const int* cs, double dHat2, double kappa,
                              double thickness, int project, double* H, int* vid)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        set_constraints_from_host(c, n, cs);
        set_bparams1(c, dHat2, kappa, thickness);
        int* fl = c->flags.ensure(4, s);
        IDP_HIP(hipMemsetAsync(fl, 0, sizeof(int), s));
        double* b = c->Hb.ensure(144 * (size_t)std::max(n, 1L), s);
        DBuf<int> dv;
        int* dvid = dv.ensure(4 * (size_t)std::max(n, 1L), s);
        if (n) {
            const int tok = c->prof.begin(IDP_K_BARRIER_HESS);  // timed like the solve's launches
            launch_barrier_hess(s, n, (const int*)c->cw.cs.p, xd, c->m.X0.p, bparams(c), project, 0, b, dvid, fl,
                                c->hform.ensure(90 * (size_t)n, s), c->hform_ok.ensure(n, s),
                                c->prof.full() ? c->bcounts.p : nullptr, c->hperm.ensure(n, s),
                                c->hkcnt.ensure(barrier_kind_scratch(n), s), c->hess_team);
            c->prof.end(IDP_K_BARRIER_HESS, tok);
        }
        int hf = 0;
        IDP_HIP(hipMemcpyAsync(&hf, fl, sizeof(int), hipMemcpyDeviceToHost, s));
        if (n) {
            IDP_HIP(hipMemcpyAsync(H, b, 144 * 8 * (size_t)n, hipMemcpyDeviceToHost, s));
            IDP_HIP(hipMemcpyAsync(vid, dvid, 4 * 4 * (size_t)n, hipMemcpyDeviceToHost, s));
        }
        IDP_HIP(hipStreamSynchronize(s));
        return hf ? IDP_E_NONPOS_DIST : IDP_OK;
    });
}

// friction stage entry points: the set of idp_friction_basis with its x_n,
// the context's coefficients, epsv h from the given step
namespace {
int friction_stage(idp_ctx* c, double h, const double* X, const double** xd)
{
    FricState& f = c->fr;
    if (!c->have_mesh || !f.have || !f.enabled() || !(h > 0)) return IDP_E_INVALID;
    *xd = stage_x(c, X);
    f.mu_eff = f.ncomp > 0 ? 1.0 : f.mu;
    f.epsvh = std::sqrt(f.epsv2 * h * h);
    f.Xn = f.xn.p;
    return IDP_OK;
}
}  // namespace

int idp_friction_basis(idp_ctx* c, const double* X, const double* Xn, long n, const int* cs, double dHat2,
                       double kappa, double thickness, long* nf)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !Xn || n < 0 || (n > 0 && !cs) || c->m.inst.B != 1) return IDP_E_INVALID;
        if (!c->fr.enabled()) return IDP_E_INVALID;
        IDP_HIP(hipSetDevice(c->device));
        const double* xd = stage_x(c, X);
        set_constraints_from_host(c, n, cs);
        c->fr.xn.upload(Xn, 3 * (size_t)c->m.nV, c->s);
        friction_basis(c, xd, dHat2, kappa, thickness);
        if (nf) *nf = c->fr.n;
        return IDP_OK;
    });
}

int idp_get_friction(idp_ctx* c, int* fcs, double* cp, double* basis, double* nforce, long cap)
{
    return guarded(c, [&]() -> int {
        FricState& f = c->fr;
        if (!f.have) return IDP_E_INVALID;
        if (cap < f.n) return IDP_E_CAPACITY;
        auto s = c->s;
        std::vector<double> d(10 * (size_t)f.n);
        if (f.n) {
            if (fcs) IDP_HIP(hipMemcpyAsync(fcs, f.fcs.p, f.n * sizeof(int4), hipMemcpyDeviceToHost, s));
            IDP_HIP(hipMemcpyAsync(d.data(), f.fdat.p, d.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        IDP_HIP(hipStreamSynchronize(s));
        for (long i = 0; i < f.n; ++i) {
            if (cp)
                for (int k = 0; k < 2; ++k) cp[2 * i + k] = d[10 * i + k];
            if (basis)
                for (int k = 0; k < 6; ++k) basis[6 * i + k] = d[10 * i + 2 + k];
            if (nforce) nforce[i] = d[10 * i + 8];
        }
        return IDP_OK;
    });
}

int idp_friction_energy(idp_ctx* c, const double* X, double h, double* E)
{
    return guarded(c, [&]() -> int {
        const double* xd = nullptr;
        int rc = friction_stage(c, h, X, &xd);
        if (rc != IDP_OK || !E) return rc != IDP_OK ? rc : IDP_E_INVALID;
        FricState& f = c->fr;
        auto s = c->s;
        double S = 0.0;
        if (f.n) {
            const int nb = grid_for(f.n, RB);
            double* pf = f.part.ensure(nb + 1, s);
            launch_fric_energy(s, f.n, f.fcs.p, f.fdat.p, xd, f.Xn, f.epsvh, pf);
            S = dev_sum_partials(c->L, s, nb, pf);
        }
        *E = f.mu_eff * S;
        return IDP_OK;
    });
}

int idp_friction_gradient(idp_ctx* c, const double* X, double h, double* g)
{
    return guarded(c, [&]() -> int {
        const double* xd = nullptr;
        int rc = friction_stage(c, h, X, &xd);
        if (rc != IDP_OK || !g) return rc != IDP_OK ? rc : IDP_E_INVALID;
        FricState& f = c->fr;
        auto s = c->s;
        contact_structure(c->L, s, c->m.nV, 0, nullptr, f.n, (const int*)f.fcs.p);
        double* b = c->gb.ensure(12 * (size_t)std::max(f.n, 1L), s);
        if (f.n) launch_fric_grad(s, f.n, f.fcs.p, f.fdat.p, xd, f.Xn, f.epsvh, f.mu_eff, b);
        double* gd = c->g.ensure(3 * (size_t)c->m.nV, s);
        contact_gradient_gather(c->L, s, c->m, xd, nullptr, nullptr, f.n ? b : nullptr, gd, nullptr);
        c->g.download(g, 3 * (size_t)c->m.nV, s);
        IDP_HIP(hipStreamSynchronize(s));
        return IDP_OK;
    });
}

int idp_friction_hessian_local(idp_ctx* c, const double* X, double h, int project, double* H)
{
    return guarded(c, [&]() -> int {
        const double* xd = nullptr;
        int rc = friction_stage(c, h, X, &xd);
        if (rc != IDP_OK || !H) return rc != IDP_OK ? rc : IDP_E_INVALID;
        FricState& f = c->fr;
        auto s = c->s;
        if (f.n) {
            double* b = c->Hb.ensure(144 * (size_t)f.n, s);
            const int tok = c->prof.begin(IDP_K_FRICTION);
            launch_fric_hess(s, f.n, f.fcs.p, f.fdat.p, xd, f.Xn, f.epsvh, f.mu_eff, project, 0, b);
            c->prof.end(IDP_K_FRICTION, tok, (96.0 + 192.0 + 1152.0) * f.n);
            IDP_HIP(hipMemcpyAsync(H, b, 144 * (size_t)f.n * sizeof(double), hipMemcpyDeviceToHost, s));
        }
        IDP_HIP(hipStreamSynchronize(s));
        return IDP_OK;
    });
}

int idp_min_dist2(idp_ctx* c, const double* X, long n, const int* cs, double thickness, double* dist2,
                  double* min_dist2)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        set_constraints_from_host(c, n, cs);
        double* d = c->tmpv.ensure(std::max(n, 1L), s);
        double md = n ? min_dist2_dev(c, xd, n, (const int*)c->cw.cs.p, d) : INFINITY;
        if (dist2 && n) {
            IDP_HIP(hipMemcpyAsync(dist2, d, 8 * (size_t)n, hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
        }
        if (min_dist2) *min_dist2 = md - thickness * thickness;
        return IDP_OK;
    });
}

int idp_ccd_stepsize(idp_ctx* c, const double* X, const double* p, double thickness, double alpha_in,
                     double* alpha_out, long* n_candidates)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !p || !alpha_out) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        c->hp.upload(p, 3 * (size_t)c->m.nV, s);
        long nc = 0;
        *alpha_out = contact_ccd(c->cw, s, c->m, xd, c->hp.p, thickness, alpha_in, &nc);
        if (n_candidates) *n_candidates = nc;
        if (*alpha_out == 0.0) return IDP_E_CCD_ZERO;  // IPC.h:2014-2031
        return IDP_OK;
    });
}

int idp_set_accd_max_steps(idp_ctx* c, int n)
{
    return guarded(c, [&]() -> int {
        c->cw.accd_max_steps = n == 0 ? IDP_ACCD_MAX_STEPS_DEFAULT : (n < 0 ? 0 : n);
        return IDP_OK;
    });
}

int idp_set_accd_certificate(idp_ctx* c, int on)
{
    return guarded(c, [&]() -> int {
        c->cw.accd_cert = on != 0;
        return IDP_OK;
    });
}

int idp_ccd_last_stats(idp_ctx* c, long* capped, long* certified)
{
    return guarded(c, [&]() -> int {
        if (capped) *capped = c->cw.last_capped;
        if (certified) *certified = c->cw.last_certified;
        return IDP_OK;
    });
}

int idp_elastic_energy(idp_ctx* c, const double* X, double w, double* E, int* valid)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        int* fl = c->flags.ensure(4, s);
        IDP_HIP(hipMemsetAsync(fl, 0, sizeof(int), s));
        const int nb = c->m.inst.tblk.nblk;
        double* pe = c->part_e.ensure(nb + 1, s);
        launch_tet_energy(s, c->m, c->m.inst.tblk, xd, w, pe, fl);
        int hf = 0;
        IDP_HIP(hipMemcpyAsync(&hf, fl, sizeof(int), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        if (valid) *valid = hf ? 0 : 1;
        if (E) *E = hf ? 0.0 : dev_sum_partials(c->L, s, nb, pe);
        return IDP_OK;
    });
}

int idp_elastic_gradient(idp_ctx* c, const double* X, double w, double* g)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        double* gl = c->gl.ensure(12 * (size_t)c->m.nT, s);
        launch_tet_grad(s, c->m, xd, w, gl);
        c->L.ncb = 0;
        double* gd = c->g.ensure(3 * (size_t)c->m.nV, s);
        contact_gradient_gather(c->L, s, c->m, xd, nullptr, gl, nullptr, gd, nullptr);
        c->g.download(g, 3 * (size_t)c->m.nV, s);
        IDP_HIP(hipStreamSynchronize(s));
        return IDP_OK;
    });
}

int idp_elastic_hessian_local(idp_ctx* c, const double* X, double w, int project, double* H)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        const double* xd = stage_x(c, X);
        double* Hl = c->Hl.ensure(144 * (size_t)c->m.nT, s);
        launch_tet_hess_dense(s, c->m, xd, w, project, Hl);
        IDP_HIP(hipMemcpyAsync(H, Hl, 144 * 8 * (size_t)c->m.nT, hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        return IDP_OK;
    });
}

int idp_newton_direction_flow(idp_ctx* c, const double* X, const double* xtilde, double h, double dHat2,
                              double kappa, double thickness, const idp_solver_opts* opts, double* g, double* sol,
                              double* relres, int* iters)
{
    if (!c) return IDP_E_INVALID;
    struct ElasticOn {
        idp_ctx* c;
        ~ElasticOn() { c->no_elastic = false; }
    } restore_elastic{c};
    c->no_elastic = true;
    return idp_newton_direction(c, X, xtilde, h, dHat2, kappa, thickness, opts, g, sol, relres, iters);
}

int idp_newton_direction(idp_ctx* c, const double* X, const double* xtilde, double h, double dHat2, double kappa,
                         double thickness, const idp_solver_opts* opts, double* g, double* sol, double* relres,
                         int* iters)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !xtilde) return IDP_E_INVALID;
        auto s = c->s;
        auto& m = c->m;
        size_t n3 = 3 * (size_t)m.nV;
        const double* xd = stage_x(c, X);
        c->xt.upload(xtilde, n3, s);
        contact_constraint_set(c->cw, s, m, xd, &dHat2, thickness);
        update_cseg(c);
        struct FricOff {
            idp_ctx* c;
            ~FricOff() { c->fr.on = false; }
        } fric_off{c};
        if (c->fr.have && c->fr.enabled()) {  // the set and x_n of idp_friction_basis
            const double* unused = nullptr;
            if (friction_stage(c, h, nullptr, &unused) == IDP_OK) c->fr.on = true;
        }
        build_contact_structure(c);
        double* gd = c->g.ensure(n3, s);
        double* rhs = c->rhs.ensure(n3, s);
        double* sd = c->sol.ensure(n3, s);
        set_bparams1(c, dHat2, kappa, thickness);
        gradient(c, xd, h, true, gd, rhs);
        hessian(c, xd, h, true);
        double rtol = (opts && opts->pcg_rtol > 0) ? opts->pcg_rtol : 1e-8;
        int maxit = (opts && opts->pcg_maxit > 0) ? opts->pcg_maxit : 10000;
        c->L.precond = opts ? opts->precond : 0;
        c->L.coarse.verts_per_agg = coarse_from(opts);
        c->L.coarse.ready = false;  // built for this system
        double rel = 0;
        c->L.xpos = xd;
        int it = pcg_solve_persistent(c->L, s, m, rhs, sd, rtol, maxit, &rel);
        c->L.xpos = nullptr;
        if (g) c->g.download(g, n3, s);
        if (sol) c->sol.download(sol, n3, s);
        IDP_HIP(hipStreamSynchronize(s));
        if (relres) *relres = rel;
        if (iters) *iters = it;
        return it < 0 ? IDP_E_NOT_SPD : IDP_OK;
    });
}

int idp_get_bsr(idp_ctx* c, long* nnzb, int* rowptr, int* col, double* val, long cap)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh) return IDP_E_INVALID;
        auto s = c->s;
        auto& m = c->m;
        auto& L = c->L;
        // union of E and C patterns, row by row on the host
        std::vector<int> erp(m.nV + 1), ecol(m.nnzb_e), crp(m.nV + 1, 0), ccol(L.ncb);
        std::vector<double> ev(9 * m.nnzb_e), cv(9 * L.ncb);
        m.e_rowptr.download(erp.data(), m.nV + 1, s);
        m.e_col.download(ecol.data(), m.nnzb_e, s);
        std::vector<double> evp(9 * m.nnzb_e), cvp(9 * L.ncb);
        L.evals.download(evp.data(), 9 * m.nnzb_e, s);
        if (L.ncb) {
            L.crp.download(crp.data(), m.nV + 1, s);
            L.ccol.download(ccol.data(), L.ncb, s);
            L.cvals.download(cvp.data(), 9 * L.ncb, s);
        }
        IDP_HIP(hipStreamSynchronize(s));
        // device values are 9 component planes; regroup per block
        for (long k = 0; k < m.nnzb_e; ++k)
            for (int e = 0; e < 9; ++e) ev[9 * k + e] = evp[e * m.nnzb_e + k];
        for (long k = 0; k < L.ncb; ++k)
            for (int e = 0; e < 9; ++e) cv[9 * k + e] = cvp[e * L.ncb + k];
        std::vector<int> orp(1, 0), oc;
        std::vector<double> ov;
        for (int r = 0; r < m.nV; ++r) {
            std::vector<std::pair<int, const double*>> ent;
            for (int k = erp[r]; k < erp[r + 1]; ++k) ent.push_back({ecol[k], &ev[9 * (size_t)k]});
            if (L.ncb)
                for (int k = crp[r]; k < crp[r + 1]; ++k) ent.push_back({ccol[k], &cv[9 * (size_t)k]});
            std::stable_sort(ent.begin(), ent.end(),
                             [](const std::pair<int, const double*>& a, const std::pair<int, const double*>& b) {
                                 return a.first < b.first;
                             });
            for (size_t i = 0; i < ent.size();) {
                size_t j = i;
                double acc[9] = {0};
                while (j < ent.size() && ent[j].first == ent[i].first) {
                    for (int k = 0; k < 9; ++k) acc[k] += ent[j].second[k];
                    ++j;
                }
                oc.push_back(ent[i].first);
                ov.insert(ov.end(), acc, acc + 9);
                i = j;
            }
            orp.push_back((int)oc.size());
        }
        *nnzb = (long)oc.size();
        if (cap < (long)oc.size()) return IDP_E_CAPACITY;
        std::memcpy(rowptr, orp.data(), orp.size() * sizeof(int));
        std::memcpy(col, oc.data(), oc.size() * sizeof(int));
        std::memcpy(val, ov.data(), ov.size() * sizeof(double));
        return IDP_OK;
    });
}


// ---------------------------------------------------------------------------
// batched independent solves
// ---------------------------------------------------------------------------
int idp_set_mesh_batch(idp_ctx* c, int B, const int* nV, const int* nT, const double* X_rest, const int* tets,
                       const double* IB, const double* vol, const double* lambda, const double* mu,
                       const double* mass)
{
    return guarded(c, [&]() -> int {
        if (B <= 0 || !nV || !nT || !X_rest || !tets) return IDP_E_INVALID;
        std::vector<long> vbeg(B + 1, 0), tbeg(B + 1, 0);
        for (int i = 0; i < B; ++i) {
            if (nV[i] <= 0 || nT[i] <= 0) return IDP_E_INVALID;
            vbeg[i + 1] = vbeg[i] + nV[i];
            tbeg[i + 1] = tbeg[i] + nT[i];
            // the merge keys of PP/PE duplicates hold 3 local vertex ids in 56 bits
            if (B > 1 && (double)nV[i] * nV[i] * (nV[i] + 1.0) >= 7.2e16) return IDP_E_INVALID;
        }
        if (vbeg[B] > INT32_MAX / 4 || tbeg[B] > INT32_MAX / 16) return IDP_E_INVALID;
        std::vector<int> gt(4 * (size_t)tbeg[B]);
        for (int i = 0; i < B; ++i)
            for (long t = tbeg[i]; t < tbeg[i + 1]; ++t)
                for (int a = 0; a < 4; ++a) {
                    const int v = tets[4 * t + a];
                    if (v < 0 || v >= nV[i]) return IDP_E_INVALID;
                    gt[4 * t + a] = (int)(v + vbeg[i]);
                }
        return setup_mesh(c, (int)vbeg[B], X_rest, (int)tbeg[B], gt.data(), IB, vol, lambda, mu, mass, vbeg, tbeg);
    });
}

int idp_batch_size(idp_ctx* c, int* B)
{
    return guarded(c, [&]() -> int {
        if (!B) return IDP_E_INVALID;
        *B = c->have_mesh ? c->m.inst.B : 0;
        return IDP_OK;
    });
}

int idp_advance_batch(idp_ctx* c, double h, const double* newton_tol, int with_collision, const double* dHat2,
                      double* kappa3, double thickness, const idp_solver_opts* opts, int* newton_iters,
                      idp_step_stats* stats)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !kappa3 || !newton_tol || !dHat2 || !newton_iters || !(h > 0)) return IDP_E_INVALID;
        if (c->fr.enabled()) {
            c->err = "friction is stepped one mesh at a time (idp_advance_one_step)";
            return IDP_E_INVALID;
        }
        IDP_HIP(hipSetDevice(c->device));
        const int B = c->m.inst.B;
        BatchEngine& E = *c->eng;
        int rc = E.configure(c, h, with_collision != 0, thickness, opts);
        if (rc) return rc;
        std::vector<int> all(B);
        for (int i = 0; i < B; ++i) all[i] = i;
        rc = E.start(c, all, newton_tol, dHat2, kappa3);
        if (rc) return rc;
        const int verbose = opts ? opts->verbose : 0;
        const double t0 = now_ms();
        for (int k = 1; E.n_active > 0; ++k) {
            rc = E.iterate(c);
            if (rc) return rc;
            if (verbose > 0 && k % verbose == 0) {
                std::fprintf(stderr, "[idp] batch Newton iteration %d: %d of %d instances active, %.2f s\n", k,
                             E.n_active, B, (now_ms() - t0) * 1e-3);
                if (E.stage_log)
                    std::fprintf(stderr,
                                 "[idp]   stage s: grad+hess %.2f pcg %.2f ls1 %.2f ccd %.2f ls2 %.2f "
                                 "kappa+stop %.2f\n",
                                 E.stage_ms[0] * 1e-3, E.stage_ms[1] * 1e-3, E.stage_ms[2] * 1e-3,
                                 E.stage_ms[3] * 1e-3, E.stage_ms[4] * 1e-3, E.stage_ms[5] * 1e-3);
                std::fflush(stderr);
            }
        }
        for (int i = 0; i < B; ++i) {
            newton_iters[i] = E.PN[i];
            kappa3[3 * i] = E.kappa[i];
            if (stats) stats[i] = E.S[i];
        }
        return IDP_OK;
    });
}

int idp_serve_configure(idp_ctx* c, double h, int with_collision, double thickness, const idp_solver_opts* opts)
{
    return guarded(c, [&]() -> int {
        if (!c->have_mesh || !(h > 0)) return IDP_E_INVALID;
        if (c->fr.enabled()) {
            c->err = "friction is stepped one mesh at a time (idp_advance_one_step)";
            return IDP_E_INVALID;
        }
        return c->eng->configure(c, h, with_collision != 0, thickness, opts);
    });
}

int idp_serve_start(idp_ctx* c, int n, const int* slots, const double* X, const double* v, const double* newton_tol,
                    const double* dHat2, const double* kappa3)
{
    return guarded(c, [&]() -> int {
        BatchEngine& E = *c->eng;
        if (!c->have_mesh || !E.configured || n <= 0 || !slots || !X || !v || !newton_tol || !dHat2 || !kappa3)
            return IDP_E_INVALID;
        const InstDev& I = c->m.inst;
        std::vector<int> sl(slots, slots + n);
        // every slot must be free and listed once; nothing is copied before that holds
        std::vector<char> seen(I.B, 0);
        for (int k = 0; k < n; ++k) {
            const int i = sl[k];
            if (i < 0 || i >= I.B || E.active[i] || seen[i]) {
                c->err = "idp_serve_start: slot out of range, busy or listed twice";
                return IDP_E_INVALID;
            }
            seen[i] = 1;
        }
        std::vector<double> tol(I.B), d2(I.B), k3(3 * I.B);
        long off = 0;
        for (int k = 0; k < n; ++k) {
            const int i = sl[k];
            const long nv = I.vbeg[i + 1] - I.vbeg[i];
            IDP_HIP(hipMemcpyAsync(c->x.p + 3 * I.vbeg[i], X + 3 * off, 3 * nv * 8, hipMemcpyHostToDevice, c->s));
            IDP_HIP(hipMemcpyAsync(c->v.p + 3 * I.vbeg[i], v + 3 * off, 3 * nv * 8, hipMemcpyHostToDevice, c->s));
            off += nv;
            tol[i] = newton_tol[k];
            d2[i] = dHat2[k];
            for (int q = 0; q < 3; ++q) k3[3 * i + q] = kappa3[3 * k + q];
        }
        IDP_HIP(hipStreamSynchronize(c->s));
        return E.start(c, sl, tol.data(), d2.data(), k3.data(), true);
    });
}

int idp_serve_iterate(idp_ctx* c, int* n_active, int* finished)
{
    return guarded(c, [&]() -> int {
        BatchEngine& E = *c->eng;
        if (!E.configured) return IDP_E_INVALID;
        const int B = c->m.inst.B;
        if (finished) std::memset(finished, 0, B * sizeof(int));
        if (E.n_active > 0) {
            const int rc = E.iterate(c);
            if (rc) return rc;
            if (finished)
                for (int i = 0; i < B; ++i) finished[i] = E.finished_now[i];
        }
        if (n_active) *n_active = E.n_active;
        return IDP_OK;
    });
}

int idp_serve_result(idp_ctx* c, int slot, double* X, double* v, double* kappa3, idp_step_stats* stats)
{
    return guarded(c, [&]() -> int {
        const InstDev& I = c->m.inst;
        BatchEngine& E = *c->eng;
        if (!c->have_mesh || slot < 0 || slot >= I.B) return IDP_E_INVALID;
        const long nv = I.vbeg[slot + 1] - I.vbeg[slot];
        if (X) IDP_HIP(hipMemcpyAsync(X, c->x.p + 3 * I.vbeg[slot], 3 * nv * 8, hipMemcpyDeviceToHost, c->s));
        if (v) IDP_HIP(hipMemcpyAsync(v, c->v.p + 3 * I.vbeg[slot], 3 * nv * 8, hipMemcpyDeviceToHost, c->s));
        IDP_HIP(hipStreamSynchronize(c->s));
        if (kappa3 && (int)E.kappa.size() == I.B) {
            kappa3[0] = E.kappa[slot];
            kappa3[1] = E.kmax[slot];
            kappa3[2] = 0;
        }
        if (stats && (int)E.S.size() == I.B) *stats = E.S[slot];
        return IDP_OK;
    });
}

}  // extern "C"
