/*
 * idp_amd_step.hpp - host C++ session over the C ABI for callers that step
 * one mesh of tets and/or shell triangles per call, with the argument
 * meaning of
 * Advance_One_Step_IE_Discrete_Shell<double,3,false,false,flow>
 * (Library/FEM/Shell/IMPLICIT_EULER.h:151-187): positions and velocities are
 * updated in place, kappa3[0] carries the barrier stiffness from step to
 * step, and the return value is the Newton iteration count.
 *
 * The mesh (topology, rest positions, inverse bases, Lame parameters, mass,
 * and for shells the triangles, hinges and their rest data) is uploaded only
 * when it differs from the one the device holds, so a
 * driver that keeps its topology calls set_mesh every step at the cost of a
 * host-side comparison.  Header-only; link with -lidp_amd.
 */
#ifndef IDP_AMD_STEP_HPP
#define IDP_AMD_STEP_HPP

#include <cstring>
#include <string>
#include <vector>

#include "idp_amd.h"

namespace idp_amd {

struct TetMesh {
    int nV = 0, nT = 0;  // nT may be 0 for a shell-only mesh
    const double* X_rest = nullptr;  // 3 nV, nodeAttr.x0
    const int* tets = nullptr;       // 4 nT
    const double* IB = nullptr;      // 9 nT, tetAttr inverse rest basis, row-major
    const double* vol = nullptr;     // nT
    const double* lambda = nullptr;  // nT
    const double* mu = nullptr;      // nT
    const double* mass = nullptr;    // nV, nodeAttr.m
};

/* The shell part of the step's arguments (Elem, elemAttr, elasticityAttr,
 * edgeStencil, edgeInfo, bendingStiffMult): membrane and hinge bending in the
 * Hinge variant, the -h/2 x^T L x term in the Flow variant. */
struct ShellPart {
    int nE = 0;
    const int* tris = nullptr;          // 3 nE, Elem
    const double* vol = nullptr;        // nE, elasticityAttr vol (area * thickness)
    const double* lambda = nullptr;     // nE
    const double* mu = nullptr;         // nE
    const double* IB = nullptr;         // 3 nE: elemAttr IB(0,0), IB(0,1), IB(1,1)
    int nH = 0;                         // 0 when bendingStiffMult == 0
    const int* hinges = nullptr;        // 4 nH, edgeStencil
    const double* hinge_rest = nullptr; // 3 nH, edgeInfo
    double bend_k = 0;                  // bendingStiffMult * elemAttr(0).P(0,0)
};

class StepSession {
public:
    explicit StepSession(int device = 0) { status_ = idp_ctx_create(device, &ctx_); }
    ~StepSession()
    {
        if (ctx_) idp_ctx_destroy(ctx_);
    }
    StepSession(const StepSession&) = delete;
    StepSession& operator=(const StepSession&) = delete;

    /* IDP_OK when the device context exists */
    int status() const { return status_; }
    const char* error() const { return ctx_ ? idp_last_error(ctx_) : "no device context"; }
    /* how many times the mesh went to the device */
    int mesh_uploads() const { return uploads_; }

    int set_mesh(const TetMesh& m, const ShellPart* sh = nullptr)
    {
        if (!ctx_) return status_;
        const bool shell = sh && sh->nE > 0;
        if (m.nV <= 0 || m.nT < 0 || (m.nT == 0 && !shell)) return IDP_E_INVALID;
        std::vector<unsigned char> key = pack(m, shell ? sh : nullptr);
        if (!held_.empty() && key == held_) return IDP_OK;
        int rc;
        if (shell) {
            rc = idp_set_mesh_shell(ctx_, m.nV, m.X_rest, m.mass, m.nT, m.tets, m.IB, m.vol, m.lambda, m.mu, sh->nE,
                                    sh->tris, sh->vol, sh->nH, sh->hinges);
            if (rc == IDP_OK) rc = idp_set_shell_elasticity(ctx_, sh->IB, sh->lambda, sh->mu, sh->hinge_rest, sh->bend_k);
        } else {
            rc = idp_set_mesh(ctx_, m.nV, m.X_rest, m.nT, m.tets, m.IB, m.vol, m.lambda, m.mu, m.mass);
        }
        if (rc != IDP_OK) {
            held_.clear();
            return rc;
        }
        held_.swap(key);
        ++uploads_;
        return IDP_OK;
    }

    /* DBC rows (node, x, y, z): node idx[i] is driven to targets[3 i..3 i+2] */
    int set_dbc(int n, const int* idx, const double* targets)
    {
        if (!ctx_) return status_;
        return idp_set_dbc_targets(ctx_, n, idx, targets);
    }

    /* One step.  X, v: 3 nV, in/out.  b: body force (3 nV) or null.  flow:
     * the Advance_One_Step_IE_Flow variant.  kappa3: {kappa, kappa_max, 0},
     * kappa3[0] in/out.  Returns the Newton count or a negative IDP_E_* status. */
    int step(double* X, double* v, const double* b, bool flow, double h, double newton_tol, bool with_collision,
             double dHat2, double* kappa3, double thickness, const idp_solver_opts* opts = nullptr,
             idp_step_stats* stats = nullptr)
    {
        if (!ctx_) return status_;
        if (held_.empty()) return IDP_E_INVALID;
        int rc = idp_set_state(ctx_, X, v);
        if (rc != IDP_OK) return rc;
        const int it = idp_advance_one_step(ctx_, h, newton_tol, with_collision ? 1 : 0, dHat2, kappa3, thickness, b,
                                            flow ? 1 : 0, opts, stats);
        if (it < 0) return it;
        rc = idp_get_state(ctx_, X, v);
        return rc == IDP_OK ? it : rc;
    }

    idp_ctx* context() { return ctx_; }

private:
    template <class T>
    static void put(std::vector<unsigned char>& o, const T* p, size_t n)
    {
        if (n == 0) return;
        const size_t k = o.size();
        o.resize(k + n * sizeof(T));
        std::memcpy(o.data() + k, p, n * sizeof(T));
    }
    static std::vector<unsigned char> pack(const TetMesh& m, const ShellPart* sh)
    {
        std::vector<unsigned char> o;
        put(o, &m.nV, 1);
        put(o, &m.nT, 1);
        put(o, m.X_rest, 3 * (size_t)m.nV);
        put(o, m.tets, 4 * (size_t)m.nT);
        put(o, m.IB, 9 * (size_t)m.nT);
        put(o, m.vol, (size_t)m.nT);
        put(o, m.lambda, (size_t)m.nT);
        put(o, m.mu, (size_t)m.nT);
        put(o, m.mass, (size_t)m.nV);
        const int nE = sh ? sh->nE : 0, nH = sh ? sh->nH : 0;
        put(o, &nE, 1);
        put(o, &nH, 1);
        if (sh) {
            put(o, sh->tris, 3 * (size_t)nE);
            put(o, sh->vol, (size_t)nE);
            put(o, sh->lambda, (size_t)nE);
            put(o, sh->mu, (size_t)nE);
            put(o, sh->IB, 3 * (size_t)nE);
            put(o, sh->hinges, 4 * (size_t)nH);
            put(o, sh->hinge_rest, 3 * (size_t)nH);
            put(o, &sh->bend_k, 1);
        }
        return o;
    }

    idp_ctx* ctx_ = nullptr;
    int status_ = IDP_E_NO_GPU;
    int uploads_ = 0;
    std::vector<unsigned char> held_;
};

}  // namespace idp_amd

#endif
