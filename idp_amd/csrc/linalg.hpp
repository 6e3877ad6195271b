// Host-side interface of the BSR assembly and PCG (linalg.hip).
#pragma once
#include <vector>

#include "kernels.hpp"

namespace idp {

// coarse level of the one-mesh PCG preconditioner (coarse.hip)
struct CoarseWork {
    int verts_per_agg = 0;  // target free vertices per aggregate; 0: no coarse level
    int rebuild_every = 4;  // Newton systems per build: A_c^-1 and P of an earlier system still give an
                            // SPD preconditioner, so they are reused in between
    bool ready = false;     // a level exists (built for this mesh and Dirichlet set)
    int age = 0, nV = 0;
    const unsigned char* mask = nullptr;  // the Dirichlet mask it was built with
    int K = 0, ncd = 0, nsb = 0;
    DBuf<int> agg, aptr, avert;
    DBuf<double> P, Ainv, rc, part;
    DBuf<double> Ps;  // P in aggregate order (18 doubles per slot of avert)
    DBuf<unsigned long long> key0, key1;
    DBuf<unsigned> val0, val1, head, hs, seg;
    DBuf<unsigned char> tmp;
    DBuf<int> info;
    std::vector<double> hx;
    void* handle = nullptr;  // rocblas_handle
    CoarseWork() = default;
    CoarseWork(const CoarseWork&) = delete;
    CoarseWork& operator=(const CoarseWork&) = delete;
    ~CoarseWork();
};

struct LinWork {
    DBuf<unsigned char> tmp;
    // contact structure of the current constraint set
    DBuf<int> vid, nv, crow, ccol, cdiag, crp;
    DBuf<unsigned> u0, u1, seg, rv0, rv1;
    DBuf<unsigned long long> key0, key1;
    long ncb = 0, nrec = 0, ncs = 0;
    int precond = 0;  // 0: cluster block Jacobi over elastic and contact couplings, 1: 3x3 block Jacobi,
                      // 2: clusters over contact couplings only
    int upd_waves = 4;   // PCG update: 64-row wavefronts per workgroup (1..8; same arithmetic for any value)
    int spmv_shape = 5;  // PCG SpMV launch shape: 8 lanes per row, scalars in a wavefront of their own
    DBuf<int4> atab, btab;  // PCG launch tables
    std::vector<int4> h_atab, h_btab;
    DBuf<int> aoff, boff;   // per launched block: its instance's iterations in earlier calls
    std::vector<int> h_aoff, h_boff, h_it0;
    std::vector<char> h_maxit;  // per instance: the last solve stopped at the iteration limit, not converged
    DBuf<unsigned char> solve_d;
    // values
    DBuf<int> erow;  // row of every elastic block
    DBuf<double> evals, cvals;
    // PCG
    DBuf<double> pinv, r, z, p, p1, q, part0, part1, part2, part3, part_bx, scal, scal2, red;
    DBuf<int> err;
    // cluster preconditioner
    DBuf<int> cl_best, cl_root, cl_size, cl_root2, cl_cnt, cl_mem, cl_cm4;
    DBuf<unsigned long long> cl_pack;
    DBuf<double> cl_work;
    Profiler* prof = nullptr;
    CoarseWork coarse;
    const double* xpos = nullptr;  // positions of the system being solved (coarse aggregation); null: none
};

// Builds the coarse level for the assembled system at positions X (device);
// false when it is not used for this system (too few vertices, several
// instances, A_c not positive definite).
bool coarse_setup(LinWork& L, hipStream_t s, const MeshDev& m, const double* X);
// z += P A_c^-1 P^T r and the partials of r_c . e_c at part + parity * nsb
void coarse_apply(LinWork& L, hipStream_t s, int nV, const double* r, double* z, const int* flag, int parity);

// the contact blocks of the ncs barrier constraints cs followed by nf more
// stencils in the same int4 encoding (the friction set); local terms of entry
// i are at gb + 12 i and Hb + 144 i over both lists
void contact_structure(LinWork& L, hipStream_t s, int nV, long ncs, const int* cs, long nf = 0,
                       const int* fcs = nullptr);
// act (device, one byte per instance; null = all): the instances whose rows are formed
void contact_gradient_gather(LinWork& L, hipStream_t s, const MeshDev& m, const double* X, const double* Xt,
                             const double* gl, const double* gb, double* g, double* rhs,
                             const unsigned char* act = nullptr);
void assemble(LinWork& L, hipStream_t s, const MeshDev& m, const double* Hl, const double* Hb, bool any_dbc,
              const unsigned char* act = nullptr);
void init_erow(LinWork& L, hipStream_t s, const MeshDev& m, const std::vector<int>& rowof);
// solve_h (host, B; null = all new): 0 not solving, 1 new solve, 2 continue
// a solve handed over by the previous call.  warm (device, B bytes; null =
// none): new solves that start from a multiple of what x holds on entry
// (their previous direction) instead of zero.  handover: once half of the
// solves have stopped, the rest get as many iterations again and are then
// handed over (iters = PCG_HANDED_OVER); a hand-over changes no arithmetic.
constexpr int PCG_HANDED_OVER = -3;
void pcg_solve_batch(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                     int maxit, const unsigned char* solve_h, int* iters, double* relres,
                     const unsigned char* warm = nullptr, bool handover = false);
int pcg_solve_persistent(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                         int maxit, double* relres, const unsigned char* warm = nullptr);
double dev_dot(LinWork& L, hipStream_t s, long n, const double* a, const double* b);
void dev_velocity(hipStream_t s, long n, const double* x, const double* xn, double h, double* v);
void dev_axpy(hipStream_t s, long n, const double* a, double sc, const double* b, double* x);
double dev_sum_partials(LinWork& L, hipStream_t s, int npart, const double* part);
// out[0] = sum of part[0, npart) in the fixed order of dev_sum_partials (device result, no copy)
void launch_sum_partials(hipStream_t s, int npart, const double* part, double* out);
void launch_tet_hess_dense(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* Hl);

}  // namespace idp
