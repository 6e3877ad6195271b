// Device data structures and kernel launchers shared by the host code.
#pragma once
#include "common.hpp"

namespace idp {

// voxel grid of the spatial hash (SPATIAL_HASH.h: leftBottomCorner,
// one_div_voxelSize, voxelCount)
struct GridDev {
    double lbc[3];
    double inv;
    double voxel;
    int cnt[3];
    int pad_;
};

// Mesh resident on the device.  Vertex arrays are 3-vectors interleaved
// (x0,y0,z0,x1,...): a vertex's coordinates share one 32-byte sector for the
// gathers of the contact and element kernels, and the Newton vectors match
// the 3x3 BSR block layout.
struct MeshDev {
    int nV = 0, nT = 0, nb = 0, ne = 0, nf = 0;
    DBuf<double> X0;      // rest positions (mollifier thresholds)
    DBuf<double> mass;    // lumped mass per vertex
    DBuf<int> tets;       // 4 per tet
    DBuf<double> IB;      // 9 per tet, row-major
    DBuf<double> vol, lam, mu;
    DBuf<unsigned char> dbc;
    DBuf<int> bn;         // boundary nodes
    DBuf<int> be;         // boundary edges, 2 per edge
    DBuf<int> bt;         // boundary triangles, 3 per tri
    DBuf<int> v2b;        // vertex -> boundary node index (-1 if interior)
    // static elastic BSR pattern (tet connectivity + diagonal)
    DBuf<int> e_rowptr, e_col;       // nV+1, nnzb_e
    DBuf<int> e_src_ptr, e_src;      // per block: contributing (tet*16 + a*4 + b)
    DBuf<int> e_diag;                // block index of (v,v)
    long nnzb_e = 0;
};

struct ContactWork {
    DBuf<unsigned char> tmp;
    DBuf<double> dparts, dparts2, mmparts, scal;
    DBuf<GridDev> grid;
    DBuf<int> vbox, pbox;
    DBuf<unsigned> u0, u1, cP, cD, oP, oD, k0, k1;
    DBuf<int> v0, v1;
    DBuf<int4> pass, dup, merged, cs;
    DBuf<unsigned long long> dk0, dk1, pk0, pk1;
    DBuf<unsigned> cstart, cend, cnt;  // dense cell table, atomic counters
    DBuf<int2> cand;
    size_t cap_pass = 0, cap_dup = 0, cap_cand = 0;
    long ncs = 0;
    Profiler* prof = nullptr;
};

// barrier.hip
void barrier_init_constants();
void launch_barrier_energy(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                           double kappa, double th2, double* part, int* err);
int barrier_energy_blocks(long n);
void launch_min_dist2(hipStream_t s, long n, const int* cs, const double* X, double* dist2,
                      unsigned long long* dmin);
void launch_barrier_grad(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                         double kappa, double th2, double* gout, int* err);
void launch_barrier_hess(hipStream_t s, long n, const int* cs, const double* X, const double* X0, double dh,
                         double kappa, double th2, int project_spd, int mode, double* out, int* vid, int* err,
                         int* counts = nullptr);
void launch_constraint_vids(hipStream_t s, long n, const int* cs, int* vid, int* nv);

// contact.hip
long contact_constraint_set(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, double dHat2,
                            double thickness);
double contact_ccd(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                   double thickness, double alpha_in, long* ncand);

// fem.hip
int tet_energy_blocks(int nT);
void launch_tet_energy(hipStream_t s, const MeshDev& m, const double* X, double w, double* part, int* invalid);
void launch_tet_grad(hipStream_t s, const MeshDev& m, const double* X, double w, double* gl);
void launch_tet_hess(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* Hl);
void launch_inertia_energy(hipStream_t s, int nV, const double* X, const double* Xt, const double* mass,
                           double* part);
int inertia_energy_blocks(int nV);

}  // namespace idp
