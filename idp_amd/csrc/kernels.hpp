// Device data structures and kernel launchers shared by the host code.
#pragma once
#include <array>
#include <vector>
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

// Blocks of a segmented reduction: the elements of instance i are
// [beg[i], beg[i+1]); each instance is cut into blocks of bs elements starting
// at its first element, so a block never straddles two instances and an
// instance's partials are [pbeg[i], pbeg[i+1]).  With one instance this is
// the plain grid of ceil(n / bs) blocks.
struct SegDev {
    int nblk = 0, bs = 0;
    DBuf<int> lo, hi, inst, pbeg;
    std::vector<int> h_pbeg, h_lo, h_hi, h_inst;
    void build(const std::vector<long>& beg, int block, hipStream_t s)
    {
        bs = block;
        const int B = (int)beg.size() - 1;
        std::vector<int> l, h, in;
        h_pbeg.assign(B + 1, 0);
        for (int i = 0; i < B; ++i) {
            h_pbeg[i] = (int)l.size();
            for (long a = beg[i]; a < beg[i + 1]; a += block) {
                l.push_back((int)a);
                h.push_back((int)std::min<long>(a + block, beg[i + 1]));
                in.push_back(i);
            }
        }
        h_pbeg[B] = (int)l.size();
        nblk = (int)l.size();
        lo.upload(l.data(), l.size(), s);
        hi.upload(h.data(), h.size(), s);
        inst.upload(in.data(), in.size(), s);
        pbeg.upload(h_pbeg.data(), h_pbeg.size(), s);
        h_lo = std::move(l);
        h_hi = std::move(h);
        h_inst = std::move(in);
    }
};

// Instance layout of a mesh holding B independent meshes, each a contiguous
// range of vertices, tets and boundary primitives (a plain mesh is B = 1).
// Contact detection, CCD step sizes, energies and the PCG are computed per
// instance; instances never interact.
struct InstDev {
    int B = 0;
    std::vector<long> vbeg, tbeg, nbeg, ebeg, fbeg;  // host ranges, B + 1 each
    DBuf<int> vinst;                                 // vertex -> instance
    DBuf<int> d_vbeg, d_nbeg, d_ebeg, d_fbeg;        // device copies
    SegDev echunk;   // boundary edges, chunks of 256 (mean edge length)
    SegDev vblk;     // vertices, blocks of RB (bounding box, inertia energy)
    SegDev nblk;     // boundary nodes, blocks of RB (swept bounding box)
    SegDev tblk;     // tets, blocks of RB (elastic energy)
    SegDev sblk;     // shell triangles, blocks of RB (flow energy)
    SegDev hblk;     // shell hinges, blocks of RB (bending energy)
    SegDev pcg_rows[4];  // vertices, blocks of 256 / G rows for the SpMV with G = 4, 8, 16, 32 lanes per row
    SegDev pcg_upd;  // vertices, 64-row blocks (PCG update)
    SegDev dblk;     // 3 x vertices, blocks of RB (per-instance dot products)
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
    // shell triangles (Elem): elements nT .. nT + nE - 1, vol = area * thickness
    int nE = 0;
    DBuf<int> tris;       // 3 per triangle
    DBuf<double> tvol;
    DBuf<int> vt_ptr, vt; // vertex -> its triangles, ascending (nV + 1, 3 nE)
    // non-flow shell terms: rest first fundamental form (3 per triangle),
    // Lame parameters, hinges (elements nT + nE + k: 4 vertex ids, rest
    // (theta, |e|, h)) and the bending stiffness
    bool shell_el = false;
    DBuf<double> tIB, tlam, tmu, hrest;
    int nH = 0;
    DBuf<int> hinge;
    double bend_k = 0.0;
    // rows/columns projected out of the gradient and the Hessian: the DBC
    // nodes, or only the ones that do not move while a moving-DBC penalty is
    // on (null: dbc)
    const unsigned char* proj = nullptr;
    const unsigned char* proj_mask() const { return proj ? proj : dbc.p; }
    DBuf<int> bn;         // boundary nodes
    DBuf<int> be;         // boundary edges, 2 per edge
    DBuf<int> bt;         // boundary triangles, 3 per tri
    DBuf<int> v2b;        // vertex -> boundary node index (-1 if interior)
    // static elastic BSR pattern (tet connectivity + diagonal)
    DBuf<int> e_rowptr, e_col;       // nV+1, nnzb_e
    DBuf<int> e_src_ptr, e_src;      // per block: contributing (element*16 + a*4 + b)
    DBuf<int> e_diag;                // block index of (v,v)
    long nnzb_e = 0;
    InstDev inst;                    // instance layout (one instance for a plain mesh)
};

// device surface extraction work buffers
struct SurfWork {
    DBuf<unsigned> cnt;
    DBuf<unsigned long long> k0, k1, v0, v1;
    DBuf<unsigned char> tmp, node;
    DBuf<int> tri, edge, flag, pos, bn;
};
// Boundary nodes / edges / triangles of a tet mesh (tets and X on the
// device) in the order of Find_Surface_TriMesh + Find_Surface_Primitives
// (MESHIO.h:486-545, 769-835), returned on the host.  false when the mesh is
// beyond the key widths (more than 2^21 vertices or 1.4M boundary triangles).
bool surface_device(hipStream_t s, SurfWork& w, int nV, const double* X, int nT, const int* tets,
                    std::vector<int>& bn, std::vector<std::array<int, 2>>& be, std::vector<std::array<int, 3>>& bt);

struct ContactWork {
    DBuf<unsigned char> tmp;
    DBuf<double> dparts, dparts2, mmparts, scal, scal2;
    DBuf<GridDev> grid;
    DBuf<int> vbox, pbox;
    DBuf<unsigned> u0, u1, cP, cD, oP, oD, k0, k1;
    DBuf<int> v0, v1;
    DBuf<int4> pass, dup, merged, cs;
    DBuf<unsigned long long> dk0, dk1, pk0, pk1;
    DBuf<unsigned> cstart, cend, cnt;  // dense cell table, atomic counters
    DBuf<int2> cand;
    // per-instance data of the batched path
    DBuf<GridDev> grids;
    DBuf<unsigned> keybase, pinst;
    DBuf<double> ialpha;                   // CCD: capped alpha per instance
    DBuf<double> rad;                      // per-instance contact radius and its square
    DBuf<unsigned long long> iamin;        // CCD: running minimum per instance
    DBuf<double> accd_log;                 // CCD: pairs whose ACCD iteration ran long
    DBuf<unsigned char> accd_res[2];       // CCD: pair states handed from one call to the next (ping-pong)
    unsigned n_res = 0;                    // records in accd_res[res_cur]
    int res_cur = 0;
    std::vector<unsigned char> res_inst;   // instances owning those records
    DBuf<unsigned char> pend;              // CCD: per instance, a pair was handed over
    DBuf<unsigned> walk_big;               // CCD: queries whose swept box is walked by a workgroup
    DBuf<unsigned> icnt;                   // per-instance counters
    DBuf<long> cbeg;                       // constraint range per instance
    std::vector<long> h_cbeg;
    DBuf<int4> cs_old;
    size_t cap_pass = 0, cap_dup = 0, cap_cand = 0;
    long ncs = 0;
    // ACCD step cap per pair and call (<= 0: none, as in CCD.h:308-327); a
    // capped pair counts as a hit at the time of impact it reached
    int accd_max_steps = 1 << 20;
    long last_capped = 0;     // pairs capped in the last CCD call
    long last_certified = 0;  // pairs ended by the no-hit certificate in the last CCD call
    bool accd_cert = true;    // long ACCD pairs try the no-hit certificate
    std::vector<unsigned> h_capped;  // ... per instance
    DBuf<unsigned> icap;
    Profiler* prof = nullptr;
};

// barrier.hip
// Barrier parameters per instance (device arrays of B entries); a constraint's
// instance is that of its first vertex (vinst null: instance 0).
struct BarrierParams {
    const double* dh;     // dHat2 + 2 sqrt(dHat2) thickness
    const double* kappa;
    const double* th2;    // thickness^2
    const int* vinst;
    const unsigned char* act;  // instances whose constraints are evaluated (null = all); gradient and Hessian only
    __device__ int inst_of(int v) const { return vinst ? vinst[v] : 0; }
    __device__ bool skip(int inst) const { return act && !act[inst]; }
};
void barrier_init_constants();
// (B B^T)^(1/2) and its inverse for the order-9 map of the point-triangle
// coordinates (y = (x0 - x1, x2 - x1, x3 - x1)) and the order-6 map
// (y = (x1 - x0, x2 - x0) up to vertex order), row-major
struct ReducedRoots {
    double S3[9], Si3[9], S2[4], Si2[4];
};
ReducedRoots reduced_roots();
void launch_barrier_energy(hipStream_t s, int nblk, const int* lo, const int* hi, const int* cs, const double* X,
                           const double* X0, const BarrierParams& bp, double* part, int* err);
void launch_min_dist2(hipStream_t s, long n, const int* cs, const double* X, const int* vinst, double* dist2,
                      unsigned long long* dmin);
void launch_barrier_grad(hipStream_t s, long n, const int* cs, const double* X, const double* X0,
                         const BarrierParams& bp, double* gout, int* err);
// perm (n ints) and kcnt (barrier_kind_scratch(n) unsigned), both device scratch: the
// launch first groups the constraints by kind (both null: one constraint per lane in
// index order)
// form (90 n doubles) and form_ok (n bytes): device scratch of the two stages; team: lanes per
// constraint of the projection stage (8, 16, 32, 64; results do not depend on it)
inline size_t barrier_kind_scratch(long n) { return 16 + 16 * (size_t)((n + 255) / 256); }
void launch_barrier_hess(hipStream_t s, long n, const int* cs, const double* X, const double* X0,
                         const BarrierParams& bp, int project_spd, int mode, double* out, int* vid, int* err,
                         double* form, unsigned char* form_ok, int* counts = nullptr, int* perm = nullptr,
                         unsigned* kcnt = nullptr, int team = 16);
void launch_constraint_vids(hipStream_t s, long n, const int* cs, int* vid, int* nv);

// contact.hip
// dHat2: host array, one per instance
long contact_constraint_set(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* dHat2,
                            double thickness, const unsigned char* active = nullptr);
double contact_ccd(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                   double thickness, double alpha_in, long* ncand);
// CCD step sizes of the active instances (device mask; null = all).
// budget > 0 bounds the ACCD steps a pair may take in this call: a pair that
// runs out is handed over, state and all, to the next call made with resume
// set, and its instance is flagged in pending (host, B) with no step size
// yet.  Handing over changes no arithmetic: the result equals an unbounded
// call's.
void contact_ccd_batch(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                       double thickness, double* alpha_io, const unsigned char* active, long* ncand,
                       int budget = 0, bool resume = false, unsigned char* pending = nullptr);

// linalg.hip: instance layout (ranges are B + 1 offsets; vinst: vertex -> instance)
void build_inst_layout(MeshDev& m, const std::vector<long>& vbeg, const std::vector<long>& tbeg,
                       const std::vector<long>& nbeg, const std::vector<long>& ebeg, const std::vector<long>& fbeg,
                       hipStream_t s);

// fem.hip (energies over block tables: lo/hi element ranges, binst block -> instance)
void launch_tet_energy(hipStream_t s, const MeshDev& m, const SegDev& seg, const double* X, double w, double* part,
                       int* invalid);
// act (device, one byte per instance; null = all): instances whose elements are evaluated
void launch_tet_grad(hipStream_t s, const MeshDev& m, const double* X, double w, double* gl,
                     const unsigned char* act = nullptr);
void launch_tet_hess(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* Hl,
                     const unsigned char* act = nullptr);
// shell.hip: the flow term over the shell triangles and the normal-flow force
void launch_flow_energy(hipStream_t s, const MeshDev& m, const double* X, double h, double* part);
void launch_flow_grad(hipStream_t s, const MeshDev& m, const double* X, double h, double* gl);
void launch_flow_hess(hipStream_t s, const MeshDev& m, double h, double* Hl);
void launch_normal_flow(hipStream_t s, const MeshDev& m, const double* X, double magnitude, double* b);
// membrane and bending (w = h^2): energy partials over sblk / hblk, gradient and
// projected Hessian into the element slots (F, ok: 90 nH doubles, nH bytes of scratch)
void launch_shell_energy(hipStream_t s, const MeshDev& m, const double* X, double w, double* pm, double* ph);
void launch_shell_grad(hipStream_t s, const MeshDev& m, const double* X, double w, double* gl);
void launch_shell_hess(hipStream_t s, const MeshDev& m, const double* X, double w, int project, double* F,
                       unsigned char* ok, double* Hl);
void launch_inertia_energy(hipStream_t s, const SegDev& seg, const double* X, const double* Xt, const double* mass,
                           double* part);

// friction.hip: lagged friction.  fric_basis_update compacts the
// non-mollified constraints of cs into fcs with their 10-double records
// (closest points 2, tangent basis 6, normal force, pad) and returns their
// count; err[0] = 1 when a vertex is in no component.  The energy writes
// per-block partial sums over grid_for(nf, RB) blocks (mu not applied); the
// gradient 12 doubles and the Hessian 144 doubles per constraint (mode 1:
// assembly blocks, mode 0: dense row-major).
long fric_basis_update(hipStream_t s, DBuf<unsigned char>& tmp, long n, const int* cs, const double* X, double dh,
                       double th2, double kappa, int ncomp, const int* comp_range, const double* mu_comp,
                       DBuf<unsigned>& flag, DBuf<unsigned>& pos, DBuf<int4>& fcs, DBuf<double>& fdat, int* err);
void launch_fric_energy(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                        const double* Xn, double epsvh, double* part);
void launch_fric_grad(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                      const double* Xn, double epsvh, double mu, double* gout);
void launch_fric_hess(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                      const double* Xn, double epsvh, double mu, int project, int mode, double* out);

}  // namespace idp
