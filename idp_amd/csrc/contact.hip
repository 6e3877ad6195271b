// Contact detection on the device:
//   constraint set  Library/FEM/IPC.h:19-740 with the static spatial hash of
//                   Library/Grid/SPATIAL_HASH.h:28-291
//   ACCD step size  Library/FEM/IPC.h:1879-2244 with the swept hash of
//                   Library/Grid/SPATIAL_HASH.h:432-690
//
// Broad phase.  The reference hashes primitives into a uniform voxel grid
// (voxel = mean boundary edge length) and pairs primitives that share a
// voxel.  Here every primitive gets an integer voxel box; boxes are expanded
// into (voxel key, primitive) records, radix-sorted by key, and each query
// primitive walks the voxels of its own box, binary-searching the sorted
// keys.  A pair found in several shared voxels is kept only in the lowest
// shared voxel (componentwise max of the two box minima), so no hash set is
// needed.  The narrow phase then applies the reference's FP64 filters in the
// reference's order (incidence, all-DBC, AABB test, type, distance), so the
// emitted set is the reference's.  Counts are produced first, then an
// exclusive scan places every query's output; the output order is therefore
// deterministic.
#include "common.hpp"
#include "geom_dev.hpp"
#include "kernels.hpp"

#include <hipcub/hipcub.hpp>

namespace idp {

// ---------------------------------------------------------------------------
// grid set-up (SPATIAL_HASH.h:60-84 and 455-511), one thread
// ---------------------------------------------------------------------------
__global__ void k_grid_setup(GridDev* g, const double* __restrict__ mean_parts, int n_mean_parts, long n_mean,
                             const double* __restrict__ mm /* minx,miny,minz,maxx,maxy,maxz */, double pad)
{
    double tot = 0.0;
    for (int c = 0; c < n_mean_parts; ++c) tot += mean_parts[c];
    double vs = n_mean ? tot / (double)n_mean : 1.0;
    g->voxel = vs;
    double lo[3] = {mm[0] - pad, mm[1] - pad, mm[2] - pad};
    double hi[3] = {mm[3] + pad, mm[4] + pad, mm[5] + pad};
    double range[3] = {hi[0] - lo[0], hi[1] - lo[1], hi[2] - lo[2]};
    double inv = 1.0 / vs;
    long vcl[3];
    for (int d = 0; d < 3; ++d) {
        long c = (long)ceil(range[d] * inv);
        vcl[d] = c < 1 ? 1 : c;
    }
    long amt = vcl[0] * vcl[1] * vcl[2];
    if (amt > 1e9) {
        vs *= pow(amt / 1.0e9, 1.0 / 3);
        inv = 1.0 / vs;
    }
    bool bad = false;
    for (int d = 0; d < 3; ++d) {
        int c = (int)ceil(range[d] * inv);
        g->cnt[d] = c < 1 ? 1 : c;
        if (g->cnt[d] <= 0) bad = true;
    }
    if (bad) {
        inv = 1.0 / (fmax(range[0], fmax(range[1], range[2])) * 1.01);
        g->cnt[0] = g->cnt[1] = g->cnt[2] = 1;
    }
    for (int d = 0; d < 3; ++d) g->lbc[d] = lo[d];
    g->inv = inv;
}

__device__ inline void locate(const GridDev& g, double x, double y, double z, int* o)
{
    o[0] = (int)floor((x - g.lbc[0]) * g.inv);
    o[1] = (int)floor((y - g.lbc[1]) * g.inv);
    o[2] = (int)floor((z - g.lbc[2]) * g.inv);
}

__device__ inline unsigned key_of(const GridDev& g, int x, int y, int z)
{
    return (unsigned)x + (unsigned)g.cnt[0] * ((unsigned)y + (unsigned)g.cnt[1] * (unsigned)z);
}

// chunked partial sums of boundary edge lengths (chunks of 256, sequential)
__global__ void k_edge_len_parts(int ne, const int2* __restrict__ E, const double* __restrict__ X,
                                 double* __restrict__ parts)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    long b = (long)c * 256;
    if (b >= ne) return;
    long e = b + 256 < ne ? b + 256 : ne;
    double s = 0.0;
    for (long i = b; i < e; ++i) {
        int2 ed = E[i];
        s += sqrt(sqn(ld3(X, ed.x) - ld3(X, ed.y)));
    }
    parts[c] = s;
}

// chunked partial sums of |p| over boundary nodes (x, y, z in order)
__global__ void k_psize_parts(int nb, const int* __restrict__ bn, const double* __restrict__ p,
                              double* __restrict__ parts)
{
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    long b = (long)c * 256, n = 3L * nb;
    if (b >= n) return;
    long e = b + 256 < n ? b + 256 : n;
    double s = 0.0;
    for (long i = b; i < e; ++i) s += fabs(p[3 * bn[i / 3] + (i % 3)]);
    parts[c] = s;
}

// min / max of positions: per-block partials, exact (order independent)
__global__ void __launch_bounds__(RB) k_minmax_parts(long n, const int* __restrict__ idx,
                                                     const double* __restrict__ X, const double* __restrict__ p,
                                                     double alpha, double* __restrict__ parts)
{
    __shared__ double sh[6][RB];
    long i = (long)blockIdx.x * RB + threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (i < n) {
        long v = idx ? idx[i] : i;
        for (int d = 0; d < 3; ++d) {
            double a = X[3 * v + d];
            mn[d] = fmin(mn[d], a);
            mx[d] = fmax(mx[d], a);
            if (p) {
                double b = X[3 * v + d] + alpha * p[3 * v + d];
                mn[d] = fmin(mn[d], b);
                mx[d] = fmax(mx[d], b);
            }
        }
    }
    for (int d = 0; d < 3; ++d) {
        sh[d][threadIdx.x] = mn[d];
        sh[3 + d][threadIdx.x] = mx[d];
    }
    __syncthreads();
    for (int s = RB / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                sh[d][threadIdx.x] = fmin(sh[d][threadIdx.x], sh[d][threadIdx.x + s]);
                sh[3 + d][threadIdx.x] = fmax(sh[3 + d][threadIdx.x], sh[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int d = 0; d < 6; ++d) parts[6 * blockIdx.x + d] = sh[d][0];
}

__global__ void __launch_bounds__(RB) k_minmax_final(int np, const double* __restrict__ parts,
                                                     double* __restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < np; ++b)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], parts[6 * b + d]);
            r[3 + d] = fmax(r[3 + d], parts[6 * b + 3 + d]);
        }
    for (int d = 0; d < 6; ++d) out[d] = r[d];
}

// ---------------------------------------------------------------------------
// boxes and cell records
// ---------------------------------------------------------------------------
// vertex boxes: static -> point voxel; swept -> voxel box of [min(x, x+a p), max(..)] -/+ th/2
__global__ void k_vertex_boxes(int nb, const int* __restrict__ bn, const double* __restrict__ X,
                               const double* __restrict__ p, double alpha, double halfth,
                               const GridDev* __restrict__ gp, int* __restrict__ vbox /* 6 per boundary node */)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const GridDev g = *gp;
    int v = bn[i];
    double a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        double x0 = X[3 * v + d];
        if (p) {
            double x1 = x0 + alpha * p[3 * v + d];
            a[d] = fmin(x0, x1) - halfth;
            b[d] = fmax(x0, x1) + halfth;
        } else {
            a[d] = x0;
            b[d] = x0;
        }
    }
    int lo[3], hi[3];
    locate(g, a[0], a[1], a[2], lo);
    locate(g, b[0], b[1], b[2], hi);
    for (int d = 0; d < 3; ++d) {
        vbox[6 * i + d] = lo[d];
        vbox[6 * i + 3 + d] = hi[d];
    }
}

// primitive boxes: union of vertex boxes; counts voxels (clamped to the grid)
template <int K>
__global__ void k_prim_boxes(int n, const int* __restrict__ prim, const int* __restrict__ v2b,
                             const int* __restrict__ vbox, const GridDev* __restrict__ gp, int* __restrict__ pbox,
                             unsigned* __restrict__ cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GridDev g = *gp;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int k = 0; k < K; ++k) {
        int b = v2b[prim[K * i + k]];
        if (b < 0) continue;  // vertex of a zero-area boundary triangle: not a boundary node
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], vbox[6 * b + d]);
            hi[d] = max(hi[d], vbox[6 * b + 3 + d]);
        }
    }
    unsigned c = 1;
    for (int d = 0; d < 3; ++d) {
        pbox[6 * i + d] = lo[d];
        pbox[6 * i + 3 + d] = hi[d];
        int l = max(lo[d], 0), h = min(hi[d], g.cnt[d] - 1);
        c *= (h >= l) ? (unsigned)(h - l + 1) : 0u;
    }
    cnt[i] = c;
}

__global__ void k_emit_cells(int n, const int* __restrict__ pbox, const unsigned* __restrict__ off,
                             const GridDev* __restrict__ gp, unsigned* __restrict__ keys, int* __restrict__ vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GridDev g = *gp;
    int l[3], h[3];
    for (int d = 0; d < 3; ++d) {
        l[d] = max(pbox[6 * i + d], 0);
        h[d] = min(pbox[6 * i + 3 + d], g.cnt[d] - 1);
    }
    unsigned o = off[i];
    for (int z = l[2]; z <= h[2]; ++z)
        for (int y = l[1]; y <= h[1]; ++y)
            for (int x = l[0]; x <= h[0]; ++x) {
                keys[o] = key_of(g, x, y, z);
                vals[o] = i;
                ++o;
            }
}

__device__ inline unsigned lower_bound_u(const unsigned* a, unsigned n, unsigned key)
{
    unsigned lo = 0, hi = n;
    while (lo < hi) {
        unsigned mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// ---------------------------------------------------------------------------
// narrow phase of the constraint set
// ---------------------------------------------------------------------------
struct CSArgs {
    const double* X;
    const double* X0;
    const unsigned char* dbc;
    const int* bn;
    const int2* be;
    const int3* bt;
    const int* v2b;
    const GridDev* g;
    const unsigned* ckeys;  // sorted cell keys
    const int* cvals;       // primitive per record
    unsigned ncell;
    const int* pbox;        // box of every stored primitive (tri or edge)
    double r, r2;
};

__device__ inline void emit(int4* out, unsigned& n, int a, int b, int c, int d)
{
    if (out) out[n] = make_int4(a, b, c, d);
    ++n;
}

// Classifies one point-triangle pair (IPC.h:158-268).  Returns the number of
// tuples written (0/1) into pass (PT) or dup (PP/PE) streams.
__device__ inline void pt_narrow(const CSArgs& A, int vI, int f, int4* pass, unsigned& np, int4* dup, unsigned& nd)
{
    int3 t = A.bt[f];
    if (vI == t.x || vI == t.y || vI == t.z) return;
    if (A.dbc[vI] && A.dbc[t.x] && A.dbc[t.y] && A.dbc[t.z]) return;
    d3 p = ld3(A.X, vI), t0 = ld3(A.X, t.x), t1 = ld3(A.X, t.y), t2 = ld3(A.X, t.z);
    if (!pt_cd_broadphase(p, t0, t1, t2, A.r)) return;
    double d;
    switch (pt_type(p, t0, t1, t2)) {
    case 0: d = d2_pp(p, t0); if (d < A.r2) emit(dup, nd, -vI - 1, t.x, -1, -1); break;
    case 1: d = d2_pp(p, t1); if (d < A.r2) emit(dup, nd, -vI - 1, t.y, -1, -1); break;
    case 2: d = d2_pp(p, t2); if (d < A.r2) emit(dup, nd, -vI - 1, t.z, -1, -1); break;
    case 3: d = d2_pe(p, t0, t1); if (d < A.r2) emit(dup, nd, -vI - 1, t.x, t.y, -1); break;
    case 4: d = d2_pe(p, t1, t2); if (d < A.r2) emit(dup, nd, -vI - 1, t.y, t.z, -1); break;
    case 5: d = d2_pe(p, t2, t0); if (d < A.r2) emit(dup, nd, -vI - 1, t.z, t.x, -1); break;
    default: d = d2_pt(p, t0, t1, t2); if (d < A.r2) emit(pass, np, -vI - 1, t.x, t.y, t.z); break;
    }
}

// Classifies one edge-edge pair (IPC.h:371-568)
__device__ inline void ee_narrow(const CSArgs& A, int eI, int eJ, int4* pass, unsigned& np, int4* dup,
                                 unsigned& nd)
{
    int2 a = A.be[eI], b = A.be[eJ];
    if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y || eI > eJ) return;
    if (A.dbc[a.x] && A.dbc[a.y] && A.dbc[b.x] && A.dbc[b.y]) return;
    d3 a0 = ld3(A.X, a.x), a1 = ld3(A.X, a.y), b0 = ld3(A.X, b.x), b1 = ld3(A.X, b.y);
    if (!ee_cd_broadphase(a0, a1, b0, b1, A.r)) return;
    const double cr = ee_cross_sqn(a0, a1, b0, b1);
    const double eps = mollifier_eps(ld3(A.X0, a.x), ld3(A.X0, a.y), ld3(A.X0, b.x), ld3(A.X0, b.y));
    const bool m = cr < eps;
    double d;
    switch (ee_type(a0, a1, b0, b1)) {
    case 0:
        d = d2_pp(a0, b0);
        if (d < A.r2) {
            if (m) emit(pass, np, a.x, b.x, -a.y - 1, -b.y - 1);
            else emit(dup, nd, -a.x - 1, b.x, -1, -1);
        }
        break;
    case 1:
        d = d2_pp(a0, b1);
        if (d < A.r2) {
            if (m) emit(pass, np, a.x, b.y, -a.y - 1, -b.x - 1);
            else emit(dup, nd, -a.x - 1, b.y, -1, -1);
        }
        break;
    case 2:
        d = d2_pe(a0, b0, b1);
        if (d < A.r2) {
            if (m) emit(pass, np, a.x, b.x, b.y, -a.y - 1);
            else emit(dup, nd, -a.x - 1, b.x, b.y, -1);
        }
        break;
    case 3:
        d = d2_pp(a1, b0);
        if (d < A.r2) {
            if (m) emit(pass, np, a.y, b.x, -a.x - 1, -b.y - 1);
            else emit(dup, nd, -a.y - 1, b.x, -1, -1);
        }
        break;
    case 4:
        d = d2_pp(a1, b1);
        if (d < A.r2) {
            if (m) emit(pass, np, a.y, b.y, -a.x - 1, -b.x - 1);
            else emit(dup, nd, -a.y - 1, b.y, -1, -1);
        }
        break;
    case 5:
        d = d2_pe(a1, b0, b1);
        if (d < A.r2) {
            if (m) emit(pass, np, a.y, b.x, b.y, -a.x - 1);
            else emit(dup, nd, -a.y - 1, b.x, b.y, -1);
        }
        break;
    case 6:
        d = d2_pe(b0, a0, a1);
        if (d < A.r2) {
            if (m) emit(pass, np, b.x, a.x, a.y, -b.y - 1);
            else emit(dup, nd, -b.x - 1, a.x, a.y, -1);
        }
        break;
    case 7:
        d = d2_pe(b1, a0, a1);
        if (d < A.r2) {
            if (m) emit(pass, np, b.y, a.x, a.y, -b.x - 1);
            else emit(dup, nd, -b.y - 1, a.x, a.y, -1);
        }
        break;
    default:
        d = d2_ee(a0, a1, b0, b1);
        if (d < A.r2) {
            if (m) emit(pass, np, a.x, a.y, -b.x - 1, b.y);
            else emit(pass, np, a.x, a.y, b.x, b.y);
        }
        break;
    }
}

// Walk the voxels of a query box; for every stored primitive in a voxel whose
// box overlaps, visit it once (in the lowest shared voxel).
template <class F>
__device__ inline void walk_box(const CSArgs& A, const int* ql, const int* qh, F&& visit)
{
    const GridDev& g = *A.g;
    int l[3], h[3];
    for (int d = 0; d < 3; ++d) {
        l[d] = max(ql[d], 0);
        h[d] = min(qh[d], g.cnt[d] - 1);
    }
    for (int z = l[2]; z <= h[2]; ++z)
        for (int y = l[1]; y <= h[1]; ++y)
            for (int x = l[0]; x <= h[0]; ++x) {
                unsigned key = key_of(g, x, y, z);
                unsigned s = lower_bound_u(A.ckeys, A.ncell, key);
                for (unsigned k = s; k < A.ncell && A.ckeys[k] == key; ++k) {
                    int f = A.cvals[k];
                    const int* pb = A.pbox + 6 * f;
                    // canonical voxel: componentwise max of the box minima (clamped)
                    int cx = max(max(ql[0], pb[0]), 0), cy = max(max(ql[1], pb[1]), 0), cz = max(max(ql[2], pb[2]), 0);
                    if (cx != x || cy != y || cz != z) continue;
                    visit(f);
                }
            }
}

// PT queries: one thread per boundary vertex.  pass == nullptr: count only.
__global__ void k_cs_pt(int nb, CSArgs A, const unsigned* __restrict__ offP, const unsigned* __restrict__ offD,
                        int4* __restrict__ pass, int4* __restrict__ dup, unsigned* __restrict__ cntP,
                        unsigned* __restrict__ cntD)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    const GridDev& g = *A.g;
    int vI = A.bn[i];
    d3 p = ld3(A.X, vI);
    int ql[3], qh[3];
    // SPATIAL_HASH.h:215-241: query box of [p - r, p + r]
    locate(g, p.x - A.r, p.y - A.r, p.z - A.r, ql);
    locate(g, p.x + A.r, p.y + A.r, p.z + A.r, qh);
    unsigned np = 0, nd = 0;
    int4* po = pass ? pass + offP[i] : nullptr;
    int4* dO = pass ? dup + offD[i] : nullptr;
    walk_box(A, ql, qh, [&](int f) { pt_narrow(A, vI, f, po, np, dO, nd); });
    if (!pass) {
        cntP[i] = np;
        cntD[i] = nd;
    }
}

// EE queries: one thread per boundary edge; stored edges eJ > eI only.
__global__ void k_cs_ee(int ne, CSArgs A, const unsigned* __restrict__ offP, const unsigned* __restrict__ offD,
                        int4* __restrict__ pass, int4* __restrict__ dup, unsigned* __restrict__ cntP,
                        unsigned* __restrict__ cntD)
{
    int eI = blockIdx.x * blockDim.x + threadIdx.x;
    if (eI >= ne) return;
    const GridDev& g = *A.g;
    int2 e = A.be[eI];
    d3 a = ld3(A.X, e.x), b = ld3(A.X, e.y);
    int ql[3], qh[3];
    // SPATIAL_HASH.h:243-291: box of the edge inflated by r
    locate(g, fmin(a.x, b.x) - A.r, fmin(a.y, b.y) - A.r, fmin(a.z, b.z) - A.r, ql);
    locate(g, fmax(a.x, b.x) + A.r, fmax(a.y, b.y) + A.r, fmax(a.z, b.z) + A.r, qh);
    unsigned np = 0, nd = 0;
    int4* po = pass ? pass + offP[eI] : nullptr;
    int4* dO = pass ? dup + offD[eI] : nullptr;
    walk_box(A, ql, qh, [&](int f) {
        if (f > eI) ee_narrow(A, eI, f, po, np, dO, nd);
    });
    if (!pass) {
        cntP[eI] = np;
        cntD[eI] = nd;
    }
}

// ---------------------------------------------------------------------------
// merge of PP/PE duplicates (IPC.h:599-654): key order = std::map order
// ---------------------------------------------------------------------------
__global__ void k_dup_keys(unsigned n, const int4* __restrict__ dup, unsigned long long V,
                           unsigned long long* __restrict__ keys)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int4 c = dup[i];
    unsigned long long s0 = (unsigned long long)(c.x + (long long)V);  // c.x in [-V,-1]
    keys[i] = (s0 * V + (unsigned long long)c.y) * (V + 1) + (unsigned long long)(c.z + 1);
}

__global__ void k_dup_heads(unsigned n, const unsigned long long* __restrict__ keys, unsigned* __restrict__ head)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_dup_write(unsigned n, const unsigned long long* __restrict__ keys,
                            const unsigned* __restrict__ headscan, unsigned long long V, int4* __restrict__ out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(i == 0 || keys[i] != keys[i - 1])) return;
    unsigned j = i + 1;
    while (j < n && keys[j] == keys[i]) ++j;
    unsigned long long k = keys[i];
    int s2 = (int)(k % (V + 1)) - 1;
    k /= (V + 1);
    int s1 = (int)(k % V);
    int s0 = (int)((long long)(k / V) - (long long)V);
    out[headscan[i]] = make_int4(s0, s1, s2, -(int)(j - i));
}

// ---------------------------------------------------------------------------
// CCD candidates and ACCD
// ---------------------------------------------------------------------------
struct CCDArgs {
    const double* X;
    const double* P;
    const unsigned char* dbc;
    const int* bn;
    const int2* be;
    const int3* bt;
    const int* vbox;  // per boundary node
    const int* pbox;  // per stored primitive
    const GridDev* g;
    const unsigned* ckeys;
    const int* cvals;
    unsigned ncell;
    double th;
};

template <class F>
__device__ inline void walk_box_ccd(const CCDArgs& A, const int* ql, const int* qh, F&& visit)
{
    const GridDev& g = *A.g;
    int l[3], h[3];
    for (int d = 0; d < 3; ++d) {
        l[d] = max(ql[d], 0);
        h[d] = min(qh[d], g.cnt[d] - 1);
    }
    for (int z = l[2]; z <= h[2]; ++z)
        for (int y = l[1]; y <= h[1]; ++y)
            for (int x = l[0]; x <= h[0]; ++x) {
                unsigned key = key_of(g, x, y, z);
                unsigned s = lower_bound_u(A.ckeys, A.ncell, key);
                for (unsigned k = s; k < A.ncell && A.ckeys[k] == key; ++k) {
                    int f = A.cvals[k];
                    const int* pb = A.pbox + 6 * f;
                    int cx = max(max(ql[0], pb[0]), 0), cy = max(max(ql[1], pb[1]), 0), cz = max(max(ql[2], pb[2]), 0);
                    if (cx != x || cy != y || cz != z) continue;
                    visit(f);
                }
            }
}

// Candidate pairs that survive the swept AABB test: count (out == nullptr) / write
__global__ void k_ccd_pt_cand(int nb, CCDArgs A, const unsigned* __restrict__ off, int2* __restrict__ out,
                              unsigned* __restrict__ cnt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int v = A.bn[i];
    d3 p = ld3(A.X, v), dp = ld3(A.P, v);
    unsigned n = 0;
    int2* o = out ? out + off[i] : nullptr;
    walk_box_ccd(A, A.vbox + 6 * i, A.vbox + 6 * i + 3, [&](int f) {
        int3 t = A.bt[f];
        if (v == t.x || v == t.y || v == t.z) return;
        if (A.dbc[v] && A.dbc[t.x] && A.dbc[t.y] && A.dbc[t.z]) return;
        if (!pt_ccd_broadphase(p, ld3(A.X, t.x), ld3(A.X, t.y), ld3(A.X, t.z), dp, ld3(A.P, t.x), ld3(A.P, t.y),
                               ld3(A.P, t.z), A.th))
            return;
        if (o) o[n] = make_int2(v, f);
        ++n;
    });
    if (!out) cnt[i] = n;
}

__global__ void k_ccd_ee_cand(int ne, CCDArgs A, const unsigned* __restrict__ off, int2* __restrict__ out,
                              unsigned* __restrict__ cnt)
{
    int eI = blockIdx.x * blockDim.x + threadIdx.x;
    if (eI >= ne) return;
    int2 a = A.be[eI];
    d3 a0 = ld3(A.X, a.x), a1 = ld3(A.X, a.y), da0 = ld3(A.P, a.x), da1 = ld3(A.P, a.y);
    unsigned n = 0;
    int2* o = out ? out + off[eI] : nullptr;
    const int* eb = A.pbox + 6 * eI;
    walk_box_ccd(A, eb, eb + 3, [&](int f) {
        if (f <= eI) return;
        int2 b = A.be[f];
        if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y) return;
        if (A.dbc[a.x] && A.dbc[a.y] && A.dbc[b.x] && A.dbc[b.y]) return;
        if (!ee_ccd_broadphase(a0, a1, ld3(A.X, b.x), ld3(A.X, b.y), da0, da1, ld3(A.P, b.x), ld3(A.P, b.y), A.th))
            return;
        if (o) o[n] = make_int2(eI, f);
        ++n;
    });
    if (!out) cnt[eI] = n;
}

// ACCD over candidate pairs; toc bound = running minimum (result independent
// of evaluation order: the minimum over pairs of their time of impact).
__global__ void k_accd(unsigned npt, unsigned nee, const int2* __restrict__ cand, CCDArgs A, double eta,
                       unsigned long long* __restrict__ amin)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= npt + nee) return;
    int2 c = cand[i];
    double toc = __longlong_as_double((long long)*((volatile unsigned long long*)amin));
    bool hit;
    if (i < npt) {
        int3 t = A.bt[c.y];
        hit = accd_pt(ld3(A.X, c.x), ld3(A.X, t.x), ld3(A.X, t.y), ld3(A.X, t.z), ld3(A.P, c.x), ld3(A.P, t.x),
                      ld3(A.P, t.y), ld3(A.P, t.z), eta, A.th, toc);
    } else {
        int2 a = A.be[c.x], b = A.be[c.y];
        hit = accd_ee(ld3(A.X, a.x), ld3(A.X, a.y), ld3(A.X, b.x), ld3(A.X, b.y), ld3(A.P, a.x), ld3(A.P, a.y),
                      ld3(A.P, b.x), ld3(A.P, b.y), eta, A.th, toc);
    }
    if (hit) atomicMin(amin, (unsigned long long)__double_as_longlong(toc));
}

// span cap (SPATIAL_HASH.h:466-482), one thread
__global__ void k_ccd_cap(const double* __restrict__ eparts, int nep, long ne, const double* __restrict__ pparts,
                          int npp, long nb, double alpha_in, double* __restrict__ alpha_out)
{
    double vs = 0.0, ps = 0.0;
    for (int c = 0; c < nep; ++c) vs += eparts[c];
    vs = ne ? vs / (double)ne : 1.0;
    for (int c = 0; c < npp; ++c) ps += pparts[c];
    ps /= (double)(nb * 3);
    double a = alpha_in;
    const double span = a * ps / vs;
    if (span > 1) a /= span;
    *alpha_out = a;
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
namespace {

template <class T>
void excl_scan(hipStream_t s, DBuf<unsigned char>& tmp, const T* in, T* out, int n)
{
    size_t bytes = 0;
    IDP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, s));
}

template <class K, class V>
void sort_pairs(hipStream_t s, DBuf<unsigned char>& tmp, const K* kin, K* kout, const V* vin, V* vout, long n,
                int end_bit)
{
    size_t bytes = 0;
    IDP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
}

template <class K>
void sort_keys(hipStream_t s, DBuf<unsigned char>& tmp, const K* kin, K* kout, long n, int end_bit)
{
    size_t bytes = 0;
    IDP_HIP(hipcub::DeviceRadixSort::SortKeys(nullptr, bytes, kin, kout, (int)n, 0, end_bit, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(hipcub::DeviceRadixSort::SortKeys(tmp.p, bytes, kin, kout, (int)n, 0, end_bit, s));
}

unsigned read_u(hipStream_t s, const unsigned* p)
{
    unsigned v;
    IDP_HIP(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    return v;
}

int bits_for(unsigned long long maxv)
{
    int b = 1;
    while (b < 64 && (maxv >> b)) ++b;
    return b;
}

// Sorted (voxel key, primitive) records of the boxes in pbox (count n).
unsigned build_cells(ContactWork& w, hipStream_t s, int n, const GridDev* g)
{
    if (n == 0) return 0;
    unsigned* cnt = w.u0.ensure(n + 1, s);
    unsigned* off = w.u1.ensure(n + 1, s);
    // counts already written by k_prim_boxes into u0; total = scan
    IDP_HIP(hipMemsetAsync(cnt + n, 0, sizeof(unsigned), s));
    excl_scan(s, w.tmp, cnt, off, n + 1);
    unsigned total = read_u(s, off + n);
    if (total == 0) return 0;
    unsigned* k0 = w.k0.ensure(total, s);
    int* v0 = w.v0.ensure(total, s);
    IDP_LAUNCH(k_emit_cells, grid_for(n, 256), 256, 0, s, n, w.pbox.p, off, g, k0, v0);
    unsigned* k1 = w.k1.ensure(total, s);
    int* v1 = w.v1.ensure(total, s);
    GridDev gh;
    IDP_HIP(hipMemcpyAsync(&gh, g, sizeof(gh), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    unsigned long long maxkey = (unsigned long long)gh.cnt[0] * gh.cnt[1] * gh.cnt[2];
    sort_pairs(s, w.tmp, k0, k1, v0, v1, total, bits_for(maxkey));
    return total;
}

void grid_from(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const int* idx, long nidx,
               const double* p, double alpha, double pad)
{
    // mean boundary edge length (chunked sum) and min/max over the given vertices
    int nep = grid_for(m.ne, 256);
    double* ep = w.dparts.ensure(nep + 1, s);
    IDP_LAUNCH(k_edge_len_parts, grid_for(nep, 64), 64, 0, s, m.ne, (const int2*)m.be.p, X, ep);
    int nmm = grid_for(nidx, RB);
    double* mm = w.mmparts.ensure(6 * (size_t)nmm + 6, s);
    IDP_LAUNCH(k_minmax_parts, nmm, RB, 0, s, nidx, idx, X, p, alpha, mm);
    IDP_LAUNCH(k_minmax_final, 1, RB, 0, s, nmm, mm, mm + 6 * (size_t)nmm);
    GridDev* g = w.grid.ensure(1, s);
    IDP_LAUNCH(k_grid_setup, 1, 1, 0, s, g, ep, nep, (long)m.ne, mm + 6 * (size_t)nmm, pad);
}

}  // namespace

// Constraint set at X (device); result in w.cs (count returned)
long contact_constraint_set(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, double dHat2,
                            double thickness)
{
    const double r = sqrt(dHat2) + thickness;  // IPC.h:51-52
    const double r2 = r * r;
    if (m.nb == 0) return 0;
    grid_from(w, s, m, X, nullptr, m.nV, nullptr, 0.0, 0.0);
    const GridDev* g = w.grid.p;
    int* vbox = w.vbox.ensure(6 * (size_t)m.nb, s);
    IDP_LAUNCH(k_vertex_boxes, grid_for(m.nb, 256), 256, 0, s, m.nb, m.bn.p, X, (const double*)nullptr, 0.0, 0.0,
               g, vbox);

    CSArgs A;
    A.X = X;
    A.X0 = m.X0.p;
    A.dbc = m.dbc.p;
    A.bn = m.bn.p;
    A.be = (const int2*)m.be.p;
    A.bt = (const int3*)m.bt.p;
    A.v2b = m.v2b.p;
    A.g = g;
    A.r = r;
    A.r2 = r2;

    // ---- PT: triangles in cells, query per boundary vertex
    unsigned nPTp = 0, nPTd = 0, nEEp = 0, nEEd = 0;
    DBuf<int4>& pass = w.pass;
    DBuf<int4>& dup = w.dup;
    unsigned* cntP = w.cP.ensure(m.nb + m.ne + 2, s);
    unsigned* cntD = w.cD.ensure(m.nb + m.ne + 2, s);
    unsigned* offP = w.oP.ensure(m.nb + m.ne + 2, s);
    unsigned* offD = w.oD.ensure(m.nb + m.ne + 2, s);
    {
        w.pbox.ensure(6 * (size_t)std::max(m.nf, 1), s);
        w.u0.ensure(m.nf + 1, s);
        IDP_LAUNCH(k_prim_boxes<3>, grid_for(m.nf, 256), 256, 0, s, m.nf, m.bt.p, m.v2b.p, vbox, g, w.pbox.p, w.u0.p);
        unsigned ncell = build_cells(w, s, m.nf, g);
        A.ckeys = w.k1.p;
        A.cvals = w.v1.p;
        A.ncell = ncell;
        A.pbox = w.pbox.p;
        IDP_LAUNCH(k_cs_pt, grid_for(m.nb, 128), 128, 0, s, m.nb, A, (const unsigned*)nullptr,
                   (const unsigned*)nullptr, (int4*)nullptr, (int4*)nullptr, cntP, cntD);
        IDP_HIP(hipMemsetAsync(cntP + m.nb, 0, sizeof(unsigned), s));
        IDP_HIP(hipMemsetAsync(cntD + m.nb, 0, sizeof(unsigned), s));
        excl_scan(s, w.tmp, cntP, offP, m.nb + 1);
        excl_scan(s, w.tmp, cntD, offD, m.nb + 1);
        nPTp = read_u(s, offP + m.nb);
        nPTd = read_u(s, offD + m.nb);
    }
    // output layout: [PT pass | EE pass] and [PT dup | EE dup]
    // (PT results are written after EE counts are known)
    unsigned* cntP2 = cntP + m.nb + 1;
    unsigned* cntD2 = cntD + m.nb + 1;
    unsigned* offP2 = offP + m.nb + 1;
    unsigned* offD2 = offD + m.nb + 1;
    CSArgs B = A;
    {
        // write PT results now (buffers sized after the count)
        pass.ensure(nPTp + 1, s);
        dup.ensure(nPTd + 1, s);
        IDP_LAUNCH(k_cs_pt, grid_for(m.nb, 128), 128, 0, s, m.nb, A, offP, offD, pass.p, dup.p, cntP, cntD);
        // ---- EE: edges in cells, query per edge
        w.pbox.ensure(6 * (size_t)std::max(m.ne, 1), s);
        w.u0.ensure(m.ne + 1, s);
        IDP_LAUNCH(k_prim_boxes<2>, grid_for(m.ne, 256), 256, 0, s, m.ne, m.be.p, m.v2b.p, vbox, g, w.pbox.p,
                   w.u0.p);
        unsigned ncell = build_cells(w, s, m.ne, g);
        B.ckeys = w.k1.p;
        B.cvals = w.v1.p;
        B.ncell = ncell;
        B.pbox = w.pbox.p;
        IDP_LAUNCH(k_cs_ee, grid_for(m.ne, 128), 128, 0, s, m.ne, B, (const unsigned*)nullptr,
                   (const unsigned*)nullptr, (int4*)nullptr, (int4*)nullptr, cntP2, cntD2);
        IDP_HIP(hipMemsetAsync(cntP2 + m.ne, 0, sizeof(unsigned), s));
        IDP_HIP(hipMemsetAsync(cntD2 + m.ne, 0, sizeof(unsigned), s));
        excl_scan(s, w.tmp, cntP2, offP2, m.ne + 1);
        excl_scan(s, w.tmp, cntD2, offD2, m.ne + 1);
        nEEp = read_u(s, offP2 + m.ne);
        nEEd = read_u(s, offD2 + m.ne);
    }
    // the PT pass tuples must stay where they are; grow with keep
    pass.ensure(nPTp + nEEp + 1, s, true);
    dup.ensure(nPTd + nEEd + 1, s, true);
    IDP_LAUNCH(k_cs_ee, grid_for(m.ne, 128), 128, 0, s, m.ne, B, offP2, offD2, pass.p + nPTp, dup.p + nPTd, cntP2,
               cntD2);

    // ---- merge duplicates
    unsigned nd = nPTd + nEEd;
    unsigned nmerged = 0;
    unsigned long long V = (unsigned long long)m.nV;
    if (nd) {
        unsigned long long* k0 = w.dk0.ensure(nd, s);
        unsigned long long* k1 = w.dk1.ensure(nd, s);
        IDP_LAUNCH(k_dup_keys, grid_for(nd, 256), 256, 0, s, nd, dup.p, V, k0);
        sort_keys(s, w.tmp, k0, k1, nd, bits_for(V * V * (V + 1)));
        unsigned* head = w.u0.ensure(nd + 1, s);
        unsigned* hs = w.u1.ensure(nd + 1, s);
        IDP_LAUNCH(k_dup_heads, grid_for(nd, 256), 256, 0, s, nd, k1, head);
        IDP_HIP(hipMemsetAsync(head + nd, 0, sizeof(unsigned), s));
        excl_scan(s, w.tmp, head, hs, nd + 1);
        nmerged = read_u(s, hs + nd);
        w.merged.ensure(nmerged + 1, s);
        IDP_LAUNCH(k_dup_write, grid_for(nd, 256), 256, 0, s, nd, k1, hs, V, w.merged.p);
    }
    long total = (long)nPTp + nEEp + nmerged;
    int4* cs = w.cs.ensure(total + 1, s);
    if (nPTp + nEEp)
        IDP_HIP(hipMemcpyAsync(cs, pass.p, sizeof(int4) * (nPTp + nEEp), hipMemcpyDeviceToDevice, s));
    if (nmerged)
        IDP_HIP(hipMemcpyAsync(cs + nPTp + nEEp, w.merged.p, sizeof(int4) * nmerged, hipMemcpyDeviceToDevice, s));
    w.ncs = total;
    return total;
}

// ACCD step size along P from X; returns alpha (host)
double contact_ccd(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                   double thickness, double alpha_in, long* ncand_out)
{
    if (m.nb == 0) return alpha_in;
    // span cap (needs mean edge length and mean |p|)
    int nep = grid_for(m.ne, 256);
    int npp = grid_for(3L * m.nb, 256);
    double* ep = w.dparts.ensure(nep + npp + 2, s);
    double* pp = ep + nep + 1;
    IDP_LAUNCH(k_edge_len_parts, grid_for(nep, 64), 64, 0, s, m.ne, (const int2*)m.be.p, X, ep);
    IDP_LAUNCH(k_psize_parts, grid_for(npp, 64), 64, 0, s, m.nb, m.bn.p, P, pp);
    double* dalpha = w.scal.ensure(4, s);
    IDP_LAUNCH(k_ccd_cap, 1, 1, 0, s, ep, nep, (long)m.ne, pp, npp, (long)m.nb, alpha_in, dalpha);
    double alpha;
    IDP_HIP(hipMemcpyAsync(&alpha, dalpha, sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));

    // swept grid over boundary nodes at [x, x + alpha p] -/+ th/2
    // (SPATIAL_HASH.h:484-511); mean edge length recomputed inside grid_from
    int nmm = grid_for(m.nb, RB);
    double* mm = w.mmparts.ensure(6 * (size_t)nmm + 6, s);
    IDP_LAUNCH(k_minmax_parts, nmm, RB, 0, s, (long)m.nb, m.bn.p, X, P, alpha, mm);
    IDP_LAUNCH(k_minmax_final, 1, RB, 0, s, nmm, mm, mm + 6 * (size_t)nmm);
    GridDev* g = w.grid.ensure(1, s);
    double* ep2 = w.dparts2.ensure(nep + 1, s);
    IDP_LAUNCH(k_edge_len_parts, grid_for(nep, 64), 64, 0, s, m.ne, (const int2*)m.be.p, X, ep2);
    IDP_LAUNCH(k_grid_setup, 1, 1, 0, s, g, ep2, nep, (long)m.ne, mm + 6 * (size_t)nmm, thickness / 2);
    int* vbox = w.vbox.ensure(6 * (size_t)m.nb, s);
    IDP_LAUNCH(k_vertex_boxes, grid_for(m.nb, 256), 256, 0, s, m.nb, m.bn.p, X, P, alpha, thickness / 2, g, vbox);

    CCDArgs A;
    A.X = X;
    A.P = P;
    A.dbc = m.dbc.p;
    A.bn = m.bn.p;
    A.be = (const int2*)m.be.p;
    A.bt = (const int3*)m.bt.p;
    A.vbox = vbox;
    A.g = g;
    A.th = thickness;

    unsigned* cnt = w.cP.ensure(std::max(m.nb, m.ne) + 2, s);
    unsigned* off = w.oP.ensure(std::max(m.nb, m.ne) + 2, s);
    // PT candidates
    w.pbox.ensure(6 * (size_t)std::max(m.nf, 1), s);
    w.u0.ensure(m.nf + 1, s);
    IDP_LAUNCH(k_prim_boxes<3>, grid_for(m.nf, 256), 256, 0, s, m.nf, m.bt.p, m.v2b.p, vbox, g, w.pbox.p, w.u0.p);
    unsigned ncell = build_cells(w, s, m.nf, g);
    A.ckeys = w.k1.p;
    A.cvals = w.v1.p;
    A.ncell = ncell;
    A.pbox = w.pbox.p;
    IDP_LAUNCH(k_ccd_pt_cand, grid_for(m.nb, 128), 128, 0, s, m.nb, A, (const unsigned*)nullptr, (int2*)nullptr, cnt);
    IDP_HIP(hipMemsetAsync(cnt + m.nb, 0, sizeof(unsigned), s));
    excl_scan(s, w.tmp, cnt, off, m.nb + 1);
    unsigned npt = read_u(s, off + m.nb);
    int2* cand = w.cand.ensure(npt + 1, s);
    IDP_LAUNCH(k_ccd_pt_cand, grid_for(m.nb, 128), 128, 0, s, m.nb, A, off, cand, cnt);
    // EE candidates
    w.pbox.ensure(6 * (size_t)std::max(m.ne, 1), s);
    w.u0.ensure(m.ne + 1, s);
    IDP_LAUNCH(k_prim_boxes<2>, grid_for(m.ne, 256), 256, 0, s, m.ne, m.be.p, m.v2b.p, vbox, g, w.pbox.p, w.u0.p);
    ncell = build_cells(w, s, m.ne, g);
    A.ckeys = w.k1.p;
    A.cvals = w.v1.p;
    A.ncell = ncell;
    A.pbox = w.pbox.p;
    IDP_LAUNCH(k_ccd_ee_cand, grid_for(m.ne, 128), 128, 0, s, m.ne, A, (const unsigned*)nullptr, (int2*)nullptr, cnt);
    IDP_HIP(hipMemsetAsync(cnt + m.ne, 0, sizeof(unsigned), s));
    excl_scan(s, w.tmp, cnt, off, m.ne + 1);
    unsigned nee = read_u(s, off + m.ne);
    cand = w.cand.ensure(npt + nee + 1, s, true);
    IDP_LAUNCH(k_ccd_ee_cand, grid_for(m.ne, 128), 128, 0, s, m.ne, A, off, cand + npt, cnt);
    // ACCD
    unsigned long long* amin = (unsigned long long*)w.scal.p + 1;
    unsigned long long abits;
    std::memcpy(&abits, &alpha, sizeof(abits));
    IDP_HIP(hipMemcpyAsync(amin, &abits, sizeof(abits), hipMemcpyHostToDevice, s));
    const int tok = w.prof ? w.prof->begin(IDP_K_CCD) : -1;
    IDP_LAUNCH(k_accd, grid_for(npt + nee, 64), 64, 0, s, npt, nee, cand, A, 0.1, amin);
    if (w.prof) w.prof->end(IDP_K_CCD, tok, (double)npt + nee);
    IDP_HIP(hipMemcpyAsync(&abits, amin, sizeof(abits), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    std::memcpy(&alpha, &abits, sizeof(alpha));
    if (ncand_out) *ncand_out = (long)npt + nee;
    return alpha;
}

}  // namespace idp
