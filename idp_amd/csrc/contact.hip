// Contact detection on the device:
//   constraint set  Library/FEM/IPC.h:19-740 with the static spatial hash of
//                   Library/Grid/SPATIAL_HASH.h:28-291
//   ACCD step size  Library/FEM/IPC.h:1879-2244 with the swept hash of
//                   Library/Grid/SPATIAL_HASH.h:432-690
//
// Broad phase.  The reference hashes primitives into a uniform voxel grid
// (voxel = mean boundary edge length) and pairs primitives that share a
// voxel.  Here every triangle and edge gets an integer voxel box; boxes are
// expanded into (2 * voxel + type, primitive) records and radix-sorted, and a
// dense [start, end) table per record key (cleared after each use) gives the
// primitives of a voxel in O(1).  Each query (boundary vertex or edge) is
// handled by a group of lanes that split the voxels of its box.  A pair found
// in several shared voxels is kept only in the lowest shared voxel
// (componentwise max of the two box minima), so no hash set is needed.  The
// narrow phase applies the reference's FP64 filters in the reference's order
// (incidence, all-DBC, AABB test, type, distance), so the emitted set is the
// reference's.  Outputs are appended with atomic counters and then sorted on
// integer keys, so the constraint order is deterministic; CCD candidates need
// no order (the step is a minimum).
#include "common.hpp"
#include "prim.hpp"
#include "geom_dev.hpp"
#include "kernels.hpp"


namespace idp {

// ---------------------------------------------------------------------------
// grid set-up (SPATIAL_HASH.h:60-84 and 455-511), one thread per instance:
// voxel = mean boundary edge length of the instance (its chunk partials added
// in order), origin and extent from the instance's bounding box
// ---------------------------------------------------------------------------
__global__ void k_grid_setup(int B, GridDev* __restrict__ grids, const double* __restrict__ mean_parts,
                             const int* __restrict__ mean_pbeg, const int* __restrict__ ebeg,
                             const double* __restrict__ mm /* 6 per instance: min xyz, max xyz */, double pad,
                             const unsigned char* __restrict__ active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    GridDev* g = grids + i;
    if (active && !active[i]) {  // an instance left out: a one-voxel grid (no records, no key space)
        g->lbc[0] = g->lbc[1] = g->lbc[2] = 0.0;
        g->inv = g->voxel = 1.0;
        g->cnt[0] = g->cnt[1] = g->cnt[2] = 1;
        return;
    }
    double tot = 0.0;
    for (int c = mean_pbeg[i]; c < mean_pbeg[i + 1]; ++c) tot += mean_parts[c];
    const long n_mean = ebeg[i + 1] - ebeg[i];
    double vs = n_mean ? tot / (double)n_mean : 1.0;
    g->voxel = vs;
    const double* m6 = mm + 6 * i;
    double lo[3] = {m6[0] - pad, m6[1] - pad, m6[2] - pad};
    double hi[3] = {m6[3] + pad, m6[4] + pad, m6[5] + pad};
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

// Sum of a chunk of 256 consecutive values, added sequentially in index
// order (the order of the CPU restatement): the block computes the values in
// parallel into LDS, one lane adds them up.
__device__ inline double chunk_sum_lds(double v, double* sh, int nvalid)
{
    sh[threadIdx.x] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int i = 0; i < nvalid; ++i) s += sh[i];
    return s;
}

// chunked partial sums of boundary edge lengths (one block per chunk of 256
// edges of one instance)
__global__ void __launch_bounds__(256) k_edge_len_parts(const int* __restrict__ lo, const int* __restrict__ hi,
                                                        const int2* __restrict__ E, const double* __restrict__ X,
                                                        double* __restrict__ parts)
{
    __shared__ double sh[256];
    const int b = lo[blockIdx.x], e = hi[blockIdx.x], i = b + threadIdx.x;
    double len = 0.0;
    if (i < e) {
        int2 ed = E[i];
        len = sqrt(sqn(ld3(X, ed.x) - ld3(X, ed.y)));
    }
    const double s = chunk_sum_lds(len, sh, e - b);
    if (threadIdx.x == 0) parts[blockIdx.x] = s;
}

// sum of |p| over an instance's boundary nodes, x, y, z of each node in node
// order, one add after another: the order of the reference's loop
// (SPATIAL_HASH.h:466-474), so the span cap rounds as it does there.  One
// thread per instance; the loads run ahead of the dependent adds.
__global__ void k_psize_seq(int B, const int* __restrict__ nbeg, const int* __restrict__ bn,
                            const double* __restrict__ p, double* __restrict__ out,
                            const unsigned char* __restrict__ active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B || (active && !active[i])) return;
    double ps = 0.0;
#pragma unroll 8
    for (int k = nbeg[i]; k < nbeg[i + 1]; ++k) {
        const int v = bn[k];
        ps += fabs(p[3 * v]);
        ps += fabs(p[3 * v + 1]);
        ps += fabs(p[3 * v + 2]);
    }
    out[i] = ps;
}

// min / max of positions (and of x + alpha_i p when p is given) over the
// elements [lo, hi) of each block (vertex ids, or indices into idx):
// per-block partials, exact (order independent)
__global__ void __launch_bounds__(RB) k_minmax_parts(const int* __restrict__ lo, const int* __restrict__ hi,
                                                     const int* __restrict__ binst, const int* __restrict__ idx,
                                                     const double* __restrict__ X, const double* __restrict__ p,
                                                     const double* __restrict__ alpha_p, double* __restrict__ parts)
{
    __shared__ double sh[6][RB];
    const int e = lo[blockIdx.x] + threadIdx.x;
    double mn[3] = {INFINITY, INFINITY, INFINITY}, mx[3] = {-INFINITY, -INFINITY, -INFINITY};
    if (e < hi[blockIdx.x]) {
        long v = idx ? idx[e] : e;
        const double alpha = p ? alpha_p[binst[blockIdx.x]] : 0.0;
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

// per instance: min / max over its block partials
__global__ void k_minmax_final(int B, const int* __restrict__ pbeg, const double* __restrict__ parts,
                               double* __restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = pbeg[i]; b < pbeg[i + 1]; ++b)
        for (int d = 0; d < 3; ++d) {
            r[d] = fmin(r[d], parts[6 * b + d]);
            r[3 + d] = fmax(r[3 + d], parts[6 * b + 3 + d]);
        }
    for (int d = 0; d < 6; ++d) out[6 * i + d] = r[d];
}

// ---------------------------------------------------------------------------
// boxes and cell records
// ---------------------------------------------------------------------------
// vertex boxes in the grid of the vertex's instance: static -> point voxel;
// swept -> voxel box of [min(x, x + a p), max(..)] -/+ th/2, a = alpha of the
// instance (device memory)
__global__ void k_vertex_boxes(int nb, const int* __restrict__ bn, const int* __restrict__ vinst,
                               const double* __restrict__ X, const double* __restrict__ p,
                               const double* __restrict__ alpha_p, double halfth, const GridDev* __restrict__ grids,
                               int* __restrict__ vbox /* 6 per boundary node */)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nb) return;
    int v = bn[i];
    const int inst = vinst[v];
    const GridDev g = grids[inst];
    double a[3], b[3];
    for (int d = 0; d < 3; ++d) {
        double x0 = X[3 * v + d];
        if (p) {
            double x1 = x0 + alpha_p[inst] * p[3 * v + d];
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

// primitive boxes: union of vertex boxes; counts voxels (clamped to the grid
// of the primitive's instance); primitives of inactive instances get none
template <int K>
__global__ void k_prim_boxes(int n, const int* __restrict__ prim, const int* __restrict__ v2b,
                             const int* __restrict__ vinst, const unsigned char* __restrict__ active,
                             const int* __restrict__ vbox, const GridDev* __restrict__ grids,
                             int* __restrict__ pbox, unsigned* __restrict__ cnt, unsigned* __restrict__ pinst)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = vinst[prim[K * i]];
    pinst[i] = (unsigned)inst;
    int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
    for (int k = 0; k < K; ++k) {
        int b = v2b[prim[K * i + k]];
        if (b < 0) continue;  // vertex of a zero-area boundary triangle: not a boundary node
        for (int d = 0; d < 3; ++d) {
            lo[d] = min(lo[d], vbox[6 * b + d]);
            hi[d] = max(hi[d], vbox[6 * b + 3 + d]);
        }
    }
    if (active && !active[inst]) {
        for (int d = 0; d < 3; ++d) {
            pbox[6 * i + d] = lo[d];
            pbox[6 * i + 3 + d] = hi[d];
        }
        cnt[i] = 0;
        return;
    }
    const GridDev g = grids[inst];
    unsigned c = 1;
    for (int d = 0; d < 3; ++d) {
        pbox[6 * i + d] = lo[d];
        pbox[6 * i + 3 + d] = hi[d];
        int l = max(lo[d], 0), h = min(hi[d], g.cnt[d] - 1);
        c *= (h >= l) ? (unsigned)(h - l + 1) : 0u;
    }
    cnt[i] = c;
}

// (keybase[inst] + 2 * voxel + type, primitive) records; primitives [0, nsplit)
// are type 0 (triangles), [nsplit, n) type 1 (edges, numbered from 0)
__global__ void k_emit_cells(int n, int nsplit, const int* __restrict__ pbox, const unsigned* __restrict__ off,
                             const unsigned* __restrict__ pinst, const unsigned* __restrict__ keybase,
                             const GridDev* __restrict__ grids, unsigned* __restrict__ keys, int* __restrict__ vals)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned o0 = off[i], o1 = off[i + 1];
    if (o1 == o0) return;
    const unsigned inst = pinst[i];
    const GridDev g = grids[inst];
    const unsigned base = keybase[inst];
    const unsigned type = i >= nsplit ? 1u : 0u;
    const int id = i - (type ? nsplit : 0);
    int l[3], h[3];
    for (int d = 0; d < 3; ++d) {
        l[d] = max(pbox[6 * i + d], 0);
        h[d] = min(pbox[6 * i + 3 + d], g.cnt[d] - 1);
    }
    unsigned o = o0;
    for (int z = l[2]; z <= h[2]; ++z)
        for (int y = l[1]; y <= h[1]; ++y)
            for (int x = l[0]; x <= h[0]; ++x) {
                keys[o] = base + 2u * key_of(g, x, y, z) + type;
                vals[o] = id;
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

// Sorted records plus, when the grid is small enough, a dense [start, end)
// table per record key (all zero between calls: entries are cleared after use).
struct CellTable {
    const unsigned* keys;
    const int* vals;
    const unsigned* start;  // null: binary search in keys
    const unsigned* end;
    unsigned n;
};

__device__ inline void cell_range(const CellTable& T, unsigned key, unsigned& s, unsigned& e)
{
    if (T.start) {
        s = T.start[key];
        e = T.end[key];
    } else {
        s = lower_bound_u(T.keys, T.n, key);
        e = lower_bound_u(T.keys, T.n, key + 1);
    }
}

// sets (clear = 0) or resets (clear = 1) the table entries of the keys present
__global__ void k_cell_bounds(unsigned n, const unsigned* __restrict__ keys, unsigned* __restrict__ start,
                              unsigned* __restrict__ end, int clear)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned k = keys[i];
    if (i == 0 || keys[i - 1] != k) start[k] = clear ? 0u : i;
    if (i == n - 1 || keys[i + 1] != k) end[k] = clear ? 0u : i + 1;
}

__global__ void k_fill_u32(size_t n, unsigned* __restrict__ a, unsigned v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// Voxels of a query box split over the G lanes of a group (lane j takes the
// linear voxel indices j, j+G, ...).  Every stored primitive of the requested
// type whose box overlaps the query box is visited exactly once, in the
// lowest shared voxel (componentwise max of the two box minima, clamped).
template <int G, class F>
__device__ inline void walk_lanes(const GridDev& g, unsigned keybase, const CellTable& T,
                                  const int* __restrict__ pbox_t, int type, const int* ql, const int* qh, int lane,
                                  F&& visit)
{
    int l[3], n[3];
    for (int d = 0; d < 3; ++d) {
        l[d] = max(ql[d], 0);
        n[d] = min(qh[d], g.cnt[d] - 1) - l[d] + 1;
        if (n[d] <= 0) return;
    }
    const long nvox = (long)n[0] * n[1] * n[2];
    for (long li = lane; li < nvox; li += G) {
        const int x = l[0] + (int)(li % n[0]);
        const long r = li / n[0];
        const int y = l[1] + (int)(r % n[1]);
        const int z = l[2] + (int)(r / n[1]);
        unsigned s, e;
        cell_range(T, keybase + 2u * key_of(g, x, y, z) + (unsigned)type, s, e);
        for (unsigned k = s; k < e; ++k) {
            const int f = T.vals[k];
            const int* pb = pbox_t + 6L * f;
            if (max(max(ql[0], pb[0]), 0) != x || max(max(ql[1], pb[1]), 0) != y || max(max(ql[2], pb[2]), 0) != z)
                continue;
            visit(f);
        }
    }
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
    const int* pbox;  // triangles [0, nf), then edges
    const GridDev* grids;
    const unsigned* keybase;
    const int* vinst;
    const unsigned char* active;  // null: every instance
    const double* rad;            // per instance: r = sqrt(dHat2) + thickness, then r^2 (2 B values)
    int B, nf;
};

// Outputs are appended with atomic counters and put in a fixed order
// afterwards by sorting on their keys, built from instance-local indices:
//   pass:  instance << 40 | part << 39 | local, part 0 = PT (vertex, triangle),
//          part 1 = EE (edge, edge)
//   dup:   [instance << 56 |] (s0, s1, s2) merge key in std::map order
// so every instance's constraints come out in the order of a solo solve.
struct CSOut {
    unsigned* cnt;  // [0] pass, [1] dup
    unsigned long long* pkey;
    int4* pval;
    unsigned long long* dkey;
    unsigned cap_p, cap_d;
    const int *vbeg, *ebeg, *fbeg;  // instance ranges (B + 1)
    int inst_shift;                 // 56 with several instances, 0 with one
};

__device__ inline unsigned long long pt_pass_key(const CSOut& O, int inst, int vI, int f)
{
    const unsigned long long nf = O.fbeg[inst + 1] - O.fbeg[inst];
    return ((unsigned long long)inst << 40) |
           ((unsigned long long)(vI - O.vbeg[inst]) * nf + (unsigned long long)(f - O.fbeg[inst]));
}

__device__ inline unsigned long long ee_pass_key(const CSOut& O, int inst, int eI, int eJ)
{
    const unsigned long long ne = O.ebeg[inst + 1] - O.ebeg[inst];
    return ((unsigned long long)inst << 40) | (1ull << 39) |
           ((unsigned long long)(eI - O.ebeg[inst]) * ne + (unsigned long long)(eJ - O.ebeg[inst]));
}

__device__ inline void out_pass(const CSOut& O, unsigned long long key, int a, int b, int c, int d)
{
    const unsigned i = atomicAdd(O.cnt, 1u);
    if (i < O.cap_p) {
        O.pkey[i] = key;
        O.pval[i] = make_int4(a, b, c, d);
    }
}

// PP / PE duplicate (-p-1, q, e1 or -1, -1): key of the (s0, s1, s2) map entry
// in std::map order, over instance-local vertex ids
__device__ inline void out_dup(const CSOut& O, int inst, int a, int b, int c)
{
    const unsigned i = atomicAdd(O.cnt + 1, 1u);
    if (i < O.cap_d) {
        const long long vb = O.vbeg[inst];
        const unsigned long long V = (unsigned long long)(O.vbeg[inst + 1] - vb);
        const unsigned long long s0 = (unsigned long long)((a + vb) + (long long)V);  // local -p-1 in [-V,-1]
        const unsigned long long s1 = (unsigned long long)(b - vb);
        const unsigned long long s2 = c < 0 ? 0ull : (unsigned long long)(c - vb + 1);
        const unsigned long long key = (s0 * V + s1) * (V + 1) + s2;
        O.dkey[i] = O.inst_shift ? (((unsigned long long)inst << O.inst_shift) | key) : key;
    }
}

// Classifies one point-triangle pair (IPC.h:158-268).  Returns the number of
// tuples written (0/1) into pass (PT) or dup (PP/PE) streams.
__device__ inline void pt_narrow(const CSArgs& A, int inst, int vI, int f, const CSOut& O)
{
    const double r = A.rad[inst], r2 = A.rad[A.B + inst];
    int3 t = A.bt[f];
    if (vI == t.x || vI == t.y || vI == t.z) return;
    if (A.dbc[vI] && A.dbc[t.x] && A.dbc[t.y] && A.dbc[t.z]) return;
    d3 p = ld3(A.X, vI), t0 = ld3(A.X, t.x), t1 = ld3(A.X, t.y), t2 = ld3(A.X, t.z);
    if (!pt_cd_broadphase(p, t0, t1, t2, r)) return;
    double d;
    switch (pt_type(p, t0, t1, t2)) {
    case 0: d = d2_pp(p, t0); if (d < r2) out_dup(O, inst, -vI - 1, t.x, -1); break;
    case 1: d = d2_pp(p, t1); if (d < r2) out_dup(O, inst, -vI - 1, t.y, -1); break;
    case 2: d = d2_pp(p, t2); if (d < r2) out_dup(O, inst, -vI - 1, t.z, -1); break;
    case 3: d = d2_pe(p, t0, t1); if (d < r2) out_dup(O, inst, -vI - 1, t.x, t.y); break;
    case 4: d = d2_pe(p, t1, t2); if (d < r2) out_dup(O, inst, -vI - 1, t.y, t.z); break;
    case 5: d = d2_pe(p, t2, t0); if (d < r2) out_dup(O, inst, -vI - 1, t.z, t.x); break;
    default: d = d2_pt(p, t0, t1, t2); if (d < r2) out_pass(O, pt_pass_key(O, inst, vI, f), -vI - 1, t.x, t.y, t.z); break;
    }
}

// Classifies one edge-edge pair (IPC.h:371-568)
__device__ inline void ee_narrow(const CSArgs& A, int inst, int eI, int eJ, const CSOut& O)
{
    const double r = A.rad[inst], r2 = A.rad[A.B + inst];
    int2 a = A.be[eI], b = A.be[eJ];
    if (a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y || eI > eJ) return;
    if (A.dbc[a.x] && A.dbc[a.y] && A.dbc[b.x] && A.dbc[b.y]) return;
    d3 a0 = ld3(A.X, a.x), a1 = ld3(A.X, a.y), b0 = ld3(A.X, b.x), b1 = ld3(A.X, b.y);
    if (!ee_cd_broadphase(a0, a1, b0, b1, r)) return;
    const double cr = ee_cross_sqn(a0, a1, b0, b1);
    const double eps = mollifier_eps(ld3(A.X0, a.x), ld3(A.X0, a.y), ld3(A.X0, b.x), ld3(A.X0, b.y));
    const bool m = cr < eps;
    const unsigned long long ekey = ee_pass_key(O, inst, eI, eJ);
    double d;
    switch (ee_type(a0, a1, b0, b1)) {
    case 0:
        d = d2_pp(a0, b0);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.x, b.x, -a.y - 1, -b.y - 1);
            else out_dup(O, inst, -a.x - 1, b.x, -1);
        }
        break;
    case 1:
        d = d2_pp(a0, b1);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.x, b.y, -a.y - 1, -b.x - 1);
            else out_dup(O, inst, -a.x - 1, b.y, -1);
        }
        break;
    case 2:
        d = d2_pe(a0, b0, b1);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.x, b.x, b.y, -a.y - 1);
            else out_dup(O, inst, -a.x - 1, b.x, b.y);
        }
        break;
    case 3:
        d = d2_pp(a1, b0);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.y, b.x, -a.x - 1, -b.y - 1);
            else out_dup(O, inst, -a.y - 1, b.x, -1);
        }
        break;
    case 4:
        d = d2_pp(a1, b1);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.y, b.y, -a.x - 1, -b.x - 1);
            else out_dup(O, inst, -a.y - 1, b.y, -1);
        }
        break;
    case 5:
        d = d2_pe(a1, b0, b1);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.y, b.x, b.y, -a.x - 1);
            else out_dup(O, inst, -a.y - 1, b.x, b.y);
        }
        break;
    case 6:
        d = d2_pe(b0, a0, a1);
        if (d < r2) {
            if (m) out_pass(O, ekey, b.x, a.x, a.y, -b.y - 1);
            else out_dup(O, inst, -b.x - 1, a.x, a.y);
        }
        break;
    case 7:
        d = d2_pe(b1, a0, a1);
        if (d < r2) {
            if (m) out_pass(O, ekey, b.y, a.x, a.y, -b.x - 1);
            else out_dup(O, inst, -b.y - 1, a.x, a.y);
        }
        break;
    default:
        d = d2_ee(a0, a1, b0, b1);
        if (d < r2) {
            if (m) out_pass(O, ekey, a.x, a.y, -b.x - 1, b.y);
            else out_pass(O, ekey, a.x, a.y, b.x, b.y);
        }
        break;
    }
}

// Constraint-set queries: q < nb are boundary vertices against triangles
// (SPATIAL_HASH.h:215-241, box [p - r, p + r]); the rest are boundary edges
// against edges eJ > eI (SPATIAL_HASH.h:243-291, edge box inflated by r).
// G lanes per query.
template <int G>
__global__ void __launch_bounds__(256) k_cs_walk(int nb, int ne, CSArgs A, CellTable T, CSOut O)
{
    const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long q = tid / G;
    const int lane = (int)(tid % G);
    if (q >= (long)nb + ne) return;
    int ql[3], qh[3];
    if (q < nb) {
        const int vI = A.bn[q];
        const int inst = A.vinst[vI];
        if (A.active && !A.active[inst]) return;
        const GridDev& g = A.grids[inst];
        const double r = A.rad[inst];
        const d3 p = ld3(A.X, vI);
        locate(g, p.x - r, p.y - r, p.z - r, ql);
        locate(g, p.x + r, p.y + r, p.z + r, qh);
        walk_lanes<G>(g, A.keybase[inst], T, A.pbox, 0, ql, qh, lane,
                      [&](int f) { pt_narrow(A, inst, vI, f, O); });
    } else {
        const int eI = (int)(q - nb);
        const int2 e = A.be[eI];
        const int inst = A.vinst[e.x];
        if (A.active && !A.active[inst]) return;
        const GridDev& g = A.grids[inst];
        const double r = A.rad[inst];
        const d3 a = ld3(A.X, e.x), b = ld3(A.X, e.y);
        locate(g, fmin(a.x, b.x) - r, fmin(a.y, b.y) - r, fmin(a.z, b.z) - r, ql);
        locate(g, fmax(a.x, b.x) + r, fmax(a.y, b.y) + r, fmax(a.z, b.z) + r, qh);
        walk_lanes<G>(g, A.keybase[inst], T, A.pbox + 6L * A.nf, 1, ql, qh, lane, [&](int f) {
            if (f > eI) ee_narrow(A, inst, eI, f, O);
        });
    }
}

// ---------------------------------------------------------------------------
// merge of PP/PE duplicates (IPC.h:599-654): key order = std::map order
// ---------------------------------------------------------------------------
__global__ void k_dup_heads(unsigned n, const unsigned long long* __restrict__ keys, unsigned* __restrict__ head)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    head[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

__global__ void k_dup_write(unsigned n, const unsigned long long* __restrict__ keys,
                            const unsigned* __restrict__ headscan, const int* __restrict__ vbeg, int inst_shift,
                            int4* __restrict__ out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(i == 0 || keys[i] != keys[i - 1])) return;
    unsigned j = i + 1;
    while (j < n && keys[j] == keys[i]) ++j;
    unsigned long long k = keys[i];
    int inst = 0;
    if (inst_shift) {
        inst = (int)(k >> inst_shift);
        k &= (1ull << inst_shift) - 1;
    }
    const long long vb = vbeg[inst];
    const unsigned long long V = (unsigned long long)(vbeg[inst + 1] - vb);
    const long long s2l = (long long)(k % (V + 1)) - 1;
    k /= (V + 1);
    const long long s1l = (long long)(k % V);
    const long long s0l = (long long)(k / V) - (long long)V;  // local -p-1
    const int s0 = (int)(s0l - vb), s1 = (int)(s1l + vb), s2 = s2l < 0 ? -1 : (int)(s2l + vb);
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
    const int* pbox;  // triangles [0, nf), then edges
    const GridDev* grids;
    const unsigned* keybase;
    const int* vinst;
    const unsigned char* active;  // null: every instance
    int nf;
    int cert;  // long pairs try the no-hit certificate
    double th;
};

// Candidate pairs that survive the swept AABB test, appended in any order:
// PT as (vertex, triangle) from the front of the buffer (counter cnt[0]), EE
// as (-eI-1, eJ) from the back (counter cnt[1]).  The step size is the minimum
// over candidates, so the order does not matter.
// Append for the lanes of a wavefront that call this together: one atomic
// per wavefront (the lanes with pred take consecutive slots in lane order);
// returns the slot, or UINT_MAX for a lane without pred.
__device__ inline unsigned wave_append(unsigned* ctr, bool pred)
{
    const unsigned long long m = __ballot(pred);
    if (m == 0) return UINT_MAX;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(ctr, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    return pred ? base + (unsigned)__popcll(m & ((1ull << lane) - 1)) : UINT_MAX;
}

constexpr long WALK_BIG_VOX = 512;  // CCD queries over more voxels are walked by a workgroup

// voxels a query box covers (0 when it lies outside the grid)
__device__ inline long walk_nvox(const GridDev& g, const int* ql, const int* qh)
{
    long n = 1;
    for (int d = 0; d < 3; ++d) {
        const int k = min(qh[d], g.cnt[d] - 1) - max(ql[d], 0) + 1;
        if (k <= 0) return 0;
        n *= k;
    }
    return n;
}

// Swept-box candidates of the CCD (SPATIAL_HASH.h:484-690): G lanes per
// query.  BIG: the first pass (BIG = false) hands every query whose box
// covers more than WALK_BIG_VOX voxels to a list that the second pass walks
// with a whole workgroup per query, so one fast-moving primitive does not
// hold the launch up on eight lanes.
template <int G, bool BIG>
__device__ inline void ccd_walk_query(long q, int lane, int nb, const CCDArgs& A, const CellTable& T,
                                      unsigned* __restrict__ cnt, int2* __restrict__ cand, unsigned cap,
                                      unsigned* __restrict__ nbig, unsigned* __restrict__ big)
{
    if (q < nb) {
        const int v = A.bn[q];
        const int inst = A.vinst[v];
        if (A.active && !A.active[inst]) return;
        const GridDev& g = A.grids[inst];
        if (!BIG && walk_nvox(g, A.vbox + 6 * q, A.vbox + 6 * q + 3) > WALK_BIG_VOX) {
            if (lane == 0) big[atomicAdd(nbig, 1u)] = (unsigned)q;
            return;
        }
        const d3 p = ld3(A.X, v), dp = ld3(A.P, v);
        walk_lanes<G>(g, A.keybase[inst], T, A.pbox, 0, A.vbox + 6 * q, A.vbox + 6 * q + 3, lane, [&](int f) {
            const int3 t = A.bt[f];
            const bool ok = !(v == t.x || v == t.y || v == t.z) &&
                            !(A.dbc[v] && A.dbc[t.x] && A.dbc[t.y] && A.dbc[t.z]) &&
                            pt_ccd_broadphase(p, ld3(A.X, t.x), ld3(A.X, t.y), ld3(A.X, t.z), dp, ld3(A.P, t.x),
                                              ld3(A.P, t.y), ld3(A.P, t.z), A.th);
            const unsigned i = wave_append(cnt, ok);
            if (ok && i < cap) cand[i] = make_int2(v, f);  // PT from the front
        });
    } else {
        const int eI = (int)(q - nb);
        const int2 a = A.be[eI];
        const int inst = A.vinst[a.x];
        if (A.active && !A.active[inst]) return;
        const GridDev& g = A.grids[inst];
        const int* pe = A.pbox + 6L * A.nf;
        const int* eb = pe + 6L * eI;
        if (!BIG && walk_nvox(g, eb, eb + 3) > WALK_BIG_VOX) {
            if (lane == 0) big[atomicAdd(nbig, 1u)] = (unsigned)q;
            return;
        }
        const d3 a0 = ld3(A.X, a.x), a1 = ld3(A.X, a.y), da0 = ld3(A.P, a.x), da1 = ld3(A.P, a.y);
        walk_lanes<G>(g, A.keybase[inst], T, pe, 1, eb, eb + 3, lane, [&](int f) {
            bool ok = f > eI;
            if (ok) {
                const int2 b = A.be[f];
                ok = !(a.x == b.x || a.x == b.y || a.y == b.x || a.y == b.y) &&
                     !(A.dbc[a.x] && A.dbc[a.y] && A.dbc[b.x] && A.dbc[b.y]) &&
                     ee_ccd_broadphase(a0, a1, ld3(A.X, b.x), ld3(A.X, b.y), da0, da1, ld3(A.P, b.x),
                                       ld3(A.P, b.y), A.th);
            }
            const unsigned i = wave_append(cnt + 1, ok);
            if (ok && i < cap) cand[cap - 1 - i] = make_int2(-eI - 1, f);  // EE from the back
        });
    }
}

template <int G, bool BIG>
__global__ void __launch_bounds__(256) k_ccd_walk(int nb, int ne, CCDArgs A, CellTable T,
                                                  unsigned* __restrict__ cnt, int2* __restrict__ cand, unsigned cap,
                                                  unsigned* __restrict__ nbig, unsigned* __restrict__ big)
{
    if (BIG) {
        // a workgroup per listed query (the list is final: the first pass has ended)
        const unsigned n = *nbig;
        for (unsigned k = blockIdx.x; k < n; k += gridDim.x)
            ccd_walk_query<G, true>((long)big[k], (int)threadIdx.x, nb, A, T, cnt, cand, cap, nbig, big);
    } else {
        const long tid = (long)blockIdx.x * blockDim.x + threadIdx.x;
        const long q = tid / G;
        if (q >= (long)nb + ne) return;
        ccd_walk_query<G, false>(q, (int)(tid % G), nb, A, T, cnt, cand, cap, nbig, big);
    }
}

// ACCD over the candidates as a work queue (amin: running minimum per instance).  Each lane carries one pair
// through the additive CCD iteration of CCD.h:280-395 (the reference's
// arithmetic, step by step); a lane whose pair ends (hit, passed the
// bound, or zero motion) takes the next pair, one atomic per wavefront for
// all idle lanes, so long and short pairs do not serialise a wavefront.
// PT candidates are queued before EE ones.  The bound is the running
// minimum, refreshed after every run of steps (with a value loaded a run earlier): a pair that cannot go below it cannot
// change the minimum, and a pair that does reaches the same time of impact
// whatever bound it started with, so the result is the minimum over pairs.
struct AccdLane {
    d3 x[4], dx[4];
    double maxDisp, d2, dc, f, gap, toc, tmax, d0, tmax0;
    double lb;  // the next step's lower bound, computed as soon as dc is known
    int guard, inst;
    int cert_at;  // step count at which the no-hit certificate is next tried
    int2 c;
    bool ee;
};

// steps a lane takes between visits to the work queue (ACCD_RUN_TAIL when its
// wavefront has at most two pairs left)
constexpr int ACCD_RUN = 32;
constexpr int ACCD_RUN_TAIL = 256;

// pairs that ran at least this many ACCD steps are logged (diagnostics)
constexpr int ACCD_LOG_STEPS = 1 << 16;
constexpr int ACCD_LOG_MAX = 8;
constexpr int ACCD_LOG_W = 34;  // doubles per entry (10 figures, then the pair's start x[4], dx[4])

__device__ inline double accd_dist(const AccdLane& L)
{
    return L.ee ? d2_ee_unclassified(L.x[0], L.x[1], L.x[2], L.x[3])
                : d2_pt_unclassified(L.x[0], L.x[1], L.x[2], L.x[3]);
}

// EE: fall back to the closest endpoint pair when the line distance is inside
// the thickness (CCD.h:350-356)
__device__ inline void accd_fix_ee(AccdLane& L, double th)
{
    if (L.ee && L.f <= 0) {
        L.d2 = fmin(fmin(sqn(L.x[0] - L.x[2]), sqn(L.x[0] - L.x[3])), fmin(sqn(L.x[1] - L.x[2]), sqn(L.x[1] - L.x[3])));
        L.f = L.d2 - th * th;
    }
}

// No-hit certificate of a long pair.  The iteration ends with a hit only where
// the distance it evaluates, minus the thickness, drops below gap = eta (d0 - th).
// Its evaluations lie in [toc, T]: every one before the last is at or below
// tmax, and the last one is at toc_k + lb_k with lb_k = (1 - eta)(d_k - th) /
// maxDisp; if d_k >= d0 there, the distance can fall to at most
// th + eta (d_k - th) (maxDisp bounds its rate), which is no hit, and otherwise
// toc_k + lb_k <= tmax + (1 - eta)(d0 - th) / maxDisp = T.  Every distance the
// iteration can evaluate (point-triangle: point-plane, point-edge or point-point;
// edge-edge: line-line, point-edge or point-point) is at least the distance
// between the point and the triangle's plane (PT) or between the two lines
// (EE), |g(t)| / |c(t)| with c = u x v and g = c . w, where u, v, w move
// linearly: g is a cubic and c a quadratic in t.  Over each of
// ACCD_CERT_PIECES pieces of [toc, T] the Bernstein coefficients of g bound
// |g| from below (when they share a sign) and the control points of c bound |c|
// from above; the pair cannot hit when the bound stays above the gap with a
// margin of 1e-3 of it, of the floating-point error of the coefficients (1e-12
// of the product of the moving vectors' sizes) and of the rounding the
// iteration's position updates can gather over the steps still to come (27
// standard deviations of a random walk of half-ulp errors, per coordinate of
// the four points).  Such a pair ends with no hit, as the steps would.
constexpr int ACCD_CERT_STEPS = 1024;  // pairs expected to take more steps try the certificate
constexpr int ACCD_CERT_PIECES = 64;

__device__ inline double ulp_of(double a)
{
    int e;
    frexp(a, &e);
    return ldexp(1.0, e - 53);
}

__device__ inline bool accd_certify(const AccdLane& L, double eta, double th)
{
    if (!(L.gap > 0) || !(L.maxDisp > 0)) return false;
    const double T = L.tmax * (1 + 1e-9) + (1 - eta) * (L.d0 - th) * (1 + 1e-6) / L.maxDisp + 1e-9;
    const double Tr = T - L.toc;
    if (!(Tr > 0)) return false;
    // EE: u = a1 - a0, v = b1 - b0, w = b0 - a0; PT: u = t1 - t0, v = t2 - t0, w = p - t0
    const int iu0 = L.ee ? 0 : 1, iu1 = L.ee ? 1 : 2, iv0 = L.ee ? 2 : 1, iv1 = 3, iw0 = L.ee ? 0 : 1,
              iw1 = L.ee ? 2 : 0;
    const d3 u0 = L.x[iu1] - L.x[iu0], u1 = L.dx[iu1] - L.dx[iu0];
    const d3 v0 = L.x[iv1] - L.x[iv0], v1 = L.dx[iv1] - L.dx[iv0];
    const d3 w0 = L.x[iw1] - L.x[iw0], w1 = L.dx[iw1] - L.dx[iw0];
    const d3 c0 = cross(u0, v0), c1 = cross(u0, v1) + cross(u1, v0), c2 = cross(u1, v1);
    const double g0 = dot(c0, w0), g1 = dot(c0, w1) + dot(c1, w0), g2 = dot(c1, w1) + dot(c2, w0),
                 g3 = dot(c2, w1);
    const double su = sqrt(sqn(u0)) + sqrt(sqn(u1)) * Tr, sv = sqrt(sqn(v0)) + sqrt(sqn(v1)) * Tr,
                 sw = sqrt(sqn(w0)) + sqrt(sqn(w1)) * Tr;
    const double errc = 1e-12 * su * sv, errg = errc * sw;
    double xm = 0.0, dm = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        xm = fmax(xm, fmax(fabs(L.x[k].x), fmax(fabs(L.x[k].y), fabs(L.x[k].z))));
        dm = fmax(dm, fmax(fabs(L.dx[k].x), fmax(fabs(L.dx[k].y), fabs(L.dx[k].z))));
    }
    const double ulp = ulp_of(xm + dm * Tr);
    const double kb = Tr * L.maxDisp / ((1 - eta) * L.gap) + 1;
    const double need = th + L.gap * (1 + 1e-3) + (32 * 1.7320508075688772 * sqrt(kb) + 64) * ulp;
    const double h = Tr / ACCD_CERT_PIECES;
    for (int i = 0; i < ACCD_CERT_PIECES; ++i) {
        const double ta = i * h;
        const double G0 = ((g3 * ta + g2) * ta + g1) * ta + g0;
        const double G1 = ((3 * g3 * ta + 2 * g2) * ta + g1) * h;
        const double G2 = (3 * g3 * ta + g2) * h * h;
        const double G3 = g3 * h * h * h;
        const double b0 = G0, b1 = G0 + G1 / 3, b2 = G0 + (2 * G1 + G2) / 3, b3 = G0 + G1 + G2 + G3;
        const double bmin = fmin(fmin(b0, b1), fmin(b2, b3)), bmax = fmax(fmax(b0, b1), fmax(b2, b3));
        double lo;
        if (bmin - errg > 0)
            lo = bmin - errg;
        else if (bmax + errg < 0)
            lo = -bmax - errg;
        else
            return false;
        const d3 C0 = ta * (ta * c2 + c1) + c0, C1 = h * (2 * ta * c2 + c1), C2 = (h * h) * c2;
        const d3 q1 = C0 + 0.5 * C1, q2 = (C0 + C1) + C2;
        const double cmax = sqrt(fmax(sqn(C0), fmax(sqn(q1), sqn(q2)))) + errc;
        if (!(lo > need * cmax)) return false;
    }
    return true;
}

// returns false when the pair has no motion (no hit)
__device__ inline bool accd_init(AccdLane& L, int2 c, const CCDArgs& A, double eta,
                                 const unsigned long long* amin)
{
    const double th = A.th;
    if (c.x >= 0) {
        const int3 t = A.bt[c.y];
        L.ee = false;
        L.inst = A.vinst[c.x];
        L.x[0] = ld3(A.X, c.x); L.x[1] = ld3(A.X, t.x); L.x[2] = ld3(A.X, t.y); L.x[3] = ld3(A.X, t.z);
        d3 dp = ld3(A.P, c.x), dt0 = ld3(A.P, t.x), dt1 = ld3(A.P, t.y), dt2 = ld3(A.P, t.z);
        const d3 mov = 0.25 * (((dt0 + dt1) + dt2) + dp);
        dt0 = dt0 - mov; dt1 = dt1 - mov; dt2 = dt2 - mov; dp = dp - mov;
        L.dx[0] = dp; L.dx[1] = dt0; L.dx[2] = dt1; L.dx[3] = dt2;
        const double m2 = fmax(sqn(dt0), fmax(sqn(dt1), sqn(dt2)));
        L.maxDisp = sqrt(sqn(dp)) + sqrt(m2);
    } else {
        const int2 a = A.be[-c.x - 1], b = A.be[c.y];
        L.ee = true;
        L.inst = A.vinst[a.x];
        L.x[0] = ld3(A.X, a.x); L.x[1] = ld3(A.X, a.y); L.x[2] = ld3(A.X, b.x); L.x[3] = ld3(A.X, b.y);
        d3 da0 = ld3(A.P, a.x), da1 = ld3(A.P, a.y), db0 = ld3(A.P, b.x), db1 = ld3(A.P, b.y);
        const d3 mov = 0.25 * (((da0 + da1) + db0) + db1);
        da0 = da0 - mov; da1 = da1 - mov; db0 = db0 - mov; db1 = db1 - mov;
        L.dx[0] = da0; L.dx[1] = da1; L.dx[2] = db0; L.dx[3] = db1;
        L.maxDisp = sqrt(fmax(sqn(da0), sqn(da1))) + sqrt(fmax(sqn(db0), sqn(db1)));
    }
    if (L.maxDisp == 0) return false;
    L.d2 = accd_dist(L);
    L.f = L.d2 - th * th;
    accd_fix_ee(L, th);
    L.dc = sqrt(L.d2);
    L.gap = eta * L.f / (L.dc + th);
    L.lb = (1 - eta) * L.f / ((L.dc + th) * L.maxDisp);
    L.toc = 0;
    L.tmax = __longlong_as_double((long long)*((volatile const unsigned long long*)(amin + L.inst)));
    L.guard = 0;
    L.d0 = L.dc;
    L.tmax0 = L.tmax;
    L.c = c;
    L.cert_at = ACCD_CERT_STEPS;
    return true;
}

// one additive step; returns 0 running, 1 hit (toc), 2 no hit, 3 stopped by
// the step cap (max_steps > 0) and taken as a hit at the toc reached.  The
// next step's bound is computed before the gap test is resolved (the same
// expression the loop would evaluate after it), so the two divisions overlap
// instead of following each other across the branch.
// The kind (EE or PT) is a template argument, so that a run of steps is one
// loop without the kind's branch.
template <bool EE>
__device__ inline int accd_step(AccdLane& L, double eta, double th, int max_steps)
{
    const double lb = L.lb;
#pragma unroll
    for (int k = 0; k < 4; ++k) L.x[k] = L.x[k] + lb * L.dx[k];
    if (EE) {
        L.d2 = d2_ee_unclassified(L.x[0], L.x[1], L.x[2], L.x[3]);
        L.f = L.d2 - th * th;
        accd_fix_ee(L, th);
    } else {
        L.d2 = d2_pt_unclassified(L.x[0], L.x[1], L.x[2], L.x[3]);
        L.f = L.d2 - th * th;
    }
    L.dc = sqrt(L.d2);
    const double den = L.dc + th;
    L.lb = (1 - eta) * L.f / (den * L.maxDisp);
    if (L.toc != 0.0 && (L.f / den < L.gap)) return 1;
    L.toc += lb;
    if (L.toc > L.tmax) return 2;
    if (++L.guard >= max_steps && max_steps > 0) return 3;
    return 0;
}

__global__ void __launch_bounds__(256) k_accd(const unsigned* __restrict__ cnt, unsigned cap,
                                              const int2* __restrict__ cand, CCDArgs A, double eta,
                                              unsigned long long* __restrict__ amin, unsigned* __restrict__ next,
                                              unsigned long long* __restrict__ stepstat /* total, max, capped, certified */,
                                              unsigned* __restrict__ nlog, double* __restrict__ lg, int budget,
                                              int max_steps, unsigned* __restrict__ capcnt,
                                              const AccdLane* __restrict__ res_in, unsigned nres,
                                              AccdLane* __restrict__ res_out, unsigned* __restrict__ nres_out,
                                              unsigned res_cap, unsigned char* __restrict__ pend)
{
    const unsigned npt = min(cnt[0], cap), nee = min(cnt[1], cap - npt);
    const unsigned n = nres + npt + nee;  // resumed pairs first, then PT, then EE candidates
    const int lane = threadIdx.x & 63;
    AccdLane L;
    bool active = false, finished = false;
    int steps = 0, slice = 0;
    double amin_pf = 0.0;  // running minimum of this lane's instance, loaded ahead of use
    for (;;) {
        const unsigned long long idle = __ballot(!active && !finished);
        if (idle) {
            const int leader = __ffsll((long long)idle) - 1;
            unsigned base = 0;
            if (lane == leader) base = atomicAdd(next, (unsigned)__popcll(idle));
            base = __shfl(base, leader);
            if (!active && !finished) {
                const unsigned j = base + (unsigned)__popcll(idle & ((1ull << lane) - 1));
                slice = 0;
                if (j >= n) {
                    finished = true;
                } else if (j < nres) {
                    // a pair stopped by the step budget of an earlier call: its state is
                    // continued exactly; only its bound may have come down since
                    L = res_in[j];
                    L.tmax = fmin(L.tmax, __longlong_as_double(
                                              (long long)*((volatile unsigned long long*)(amin + L.inst))));
                    amin_pf = L.tmax;
                    active = true;
                } else {
                    const unsigned q = j - nres;
                    const int2 c = q < npt ? cand[q] : cand[cap - 1 - (q - npt)];
                    active = accd_init(L, c, A, eta, amin);
                    amin_pf = L.tmax;
                    // a pair whose first step predicts a long run tries the certificate
                    // after its first run of steps, every pair after ACCD_CERT_STEPS
                    if (!A.cert)
                        L.cert_at = 0x7fffffff;
                    else if (L.tmax * L.maxDisp > ACCD_CERT_STEPS * (1 - eta) * (L.dc - A.th))
                        L.cert_at = 0;
                }
            }
        }
        if (__ballot(active) == 0) {
            if (__ballot(!finished) == 0) break;
            continue;
        }
        if (active) {
            // a run of steps between visits to the queue (a long pair's lane does
            // little besides stepping); it stops at the pair's end or the budget
            // a pair with at most one other left in its wavefront (the tail of a long
            // pair) takes longer runs between visits to the queue
            const int run = __popcll(__ballot(1)) <= 2 ? ACCD_RUN_TAIL : ACCD_RUN;
            const int room = budget > 0 ? min(run, budget - slice) : run;
            int r = 0, k = 0;
            if (L.ee) {
                do {
                    r = accd_step<true>(L, eta, A.th, max_steps);
                    ++k;
                } while (r == 0 && k < room);
            } else {
                do {
                    r = accd_step<false>(L, eta, A.th, max_steps);
                    ++k;
                } while (r == 0 && k < room);
            }
            steps += k;
            slice += k;
            if (r == 0 && L.guard >= L.cert_at) {
                // a long pair tries the certificate again at 4x the steps it has taken
                if (accd_certify(L, eta, A.th)) {
                    r = 2;
                    atomicAdd(stepstat + 3, 1ull);
                } else {
                    L.cert_at = L.guard > (1 << 28) ? 0x7fffffff : max(4 * L.guard, 256);
                }
            }
            if (r == 1 || r == 3) atomicMin(amin + L.inst, (unsigned long long)__double_as_longlong(L.toc));
            if (r == 3) {
                atomicAdd(stepstat + 2, 1ull);
                atomicAdd(capcnt + L.inst, 1u);
            }
            if (r != 0) {
                active = false;
                atomicMax(stepstat + 1, (unsigned long long)L.guard + 1);
                if (L.guard + 1 >= ACCD_LOG_STEPS) {
                    const unsigned k = atomicAdd(nlog, 1u);
                    if (k < ACCD_LOG_MAX) {
                        double* e = lg + ACCD_LOG_W * k;
                        e[0] = L.c.x;
                        e[1] = L.c.y;
                        e[2] = L.inst;
                        e[3] = L.d0;
                        e[4] = L.maxDisp;
                        e[5] = L.tmax0;
                        e[6] = L.toc;
                        e[7] = L.guard + 1;
                        e[8] = r;
                        e[9] = L.dc;
                        AccdLane L0;
                        if (accd_init(L0, L.c, A, eta, amin))
                            for (int v = 0; v < 4; ++v) {
                                e[10 + 3 * v] = L0.x[v].x, e[11 + 3 * v] = L0.x[v].y, e[12 + 3 * v] = L0.x[v].z;
                                e[22 + 3 * v] = L0.dx[v].x, e[23 + 3 * v] = L0.dx[v].y, e[24 + 3 * v] = L0.dx[v].z;
                            }
                    }
                }
            } else if (budget > 0 && slice >= budget) {
                // step budget of this call spent: hand the pair's state to the next call
                const unsigned k = atomicAdd(nres_out, 1u);
                if (k < res_cap) {
                    res_out[k] = L;
                    pend[L.inst] = 1;
                    active = false;
                } else {
                    slice = -(1 << 30);  // no room: run this pair to its end (past the guard)
                }
            } else {
                // the bound loaded one run ago (its latency hidden behind the run); reload
                L.tmax = fmin(L.tmax, amin_pf);
                amin_pf = __longlong_as_double((long long)*((volatile unsigned long long*)(amin + L.inst)));
            }
        }
    }
    // total steps of this lane (all its pairs)
    const unsigned long long tot = (unsigned long long)steps;
    if (tot) atomicAdd(stepstat, tot);
}

// span cap (SPATIAL_HASH.h:466-482), one thread per instance; also seeds the
// instance's ACCD bound
__global__ void k_ccd_cap(int B, const double* __restrict__ eparts, const int* __restrict__ epbeg,
                          const int* __restrict__ ebeg, const double* __restrict__ psum, const int* __restrict__ nbeg,
                          const double* __restrict__ alpha_in, double* __restrict__ alpha_out,
                          unsigned long long* __restrict__ amin, const unsigned char* __restrict__ active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B || (active && !active[i])) return;
    double vs = 0.0, ps = 0.0;
    for (int c = epbeg[i]; c < epbeg[i + 1]; ++c) vs += eparts[c];
    const long ne = ebeg[i + 1] - ebeg[i], nb = nbeg[i + 1] - nbeg[i];
    vs = ne ? vs / (double)ne : 1.0;
    ps = psum[i] / (double)(nb * 3);
    double a = alpha_in[i];
    const double span = a * ps / vs;
    if (span > 1) a /= span;
    alpha_out[i] = a;
    amin[i] = (unsigned long long)__double_as_longlong(a);
}

__global__ void k_reset_amin(int B, const double* __restrict__ alpha, unsigned long long* __restrict__ amin,
                             const unsigned char* __restrict__ active)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < B && (!active || active[i])) amin[i] = (unsigned long long)__double_as_longlong(alpha[i]);
}

__global__ void k_zero_u64(unsigned long long* __restrict__ a, int n)
{
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0ull;
}

__global__ void k_zero_u8(unsigned char* __restrict__ a, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

__global__ void k_zero_u32(unsigned* __restrict__ a, int n)
{
    if ((int)threadIdx.x < n) a[threadIdx.x] = 0u;
}

// per-instance counts of sorted pass tuples (instance in the key's high bits)
// and of merged tuples (instance of their first vertex)
__global__ void k_count_pass(unsigned n, const unsigned long long* __restrict__ keys, unsigned* __restrict__ cnt)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(cnt + (unsigned)(keys[i] >> 40), 1u);
}
__global__ void k_count_merged(unsigned n, const int4* __restrict__ t, const int* __restrict__ vinst,
                               unsigned* __restrict__ cnt)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(cnt + vinst[-t[i].x - 1], 1u);
}

// place sorted pass tuples and merged tuples at their instance's range:
// instance i gets [pass_i | merged_i] at cbeg[i]
__global__ void k_place_pass(unsigned n, const unsigned long long* __restrict__ keys, const int4* __restrict__ t,
                             const long* __restrict__ cbeg, const long* __restrict__ pstart,
                             int4* __restrict__ out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned inst = (unsigned)(keys[i] >> 40);
    out[cbeg[inst] + (i - pstart[inst])] = t[i];
}
__global__ void k_place_merged(unsigned n, const int4* __restrict__ t, const int* __restrict__ vinst,
                               const long* __restrict__ cbeg, const long* __restrict__ mstart,
                               const unsigned* __restrict__ pcnt, int4* __restrict__ out)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int inst = vinst[-t[i].x - 1];
    out[cbeg[inst] + pcnt[inst] + (i - mstart[inst])] = t[i];
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------
namespace {

constexpr int WALK_G = 8;                        // lanes per query
constexpr unsigned long long DENSE_MAX = 1ull << 25;  // max record keys of a dense cell table

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

int walk_blocks(long nq) { return grid_for(nq * WALK_G, 256); }

// Boxes of the triangles and edges from the vertex boxes, their sorted
// (keybase[instance] + 2 * voxel + type, primitive) records and the cell
// table.  One host synchronisation (record count and the instance grids).
CellTable build_cells(ContactWork& w, hipStream_t s, const MeshDev& m, const int* vbox,
                      const unsigned char* active)
{
    const InstDev& I = m.inst;
    const int B = I.B;
    const int n = m.nf + m.ne;
    int* pbox = w.pbox.ensure(6 * (size_t)std::max(n, 1), s);
    unsigned* cnt = w.u0.ensure(n + 1, s);
    unsigned* off = w.u1.ensure(n + 1, s);
    unsigned* pinst = w.pinst.ensure(n + 1, s);
    IDP_LAUNCH(k_prim_boxes<3>, grid_for(m.nf, 256), 256, 0, s, m.nf, m.bt.p, m.v2b.p, I.vinst.p, active, vbox,
               w.grids.p, pbox, cnt, pinst);
    IDP_LAUNCH(k_prim_boxes<2>, grid_for(m.ne, 256), 256, 0, s, m.ne, m.be.p, m.v2b.p, I.vinst.p, active, vbox,
               w.grids.p, pbox + 6 * (size_t)m.nf, cnt + m.nf, pinst + m.nf);
    IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, cnt + n, 1);
    excl_scan(s, w.tmp, cnt, off, n + 1);
    unsigned total = 0;
    std::vector<GridDev> hg(B);
    IDP_HIP(hipMemcpyAsync(&total, off + n, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipMemcpyAsync(hg.data(), w.grids.p, B * sizeof(GridDev), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    CellTable T{};
    T.n = total;
    // key bases: instance i owns keys [base_i, base_i + 2 * voxels_i)
    std::vector<unsigned> base(B + 1, 0);
    unsigned long long acc = 0;
    for (int i = 0; i < B; ++i) {
        base[i] = (unsigned)acc;
        acc += 2ull * hg[i].cnt[0] * hg[i].cnt[1] * hg[i].cnt[2];
        if (acc >= 0xFFFFFFFFull) throw HipError(IDP_E_CAPACITY, "voxel key space exceeds 32 bits");
    }
    base[B] = (unsigned)acc;
    w.keybase.upload(base.data(), B + 1, s);
    if (total == 0) return T;
    unsigned* k0 = w.k0.ensure(total, s);
    int* v0 = w.v0.ensure(total, s);
    IDP_LAUNCH(k_emit_cells, grid_for(n, 256), 256, 0, s, n, m.nf, pbox, off, pinst, w.keybase.p, w.grids.p, k0,
               v0);
    unsigned* k1 = w.k1.ensure(total, s);
    int* v1 = w.v1.ensure(total, s);
    const unsigned long long nkeys = acc;
    sort_pairs(s, w.tmp, k0, k1, v0, v1, total, bits_for(nkeys));
    T.keys = k1;
    T.vals = v1;
    if (nkeys <= DENSE_MAX) {
        if (w.cstart.cap < nkeys || !w.cstart.p) {
            w.cstart.ensure(nkeys, s);
            w.cend.ensure(nkeys, s);
            IDP_LAUNCH(k_fill_u32, grid_for((long)w.cstart.cap, 256), 256, 0, s, w.cstart.cap, w.cstart.p, 0u);
            IDP_LAUNCH(k_fill_u32, grid_for((long)w.cend.cap, 256), 256, 0, s, w.cend.cap, w.cend.p, 0u);
        }
        IDP_LAUNCH(k_cell_bounds, grid_for(total, 256), 256, 0, s, total, k1, w.cstart.p, w.cend.p, 0);
        T.start = w.cstart.p;
        T.end = w.cend.p;
    }
    return T;
}

void clear_cells(ContactWork& w, hipStream_t s, const CellTable& T)
{
    if (T.start && T.n)
        IDP_LAUNCH(k_cell_bounds, grid_for(T.n, 256), 256, 0, s, T.n, T.keys, w.cstart.p, w.cend.p, 1);
}

// per-instance grids from the chunked mean edge length (ep, echunk table)
// and the bounding box of the elements of seg (idx: element -> vertex)
void grid_from(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const SegDev& seg, const int* idx,
               const double* p, const double* alpha_p, double pad, const double* ep, const unsigned char* active)
{
    const InstDev& I = m.inst;
    const int B = I.B;
    double* mm = w.mmparts.ensure(6 * (size_t)seg.nblk + 6 * (size_t)B + 6, s);
    double* mmf = mm + 6 * (size_t)seg.nblk;
    IDP_LAUNCH(k_minmax_parts, seg.nblk, RB, 0, s, seg.lo.p, seg.hi.p, seg.inst.p, idx, X, p, alpha_p, mm);
    IDP_LAUNCH(k_minmax_final, grid_for(B, 64), 64, 0, s, B, seg.pbeg.p, mm, mmf);
    w.grids.ensure(B, s);
    IDP_LAUNCH(k_grid_setup, grid_for(B, 64), 64, 0, s, B, w.grids.p, ep, I.echunk.pbeg.p, I.d_ebeg.p, mmf, pad,
               active);
}

}  // namespace

// Constraint set at X (device) for the instances of the active mask (device,
// one byte per instance; null = all); the other instances keep theirs.
// Result in w.cs, instance i's constraints at [w.h_cbeg[i], w.h_cbeg[i+1]),
// each in the order PT pass tuples by (vertex, triangle), EE pass tuples by
// (edge, edge), then the merged PP/PE entries in map order.
long contact_constraint_set(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* dHat2,
                            double thickness, const unsigned char* active)
{
    const InstDev& I = m.inst;
    const int B = I.B;
    std::vector<double> rad(2 * B);
    for (int i = 0; i < B; ++i) {
        const double r = sqrt(dHat2[i]) + thickness;  // IPC.h:51-52
        rad[i] = r;
        rad[B + i] = r * r;
    }
    const std::vector<long> old_cbeg = w.h_cbeg.size() == (size_t)B + 1 ? w.h_cbeg : std::vector<long>(B + 1, 0);
    const long old_ncs = w.ncs;
    // rad is uploaded below, after the early exit
    if (m.nb == 0) {
        w.ncs = 0;
        w.h_cbeg.assign(B + 1, 0);
        w.cbeg.upload(w.h_cbeg.data(), B + 1, s);
        return 0;
    }
    double* ep = w.dparts.ensure(I.echunk.nblk + 1, s);
    IDP_LAUNCH(k_edge_len_parts, I.echunk.nblk, 256, 0, s, I.echunk.lo.p, I.echunk.hi.p, (const int2*)m.be.p, X, ep);
    grid_from(w, s, m, X, I.vblk, nullptr, nullptr, nullptr, 0.0, ep, active);
    int* vbox = w.vbox.ensure(6 * (size_t)m.nb, s);
    IDP_LAUNCH(k_vertex_boxes, grid_for(m.nb, 256), 256, 0, s, m.nb, m.bn.p, I.vinst.p, X, (const double*)nullptr,
               (const double*)nullptr, 0.0, w.grids.p, vbox);
    CellTable T = build_cells(w, s, m, vbox, active);

    CSArgs A;
    A.X = X;
    A.X0 = m.X0.p;
    A.dbc = m.dbc.p;
    A.bn = m.bn.p;
    A.be = (const int2*)m.be.p;
    A.bt = (const int3*)m.bt.p;
    A.pbox = w.pbox.p;
    A.grids = w.grids.p;
    A.keybase = w.keybase.p;
    A.vinst = I.vinst.p;
    A.active = active;
    A.rad = w.rad.upload_get(rad.data(), rad.size(), s);
    A.B = B;
    A.nf = m.nf;
    CSOut O;
    O.vbeg = I.d_vbeg.p;
    O.ebeg = I.d_ebeg.p;
    O.fbeg = I.d_fbeg.p;
    O.inst_shift = B > 1 ? 56 : 0;
    unsigned* cnt = w.cnt.ensure(6, s);
    if (w.cap_pass == 0) w.cap_pass = 4u * (m.nb + m.ne) + 1024;
    if (w.cap_dup == 0) w.cap_dup = 4u * (m.nb + m.ne) + 1024;
    unsigned hc[2] = {0, 0};
    if (T.n) {
        for (int attempt = 0; attempt < 2; ++attempt) {
            O.cnt = cnt;
            O.pkey = w.pk0.ensure(w.cap_pass, s);
            O.pval = w.pass.ensure(w.cap_pass, s);
            O.dkey = w.dk0.ensure(w.cap_dup, s);
            O.cap_p = (unsigned)std::min(w.pk0.cap, w.pass.cap);
            O.cap_d = (unsigned)w.dk0.cap;
            IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, cnt, 2);
            IDP_LAUNCH(k_cs_walk<WALK_G>, walk_blocks((long)m.nb + m.ne), 256, 0, s, m.nb, m.ne, A, T, O);
            IDP_HIP(hipMemcpyAsync(hc, cnt, 2 * sizeof(unsigned), hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
            if (hc[0] <= O.cap_p && hc[1] <= O.cap_d) break;
            w.cap_pass = std::max<size_t>(w.cap_pass, hc[0] + hc[0] / 4 + 1024);
            w.cap_dup = std::max<size_t>(w.cap_dup, hc[1] + hc[1] / 4 + 1024);
            if (attempt == 1) throw HipError(IDP_E_CAPACITY, "constraint-set output overflow");
        }
        clear_cells(w, s, T);
    }
    const unsigned npass = hc[0], nd = hc[1];
    // ---- order the pass tuples
    if (npass) {
        unsigned long long* pk1 = w.pk1.ensure(npass, s);
        int4* pv1 = w.dup.ensure(npass, s);
        sort_pairs(s, w.tmp, w.pk0.p, pk1, w.pass.p, pv1, npass, bits_for(((unsigned long long)B << 40) | (1ull << 40)));
    }
    // ---- merge duplicates (IPC.h:599-654)
    unsigned nmerged = 0;
    if (nd) {
        unsigned long long* k1 = w.dk1.ensure(nd, s);
        sort_keys(s, w.tmp, w.dk0.p, k1, nd, 64);
        unsigned* head = w.u0.ensure(nd + 1, s);
        unsigned* hs = w.u1.ensure(nd + 1, s);
        IDP_LAUNCH(k_dup_heads, grid_for(nd, 256), 256, 0, s, nd, k1, head);
        IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, head + nd, 1);
        excl_scan(s, w.tmp, head, hs, nd + 1);
        nmerged = read_u(s, hs + nd);
        w.merged.ensure(nmerged + 1, s);
        IDP_LAUNCH(k_dup_write, grid_for(nd, 256), 256, 0, s, nd, k1, hs, I.d_vbeg.p, O.inst_shift, w.merged.p);
    }
    if (B == 1 && !active) {
        const long total = (long)npass + nmerged;
        int4* cs = w.cs.ensure(total + 1, s);
        if (npass) IDP_HIP(hipMemcpyAsync(cs, w.dup.p, sizeof(int4) * npass, hipMemcpyDeviceToDevice, s));
        if (nmerged)
            IDP_HIP(hipMemcpyAsync(cs + npass, w.merged.p, sizeof(int4) * nmerged, hipMemcpyDeviceToDevice, s));
        w.ncs = total;
        w.h_cbeg = {0, total};
        w.cbeg.upload(w.h_cbeg.data(), 2, s);
        return total;
    }
    // ---- several instances: [pass_i | merged_i] per instance; inactive
    // instances keep their previous constraints
    std::vector<unsigned char> act(B, 1);
    if (active) {
        IDP_HIP(hipMemcpyAsync(act.data(), active, B, hipMemcpyDeviceToHost, s));
    }
    unsigned* ic = w.icnt.ensure(2 * (size_t)B, s);
    IDP_HIP(hipMemsetAsync(ic, 0, 2 * B * sizeof(unsigned), s));
    if (npass) IDP_LAUNCH(k_count_pass, grid_for(npass, 256), 256, 0, s, npass, w.pk1.p, ic);
    if (nmerged) IDP_LAUNCH(k_count_merged, grid_for(nmerged, 256), 256, 0, s, nmerged, w.merged.p, I.vinst.p, ic + B);
    std::vector<unsigned> hcnt(2 * B);
    IDP_HIP(hipMemcpyAsync(hcnt.data(), ic, 2 * B * sizeof(unsigned), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    std::vector<long> cbeg(B + 1, 0), pstart(B, 0), mstart(B, 0);
    long ps = 0, ms = 0;
    for (int i = 0; i < B; ++i) {
        pstart[i] = ps;
        mstart[i] = ms;
        ps += hcnt[i];
        ms += hcnt[B + i];
        const long n_i = act[i] ? (long)hcnt[i] + hcnt[B + i] : old_cbeg[i + 1] - old_cbeg[i];
        cbeg[i + 1] = cbeg[i] + n_i;
    }
    const long total = cbeg[B];
    // previous constraints of inactive instances
    bool keep = false;
    for (int i = 0; i < B; ++i) keep |= !act[i] && old_cbeg[i + 1] > old_cbeg[i];
    if (keep) {
        w.cs_old.ensure(old_ncs + 1, s);
        IDP_HIP(hipMemcpyAsync(w.cs_old.p, w.cs.p, sizeof(int4) * old_ncs, hipMemcpyDeviceToDevice, s));
    }
    int4* cs = w.cs.ensure(total + 1, s, keep);
    w.cbeg.upload(cbeg.data(), B + 1, s);
    long* dps = (long*)w.dparts2.ensure(2 * (size_t)B + 2, s);
    IDP_HIP(hipMemcpyAsync(dps, pstart.data(), B * sizeof(long), hipMemcpyHostToDevice, s));
    IDP_HIP(hipMemcpyAsync(dps + B, mstart.data(), B * sizeof(long), hipMemcpyHostToDevice, s));
    if (keep)
        for (int i = 0; i < B; ++i)
            if (!act[i] && old_cbeg[i + 1] > old_cbeg[i])
                IDP_HIP(hipMemcpyAsync(cs + cbeg[i], w.cs_old.p + old_cbeg[i],
                                       sizeof(int4) * (old_cbeg[i + 1] - old_cbeg[i]), hipMemcpyDeviceToDevice, s));
    if (npass) IDP_LAUNCH(k_place_pass, grid_for(npass, 256), 256, 0, s, npass, w.pk1.p, w.dup.p, w.cbeg.p, dps, cs);
    if (nmerged)
        IDP_LAUNCH(k_place_merged, grid_for(nmerged, 256), 256, 0, s, nmerged, w.merged.p, I.vinst.p, w.cbeg.p,
                   dps + B, ic, cs);
    IDP_HIP(hipStreamSynchronize(s));  // host vectors above are read by async copies
    w.ncs = total;
    w.h_cbeg = cbeg;
    return total;
}

// ACCD step sizes along P from X, per instance: alpha_io (host, B) holds the
// upper bounds on entry and the step sizes on return; instances outside the
// active mask (device, null = all) are left untouched.
void contact_ccd_batch(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                       double thickness, double* alpha_io, const unsigned char* active, long* ncand_out,
                       int budget, bool resume, unsigned char* pending)
{
    const InstDev& I = m.inst;
    const int B = I.B;
    if (ncand_out) *ncand_out = 0;
    if (pending) std::fill(pending, pending + B, (unsigned char)0);
    if (m.nb == 0) return;
    // pairs handed over by the previous call (step budget), and room for this call's
    constexpr unsigned RES_CAP = 4096;
    const unsigned nres = resume ? w.n_res : 0u;
    const AccdLane* res_in = (const AccdLane*)w.accd_res[w.res_cur].ensure(RES_CAP * sizeof(AccdLane), s);
    AccdLane* res_out = (AccdLane*)w.accd_res[w.res_cur ^ 1].ensure(RES_CAP * sizeof(AccdLane), s);
    unsigned char* pend = w.pend.ensure(B + 1, s);
    // span cap (mean edge length and mean |p| per instance), kept on the device
    double* ep = w.dparts.ensure(I.echunk.nblk + B + 2, s);
    double* pp = ep + I.echunk.nblk + 1;
    IDP_LAUNCH(k_edge_len_parts, I.echunk.nblk, 256, 0, s, I.echunk.lo.p, I.echunk.hi.p, (const int2*)m.be.p, X, ep);
    IDP_LAUNCH(k_psize_seq, grid_for(B, 64), 64, 0, s, B, I.d_nbeg.p, m.bn.p, P, pp, active);
    double* ain = (double*)w.scal2.ensure(2 * (size_t)B + 4, s);
    IDP_HIP(hipMemcpyAsync(ain, alpha_io, B * sizeof(double), hipMemcpyHostToDevice, s));
    double* dalpha = w.ialpha.ensure(B, s);
    unsigned long long* amin = w.iamin.ensure(B, s);
    unsigned* cnt = w.cnt.ensure(6, s);
    unsigned long long* stepstat = (unsigned long long*)w.scal.ensure(6, s) + 2;  // total, max, capped, certified
    IDP_LAUNCH(k_ccd_cap, grid_for(B, 64), 64, 0, s, B, ep, I.echunk.pbeg.p, I.d_ebeg.p, pp, I.d_nbeg.p, ain, dalpha,
               amin, active);
    IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, cnt, 6);
    IDP_LAUNCH(k_zero_u64, 1, 64, 0, s, stepstat, 4);
    IDP_LAUNCH(k_zero_u8, grid_for(B, 256), 256, 0, s, pend, B);
    unsigned* capcnt = w.icap.ensure(B, s);
    IDP_LAUNCH(k_fill_u32, grid_for(B, 256), 256, 0, s, (size_t)B, capcnt, 0u);
    double* lg = w.accd_log.ensure(ACCD_LOG_W * ACCD_LOG_MAX, s);

    // swept grid over boundary nodes at [x, x + alpha p] -/+ th/2 (SPATIAL_HASH.h:484-511)
    grid_from(w, s, m, X, I.nblk, m.bn.p, P, dalpha, thickness / 2, ep, active);
    int* vbox = w.vbox.ensure(6 * (size_t)m.nb, s);
    IDP_LAUNCH(k_vertex_boxes, grid_for(m.nb, 256), 256, 0, s, m.nb, m.bn.p, I.vinst.p, X, P, dalpha,
               thickness / 2, w.grids.p, vbox);
    CellTable T = build_cells(w, s, m, vbox, active);

    CCDArgs A;
    A.X = X;
    A.P = P;
    A.dbc = m.dbc.p;
    A.bn = m.bn.p;
    A.be = (const int2*)m.be.p;
    A.bt = (const int3*)m.bt.p;
    A.vbox = vbox;
    A.pbox = w.pbox.p;
    A.grids = w.grids.p;
    A.keybase = w.keybase.p;
    A.vinst = I.vinst.p;
    A.active = active;
    A.nf = m.nf;
    A.cert = w.accd_cert ? 1 : 0;
    A.th = thickness;
    if (w.cap_cand == 0) w.cap_cand = 16u * (m.nb + m.ne) + 4096;
    std::vector<unsigned long long> ha(B);
    unsigned long long hsteps[4] = {0, 0, 0, 0};
    unsigned hn[6] = {0, 0, 0, 0, 0, 0};
    for (int attempt = 0; attempt < 2; ++attempt) {
        int2* cand = w.cand.ensure(w.cap_cand, s);
        const unsigned cap = (unsigned)w.cand.cap;
        if (attempt) {
            IDP_LAUNCH(k_reset_amin, grid_for(B, 64), 64, 0, s, B, dalpha, amin, active);
            IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, cnt, 6);
            IDP_LAUNCH(k_zero_u64, 1, 64, 0, s, stepstat, 4);
            IDP_LAUNCH(k_zero_u8, grid_for(B, 256), 256, 0, s, pend, B);
            IDP_LAUNCH(k_fill_u32, grid_for(B, 256), 256, 0, s, (size_t)B, capcnt, 0u);
        }
        if (T.n) {
            unsigned* nbig = cnt + 5;
            unsigned* big = w.walk_big.ensure((size_t)m.nb + m.ne + 1, s);
            IDP_LAUNCH(k_zero_u32, 1, 64, 0, s, nbig, 1);
            IDP_LAUNCH((k_ccd_walk<WALK_G, false>), walk_blocks((long)m.nb + m.ne), 256, 0, s, m.nb, m.ne, A, T,
                       cnt, cand, cap, nbig, big);
            // large queries: one workgroup each
            IDP_LAUNCH((k_ccd_walk<256, true>), (int)std::min<long>((long)m.nb + m.ne, 2048), 256, 0, s, m.nb,
                       m.ne, A, T, cnt, cand, cap, nbig, big);
        }
        const int tok = w.prof ? w.prof->begin(IDP_K_CCD) : -1;
        static const bool log_time = std::getenv("IDP_ACCD_LOG") != nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        if (log_time) {
            IDP_HIP(hipEventCreate(&e0));
            IDP_HIP(hipEventCreate(&e1));
        }
        // a work queue needs no more lanes than the chip holds at once
        IDP_LAUNCH_TIMED(k_accd, 2048, 256, 0, s, e0, e1, cnt, cap, cand, A, 0.1, amin, cnt + 2, stepstat, cnt + 3,
                         lg, budget, w.accd_max_steps, capcnt, res_in, nres, res_out, cnt + 4, RES_CAP, pend);
        if (w.prof) w.prof->end(IDP_K_CCD, tok);
        if (log_time) {
            // kernel time against the longest pair's steps in this call
            unsigned long long st[4];
            float ms = 0.f;
            IDP_HIP(hipMemcpyAsync(st, stepstat, sizeof(st), hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
            IDP_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (st[1] >= 4096)
                std::fprintf(stderr,
                             "[idp] accd call: %.3f ms, longest pair %llu steps (%.1f ns per step), %llu steps, %llu "
                             "certified\n",
                             ms, st[1], 1e6 * ms / (double)st[1], st[0], st[3]);
            IDP_HIP(hipEventDestroy(e0));
            IDP_HIP(hipEventDestroy(e1));
        }
        IDP_HIP(hipMemcpyAsync(ha.data(), amin, B * sizeof(unsigned long long), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipMemcpyAsync(hsteps, stepstat, sizeof(hsteps), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipMemcpyAsync(hn, cnt, sizeof(hn), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        const size_t total = (size_t)hn[0] + hn[1];
        if (total <= cap) break;
        w.cap_cand = std::max<size_t>(w.cap_cand, total + total / 4 + 4096);
        if (attempt == 1) throw HipError(IDP_E_CAPACITY, "CCD candidate overflow");
    }
    clear_cells(w, s, T);
    static const bool log_long = std::getenv("IDP_ACCD_LOG") != nullptr;
    if (log_long && hn[3]) {
        // pairs whose ACCD iteration ran long: candidate, instance, initial
        // distance, motion bound, bound on entry, toc, steps, result, distance
        const int k = (int)std::min<unsigned>(hn[3], ACCD_LOG_MAX);
        std::vector<double> e(ACCD_LOG_W * k);
        IDP_HIP(hipMemcpyAsync(e.data(), lg, e.size() * sizeof(double), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        for (int j = 0; j < k; ++j) {
            const double* q = e.data() + ACCD_LOG_W * j;
            std::fprintf(stderr,
                         "[idp] long ACCD pair (%d,%d) inst %d d0 %.3e maxDisp %.3e tmax %.3e toc %.3e steps %.0f "
                         "result %.0f d %.3e\n",
                         (int)q[0], (int)q[1], (int)q[2], q[3], q[4], q[5], q[6], q[7], q[8], q[9]);
            std::fprintf(stderr, "[idp] long ACCD pair start");
            for (int v = 10; v < ACCD_LOG_W; ++v) std::fprintf(stderr, " %.17g", q[v]);
            std::fprintf(stderr, "\n");
        }
    }
    if (w.prof) {
        w.prof->add_work(IDP_K_CCD, (double)hn[0] + hn[1]);
        w.prof->ccd_steps += (double)hsteps[0];
        w.prof->ccd_max_steps = std::max(w.prof->ccd_max_steps, (double)hsteps[1]);
    }
    if (ncand_out) *ncand_out = (long)hn[0] + hn[1];
    w.last_capped = (long)hsteps[2];
    w.last_certified = (long)hsteps[3];
    w.h_capped.resize(B);
    if (hsteps[2]) {
        IDP_HIP(hipMemcpyAsync(w.h_capped.data(), capcnt, B * sizeof(unsigned), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
    } else {
        std::fill(w.h_capped.begin(), w.h_capped.end(), 0u);
    }
    std::vector<unsigned char> act(B, 1);
    if (active) {
        IDP_HIP(hipMemcpyAsync(act.data(), active, B, hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
    }
    std::vector<unsigned char> hp(B, 0);
    if (budget > 0) {
        IDP_HIP(hipMemcpyAsync(hp.data(), pend, B, hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
    }
    w.n_res = budget > 0 ? std::min(hn[4], RES_CAP) : 0u;
    w.res_cur ^= 1;
    // an instance's step is final unless one of its pairs was handed over;
    // resumed instances are those whose pairs came in (all of the last call's)
    for (int i = 0; i < B; ++i) {
        if (pending) pending[i] = hp[i];
        if ((act[i] || (nres && w.res_inst[i])) && !hp[i]) std::memcpy(&alpha_io[i], &ha[i], sizeof(double));
    }
    w.res_inst.assign(hp.begin(), hp.end());
}

double contact_ccd(ContactWork& w, hipStream_t s, const MeshDev& m, const double* X, const double* P,
                   double thickness, double alpha_in, long* ncand_out)
{
    double a = alpha_in;
    contact_ccd_batch(w, s, m, X, P, thickness, &a, nullptr, ncand_out, 0, false, nullptr);
    return a;
}

}  // namespace idp
