// Assembly of the projected Hessian into 3x3-block BSR and its solve.
//   triplets -> CSR, += M     Library/Math/CSR_MATRIX.h:49-56, INC_POTENTIAL.h:382-386
//   DBC projection            Library/Math/CSR_MATRIX.h:130-141
//   Newton direction solve    Library/Math/DIRECT_SOLVER.h:14-89 (replaced by PCG)
//
// The system is H = E + C: E has the static tet pattern (plus the lumped
// mass on the diagonal) and is refilled every Newton iteration by a gather
// over precomputed per-block contribution lists; C holds the contact blocks
// of the current constraint set, built by radix-sorting (row,col) keys of all
// local blocks and reducing each run in sort order.  Every sum has a fixed
// order, so results are bitwise reproducible.  PCG uses the 3x3 block-Jacobi
// preconditioner; its dot products are deterministic two-level reductions
// whose final sum is recomputed by each consumer block.
#include "common.hpp"
#include "kernels.hpp"
#include "linalg.hpp"

#include <hip/hip_cooperative_groups.h>
#include <hipcub/hipcub.hpp>

namespace idp {

// ---------------------------------------------------------------------------
// elastic BSR values and gradient gather
// ---------------------------------------------------------------------------
__global__ void k_elastic_values(long nnzb, const int* __restrict__ rowof, const int* __restrict__ col,
                                 const int* __restrict__ sptr, const int* __restrict__ src,
                                 const double* __restrict__ Hl, const double* __restrict__ mass,
                                 double* __restrict__ val)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    for (int j = sptr[k]; j < sptr[k + 1]; ++j) {
        const double* h = Hl + 9 * (size_t)src[j];
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += h[i];
    }
    int r = rowof[k];
    if (r == col[k]) {
        double m = mass[r];
        acc[0] += m;
        acc[4] += m;
        acc[8] += m;
    }
    double* o = val + 9 * (size_t)k;
#pragma unroll
    for (int i = 0; i < 9; ++i) o[i] = acc[i];
}

// g_v = elastic (sum over incident tets) + M (x - x~) + barrier (sum over the
// contact diagonal run of v); DBC rows zeroed; rhs = -g
__global__ void k_gradient_gather(int nV, const int* __restrict__ ediag, const int* __restrict__ sptr,
                                  const int* __restrict__ src, const double* __restrict__ gl,
                                  const double* __restrict__ X, const double* __restrict__ Xt,
                                  const double* __restrict__ mass, const int* __restrict__ cdiag,
                                  const unsigned* __restrict__ cseg, const unsigned* __restrict__ cval,
                                  const double* __restrict__ gb, const unsigned char* __restrict__ dbc,
                                  double* __restrict__ g, double* __restrict__ rhs)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    if (gl) {
        int k = ediag[v];
        for (int j = sptr[k]; j < sptr[k + 1]; ++j) {
            int s = src[j];
            int t = s >> 4, a = (s >> 2) & 3;
            const double* h = gl + 12 * (size_t)t + 3 * a;
            a0 += h[0];
            a1 += h[1];
            a2 += h[2];
        }
    }
    if (Xt) {
        double m = mass[v];
        a0 += m * (X[3 * v] - Xt[3 * v]);
        a1 += m * (X[3 * v + 1] - Xt[3 * v + 1]);
        a2 += m * (X[3 * v + 2] - Xt[3 * v + 2]);
    }
    if (cdiag) {
        int cb = cdiag[v];
        if (cb >= 0)
            for (unsigned j = cseg[cb]; j < cseg[cb + 1]; ++j) {
                unsigned e = cval[j];  // constraint*16 + a*nv + a  (nv packed in bits 28..)
                unsigned i = (e & 0x0FFFFFFFu) >> 4, slot = e >> 28;
                const double* h = gb + 12 * (size_t)i + 3 * slot;
                a0 += h[0];
                a1 += h[1];
                a2 += h[2];
            }
    }
    if (dbc && dbc[v]) a0 = a1 = a2 = 0.0;
    if (g) {
        g[3 * v] = a0;
        g[3 * v + 1] = a1;
        g[3 * v + 2] = a2;
    }
    if (rhs) {
        rhs[3 * v] = -a0;
        rhs[3 * v + 1] = -a1;
        rhs[3 * v + 2] = -a2;
    }
}

// ---------------------------------------------------------------------------
// contact blocks: keys, sort, run heads, reduction
// ---------------------------------------------------------------------------
__global__ void k_contact_nv2(long n, const int* __restrict__ nv, unsigned* __restrict__ cnt)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    cnt[i] = (unsigned)(nv[i] * nv[i]);
}

// record value: (slot_row << 28) | (constraint*16 + a*nv + b); slot_row = a
__global__ void k_contact_keys(long n, const int* __restrict__ vid, const int* __restrict__ nv,
                               const unsigned* __restrict__ off, unsigned long long V,
                               unsigned long long* __restrict__ keys, unsigned* __restrict__ vals)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int k = nv[i];
    unsigned o = off[i];
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            keys[o] = (unsigned long long)vid[4 * i + a] * V + (unsigned long long)vid[4 * i + b];
            vals[o] = ((unsigned)a << 28) | (unsigned)(i * 16 + a * k + b);
            ++o;
        }
}

__global__ void k_run_heads(unsigned n, const unsigned long long* __restrict__ keys, unsigned* __restrict__ head)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    head[i] = (i < n && (i == 0 || keys[i] != keys[i - 1])) ? 1u : 0u;
}

__global__ void k_contact_pattern(unsigned n, const unsigned long long* __restrict__ keys,
                                  const unsigned* __restrict__ hs, unsigned long long V, unsigned* __restrict__ seg,
                                  int* __restrict__ crow, int* __restrict__ ccol, int* __restrict__ cdiag)
{
    unsigned i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (!(i == 0 || keys[i] != keys[i - 1])) return;
    unsigned b = hs[i];
    seg[b] = i;
    int r = (int)(keys[i] / V), c = (int)(keys[i] % V);
    crow[b] = r;
    ccol[b] = c;
    if (r == c) cdiag[r] = (int)b;
}

__global__ void k_rowptr_from_rows(unsigned nb, const int* __restrict__ row, int nV, int* __restrict__ rowptr)
{
    // rowptr[r] = first block with row >= r (rows sorted)
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r > nV) return;
    unsigned lo = 0, hi = nb;
    while (lo < hi) {
        unsigned mid = (lo + hi) >> 1;
        if (row[mid] < r) lo = mid + 1;
        else hi = mid;
    }
    rowptr[r] = (int)lo;
}

__global__ void k_contact_values(unsigned nb, const unsigned* __restrict__ seg, const unsigned* __restrict__ vals,
                                 const double* __restrict__ Hb, double* __restrict__ cval)
{
    unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    for (unsigned j = seg[b]; j < seg[b + 1]; ++j) {
        const double* h = Hb + 9 * (size_t)(vals[j] & 0x0FFFFFFFu);
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += h[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) cval[9 * (size_t)b + i] = acc[i];
}

// DBC projection (CSR_MATRIX.h:130-141): blocks touching a DBC node -> delta_ij
__global__ void k_project_dbc(long nnzb, const int* __restrict__ row, const int* __restrict__ col,
                              const unsigned char* __restrict__ dbc, int keep_diag_identity, double* __restrict__ val)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    int r = row[k], c = col[k];
    if (!(dbc[r] || dbc[c])) return;
    double d = (r == c && keep_diag_identity) ? 1.0 : 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) val[9 * (size_t)k + i] = (i % 4 == 0) ? d : 0.0;
}

// ---------------------------------------------------------------------------
// PCG kernels
// ---------------------------------------------------------------------------
// block-Jacobi: inverse of E_vv + C_vv
__global__ void k_block_jacobi(int nV, const int* __restrict__ ediag, const double* __restrict__ evals,
                               const int* __restrict__ cdiag, const double* __restrict__ cvals,
                               double* __restrict__ Pinv, int* __restrict__ err)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double a[9];
    const double* e = evals + 9 * (size_t)ediag[v];
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = e[i];
    if (cdiag) {
        int c = cdiag[v];
        if (c >= 0)
#pragma unroll
            for (int i = 0; i < 9; ++i) a[i] += cvals[9 * (size_t)c + i];
    }
    double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8], c2 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
    double* o = Pinv + 9 * (size_t)v;
    if (!(det > 0)) {
        atomicOr(err, 1);
        det = 1.0;
    }
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
}

__device__ inline void row_spmv(int v, const int* __restrict__ rp, const int* __restrict__ col,
                                const double* __restrict__ val, const double* __restrict__ x, double& y0,
                                double& y1, double& y2)
{
    for (int k = rp[v]; k < rp[v + 1]; ++k) {
        const double* b = val + 9 * (size_t)k;
        const double* xc = x + 3 * (size_t)col[k];
        const double x0 = xc[0], x1 = xc[1], x2 = xc[2];
        y0 += b[0] * x0 + b[1] * x1 + b[2] * x2;
        y1 += b[3] * x0 + b[4] * x1 + b[5] * x2;
        y2 += b[6] * x0 + b[7] * x1 + b[8] * x2;
    }
}

// q = (E + C) p, partial sums of p.q
__global__ void __launch_bounds__(RB) k_spmv_dot(int nV, const int* __restrict__ erp, const int* __restrict__ ecol,
                                                 const double* __restrict__ eval, const int* __restrict__ crp,
                                                 const int* __restrict__ ccol, const double* __restrict__ cval,
                                                 const double* __restrict__ p, double* __restrict__ q,
                                                 double* __restrict__ part)
{
    __shared__ double sh[RB];
    int v = blockIdx.x * RB + threadIdx.x;
    double d = 0.0;
    if (v < nV) {
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        row_spmv(v, erp, ecol, eval, p, y0, y1, y2);
        if (crp) row_spmv(v, crp, ccol, cval, p, y0, y1, y2);
        q[3 * v] = y0;
        q[3 * v + 1] = y1;
        q[3 * v + 2] = y2;
        d = p[3 * v] * y0 + p[3 * v + 1] * y1 + p[3 * v + 2] * y2;
    }
    double s = block_sum<RB>(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// scalars layout: [0] rz (even iters) [1] rz (odd) [2] rr [3] breakdown flag
// alpha = rz / (p.q); x += alpha p; r -= alpha q; z = P r; partials of r.z, r.r
__global__ void __launch_bounds__(RB) k_pcg_update(int nV, int npart, const double* __restrict__ part_pq,
                                                   const double* __restrict__ scal, int par,
                                                   const double* __restrict__ p, const double* __restrict__ q,
                                                   const double* __restrict__ Pinv, double* __restrict__ x,
                                                   double* __restrict__ r, double* __restrict__ z,
                                                   double* __restrict__ part_rz, double* __restrict__ part_rr)
{
    __shared__ double sh[RB];
    const double pq = sum_partials_block(part_pq, npart, sh);
    const double rz = scal[par];
    const double alpha = (pq > 0) ? rz / pq : 0.0;
    int v = blockIdx.x * RB + threadIdx.x;
    double drz = 0.0, drr = 0.0;
    if (v < nV) {
        double rv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3 * v + i] += alpha * p[3 * v + i];
            rv[i] = r[3 * v + i] - alpha * q[3 * v + i];
            r[3 * v + i] = rv[i];
        }
        const double* P = Pinv + 9 * (size_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
            z[3 * v + i] = zi;
            drz += rv[i] * zi;
            drr += rv[i] * rv[i];
        }
    }
    double a = block_sum<RB>(drz, sh);
    double b = block_sum<RB>(drr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        part_rr[blockIdx.x] = b;
    }
}

// beta = rz_new / rz_old; p = z + beta p; block 0 stores rz_new, rr and the
// breakdown flag (p.q <= 0)
__global__ void __launch_bounds__(RB) k_pcg_direction(int nV, int npart, const double* __restrict__ part_rz,
                                                      const double* __restrict__ part_rr,
                                                      const double* __restrict__ part_pq, double* __restrict__ scal,
                                                      int par, const double* __restrict__ z, double* __restrict__ p)
{
    __shared__ double sh[RB];
    const double rzn = sum_partials_block(part_rz, npart, sh);
    const double rzo = scal[par];
    const double beta = rzo != 0.0 ? rzn / rzo : 0.0;
    int v = blockIdx.x * RB + threadIdx.x;
    if (v < nV)
#pragma unroll
        for (int i = 0; i < 3; ++i) p[3 * v + i] = z[3 * v + i] + beta * p[3 * v + i];
    if (blockIdx.x == 0) {
        const double rr = sum_partials_block(part_rr, npart, sh);
        const double pq = sum_partials_block(part_pq, npart, sh);
        if (threadIdx.x == 0) {
            scal[par ^ 1] = rzn;
            scal[2] = rr;
            if (!(pq > 0)) scal[3] = 1.0;
        }
    }
}

// initial: x = 0, r = b, z = P b, p = z, partials of r.z and r.r
__global__ void __launch_bounds__(RB) k_pcg_init(int nV, const double* __restrict__ b,
                                                 const double* __restrict__ Pinv, double* __restrict__ x,
                                                 double* __restrict__ r, double* __restrict__ z,
                                                 double* __restrict__ p, double* __restrict__ part_rz,
                                                 double* __restrict__ part_rr)
{
    __shared__ double sh[RB];
    int v = blockIdx.x * RB + threadIdx.x;
    double drz = 0.0, drr = 0.0;
    if (v < nV) {
        double rv[3] = {b[3 * v], b[3 * v + 1], b[3 * v + 2]};
        const double* P = Pinv + 9 * (size_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            x[3 * v + i] = 0.0;
            r[3 * v + i] = rv[i];
            double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
            z[3 * v + i] = zi;
            p[3 * v + i] = zi;
            drz += rv[i] * zi;
            drr += rv[i] * rv[i];
        }
    }
    double a = block_sum<RB>(drz, sh);
    double c = block_sum<RB>(drr, sh);
    if (threadIdx.x == 0) {
        part_rz[blockIdx.x] = a;
        part_rr[blockIdx.x] = c;
    }
}

__global__ void __launch_bounds__(RB) k_pcg_init_scalars(int npart, const double* __restrict__ part_rz,
                                                         const double* __restrict__ part_rr,
                                                         double* __restrict__ scal)
{
    __shared__ double sh[RB];
    const double rz = sum_partials_block(part_rz, npart, sh);
    const double rr = sum_partials_block(part_rr, npart, sh);
    if (threadIdx.x == 0) {
        scal[0] = rz;
        scal[1] = rz;
        scal[2] = rr;
        scal[3] = 0.0;
        scal[4] = rr;  // |b|^2
    }
}


// ---------------------------------------------------------------------------
// Persistent PCG: the whole iteration loop in one cooperative launch.
// Two grid barriers per iteration: the SpMV of the new direction reads
// p_new_j = z_j + beta p_old_j of neighbouring rows on the fly (p ping-pongs
// between two buffers), so q = A p_new is exact and no residual drift builds
// up.  Every block re-reduces the per-block partials in the same order, so
// all blocks take identical, deterministic decisions.
// ---------------------------------------------------------------------------
struct PcgArgs {
    int nV;
    const int* erp;
    const int* ecol;
    const double* eval;
    const int* crp;  // null when there are no contact blocks
    const int* ccol;
    const double* cval;
    const double* b;
    const double* Pinv;
    double* x;
    double* r;
    double* z;
    double* p0;
    double* p1;
    double* q;
    double* part_a;
    double* part_b;
    double* part_c;
    double* out;  // [0] iterations, [1] |r|^2, [2] |b|^2, [3] breakdown
    double rtol2;
    int maxit;
};

__device__ inline void row_spmv_dir(int v, const int* __restrict__ rp, const int* __restrict__ col,
                                    const double* __restrict__ val, const double* __restrict__ z,
                                    const double* __restrict__ po, double beta, double& y0, double& y1, double& y2)
{
    for (int k = rp[v]; k < rp[v + 1]; ++k) {
        const double* bk = val + 9 * (size_t)k;
        const size_t c = 3 * (size_t)col[k];
        const double x0 = z[c] + beta * po[c], x1 = z[c + 1] + beta * po[c + 1], x2 = z[c + 2] + beta * po[c + 2];
        y0 += bk[0] * x0 + bk[1] * x1 + bk[2] * x2;
        y1 += bk[3] * x0 + bk[4] * x1 + bk[5] * x2;
        y2 += bk[6] * x0 + bk[7] * x1 + bk[8] * x2;
    }
}

__global__ void __launch_bounds__(RB) k_pcg_persistent(PcgArgs A)
{
    namespace cgr = cooperative_groups;
    cgr::grid_group grid = cgr::this_grid();
    __shared__ double sh[RB];
    const int nb = gridDim.x;
    const int stride = nb * RB;
    // init: x = 0, r = b, z = P r, p_old = 0 (so p = z with beta = 0)
    double drz = 0.0, drr = 0.0;
    for (int v = blockIdx.x * RB + threadIdx.x; v < A.nV; v += stride) {
        double rv[3] = {A.b[3 * v], A.b[3 * v + 1], A.b[3 * v + 2]};
        const double* P = A.Pinv + 9 * (size_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A.x[3 * v + i] = 0.0;
            A.r[3 * v + i] = rv[i];
            A.p0[3 * v + i] = 0.0;
            double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
            A.z[3 * v + i] = zi;
            drz += rv[i] * zi;
            drr += rv[i] * rv[i];
        }
    }
    {
        double a = block_sum<RB>(drz, sh), c = block_sum<RB>(drr, sh);
        if (threadIdx.x == 0) {
            A.part_b[blockIdx.x] = a;
            A.part_c[blockIdx.x] = c;
        }
    }
    grid.sync();
    double rz = sum_partials_block(A.part_b, nb, sh);
    const double bb = sum_partials_block(A.part_c, nb, sh);
    const double stop2 = A.rtol2 * bb;
    double beta = 0.0, rr = bb;
    int it = 0, status = 0;
    double* po = A.p0;
    double* pn = A.p1;
    if (bb > 0.0) {
        for (; it < A.maxit; ++it) {
            // phase 1: p_new = z + beta p_old ; q = A p_new ; partial p.q
            double dpq = 0.0;
            for (int v = blockIdx.x * RB + threadIdx.x; v < A.nV; v += stride) {
                double y0 = 0.0, y1 = 0.0, y2 = 0.0;
                row_spmv_dir(v, A.erp, A.ecol, A.eval, A.z, po, beta, y0, y1, y2);
                if (A.crp) row_spmv_dir(v, A.crp, A.ccol, A.cval, A.z, po, beta, y0, y1, y2);
                const double p0v = A.z[3 * v] + beta * po[3 * v], p1v = A.z[3 * v + 1] + beta * po[3 * v + 1],
                             p2v = A.z[3 * v + 2] + beta * po[3 * v + 2];
                pn[3 * v] = p0v;
                pn[3 * v + 1] = p1v;
                pn[3 * v + 2] = p2v;
                A.q[3 * v] = y0;
                A.q[3 * v + 1] = y1;
                A.q[3 * v + 2] = y2;
                dpq += p0v * y0 + p1v * y1 + p2v * y2;
            }
            {
                double a = block_sum<RB>(dpq, sh);
                if (threadIdx.x == 0) A.part_a[blockIdx.x] = a;
            }
            grid.sync();
            const double pq = sum_partials_block(A.part_a, nb, sh);
            if (!(pq > 0.0)) {
                status = 1;
                break;
            }
            const double alpha = rz / pq;
            // phase 2: x += alpha p; r -= alpha q; z = P r; partials r.z, r.r
            drz = 0.0;
            drr = 0.0;
            for (int v = blockIdx.x * RB + threadIdx.x; v < A.nV; v += stride) {
                double rv[3];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    A.x[3 * v + i] += alpha * pn[3 * v + i];
                    rv[i] = A.r[3 * v + i] - alpha * A.q[3 * v + i];
                    A.r[3 * v + i] = rv[i];
                }
                const double* P = A.Pinv + 9 * (size_t)v;
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
                    A.z[3 * v + i] = zi;
                    drz += rv[i] * zi;
                    drr += rv[i] * rv[i];
                }
            }
            {
                double a = block_sum<RB>(drz, sh), c = block_sum<RB>(drr, sh);
                if (threadIdx.x == 0) {
                    A.part_b[blockIdx.x] = a;
                    A.part_c[blockIdx.x] = c;
                }
            }
            grid.sync();
            const double rzn = sum_partials_block(A.part_b, nb, sh);
            rr = sum_partials_block(A.part_c, nb, sh);
            beta = rzn / rz;
            rz = rzn;
            double* t = po;
            po = pn;
            pn = t;
            if (rr <= stop2) {
                ++it;
                break;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        A.out[0] = (double)it;
        A.out[1] = rr;
        A.out[2] = bb;
        A.out[3] = (double)status;
    }
}

// generic deterministic sum of partials -> out[0]
__global__ void __launch_bounds__(RB) k_sum_partials(int npart, const double* __restrict__ part,
                                                     double* __restrict__ out)
{
    __shared__ double sh[RB];
    const double s = sum_partials_block(part, npart, sh);
    if (threadIdx.x == 0) out[0] = s;
}

// x = a + s * b
__global__ void k_axpy3(long n, const double* __restrict__ a, double s, const double* __restrict__ b,
                        double* __restrict__ x)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    x[i] = a[i] + s * b[i];
}

// v = (x - xn) / h  (IMPLICIT_EULER.h:882-888)
__global__ void k_velocity(long n, const double* __restrict__ x, const double* __restrict__ xn, double h,
                           double* __restrict__ v)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    v[i] = (x[i] - xn[i]) / h;
}

// partial sums of a.b
__global__ void __launch_bounds__(RB) k_dot_parts(long n, const double* __restrict__ a,
                                                  const double* __restrict__ b, double* __restrict__ part)
{
    __shared__ double sh[RB];
    long i = (long)blockIdx.x * RB + threadIdx.x;
    double d = (i < n) ? a[i] * b[i] : 0.0;
    double s = block_sum<RB>(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {
template <class K, class V>
void sort_pairs(hipStream_t s, DBuf<unsigned char>& tmp, const K* kin, K* kout, const V* vin, V* vout, long n,
                int end_bit)
{
    size_t bytes = 0;
    IDP_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(hipcub::DeviceRadixSort::SortPairs(tmp.p, bytes, kin, kout, vin, vout, (int)n, 0, end_bit, s));
}
void excl_scan_u(hipStream_t s, DBuf<unsigned char>& tmp, const unsigned* in, unsigned* out, int n)
{
    size_t bytes = 0;
    IDP_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, in, out, n, s));
    tmp.ensure(bytes + 16, s);
    IDP_HIP(hipcub::DeviceScan::ExclusiveSum(tmp.p, bytes, in, out, n, s));
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
}  // namespace

void contact_structure(LinWork& L, hipStream_t s, int nV, long ncs, const int* cs)
{
    L.ncb = 0;
    L.nrec = 0;
    if (ncs == 0) return;
    int* vid = L.vid.ensure(4 * (size_t)ncs, s);
    int* nv = L.nv.ensure(ncs, s);
    launch_constraint_vids(s, ncs, cs, vid, nv);
    unsigned* cnt = L.u0.ensure(ncs + 1, s);
    unsigned* off = L.u1.ensure(ncs + 1, s);
    IDP_LAUNCH(k_contact_nv2, grid_for(ncs, 256), 256, 0, s, ncs, nv, cnt);
    IDP_HIP(hipMemsetAsync(cnt + ncs, 0, sizeof(unsigned), s));
    excl_scan_u(s, L.tmp, cnt, off, (int)ncs + 1);
    unsigned nrec = read_u(s, off + ncs);
    unsigned long long V = (unsigned long long)nV;
    unsigned long long* k0 = L.key0.ensure(nrec, s);
    unsigned* v0 = L.rv0.ensure(nrec, s);
    IDP_LAUNCH(k_contact_keys, grid_for(ncs, 256), 256, 0, s, ncs, vid, nv, off, V, k0, v0);
    unsigned long long* k1 = L.key1.ensure(nrec, s);
    unsigned* v1 = L.rv1.ensure(nrec, s);
    sort_pairs(s, L.tmp, k0, k1, v0, v1, nrec, bits_for(V * V));
    unsigned* head = L.u0.ensure(nrec + 1, s);
    unsigned* hs = L.u1.ensure(nrec + 1, s);
    IDP_LAUNCH(k_run_heads, grid_for(nrec + 1, 256), 256, 0, s, nrec, k1, head);
    excl_scan_u(s, L.tmp, head, hs, (int)nrec + 1);
    unsigned nb = read_u(s, hs + nrec);
    unsigned* seg = L.seg.ensure(nb + 1, s);
    int* crow = L.crow.ensure(nb + 1, s);
    int* ccol = L.ccol.ensure(nb + 1, s);
    int* cdiag = L.cdiag.ensure(nV, s);
    IDP_HIP(hipMemsetAsync(cdiag, 0xff, sizeof(int) * nV, s));
    IDP_LAUNCH(k_contact_pattern, grid_for(nrec, 256), 256, 0, s, nrec, k1, hs, V, seg, crow, ccol, cdiag);
    IDP_HIP(hipMemcpyAsync(seg + nb, &nrec, sizeof(unsigned), hipMemcpyHostToDevice, s));
    int* crp = L.crp.ensure(nV + 1, s);
    IDP_LAUNCH(k_rowptr_from_rows, grid_for(nV + 1, 256), 256, 0, s, nb, crow, nV, crp);
    L.ncb = nb;
    L.nrec = nrec;
    L.ncs = ncs;
    IDP_HIP(hipStreamSynchronize(s));
}

void contact_gradient_gather(LinWork& L, hipStream_t s, const MeshDev& m, const double* X, const double* Xt,
                             const double* gl, const double* gb, double* g, double* rhs)
{
    bool haveC = L.ncb > 0 && gb;
    IDP_LAUNCH(k_gradient_gather, grid_for(m.nV, 256), 256, 0, s, m.nV, m.e_diag.p, m.e_src_ptr.p, m.e_src.p, gl, X,
               Xt, m.mass.p, haveC ? L.cdiag.p : (const int*)nullptr, L.seg.p, L.rv1.p, gb, m.dbc.p, g, rhs);
}

void assemble(LinWork& L, hipStream_t s, const MeshDev& m, const double* Hl, const double* Hb, bool any_dbc)
{
    // algorithmic bytes: 80 per contribution (9 doubles + key), 76 per block, rowptr
    const double ncontrib = 16.0 * m.nT + (double)L.nrec + m.nV;
    const int tok = L.prof ? L.prof->begin(IDP_K_ASSEMBLY) : -1;
    double* ev = L.evals.ensure(9 * (size_t)m.nnzb_e, s);
    IDP_LAUNCH(k_elastic_values, grid_for(m.nnzb_e, 256), 256, 0, s, m.nnzb_e, L.erow.p, m.e_col.p, m.e_src_ptr.p,
               m.e_src.p, Hl, m.mass.p, ev);
    if (L.ncb) {
        double* cv = L.cvals.ensure(9 * (size_t)L.ncb, s);
        IDP_LAUNCH(k_contact_values, grid_for(L.ncb, 256), 256, 0, s, (unsigned)L.ncb, L.seg.p, L.rv1.p, Hb, cv);
    }
    if (any_dbc) {
        IDP_LAUNCH(k_project_dbc, grid_for(m.nnzb_e, 256), 256, 0, s, m.nnzb_e, L.erow.p, m.e_col.p, m.dbc.p, 1, ev);
        if (L.ncb)
            IDP_LAUNCH(k_project_dbc, grid_for(L.ncb, 256), 256, 0, s, (long)L.ncb, L.crow.p, L.ccol.p, m.dbc.p, 0,
                       L.cvals.p);
    }
    if (L.prof)
        L.prof->end(IDP_K_ASSEMBLY, tok, 80.0 * ncontrib + 76.0 * ((double)m.nnzb_e + L.ncb) + 4.0 * (m.nV + 1));
}

void init_erow(LinWork& L, hipStream_t s, const MeshDev& m, const std::vector<int>& rowof)
{
    L.erow.upload(rowof.data(), rowof.size(), s);
}

int pcg_solve(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol, int maxit,
              double* relres, int check_every)
{
    const int nV = m.nV;
    const int nblk = grid_for(nV, RB);
    double* Pinv = L.pinv.ensure(9 * (size_t)nV, s);
    double* r = L.r.ensure(3 * (size_t)nV, s);
    double* z = L.z.ensure(3 * (size_t)nV, s);
    double* p = L.p.ensure(3 * (size_t)nV, s);
    double* q = L.q.ensure(3 * (size_t)nV, s);
    double* ppq = L.part0.ensure(nblk, s);
    double* prz = L.part1.ensure(nblk, s);
    double* prr = L.part2.ensure(nblk, s);
    double* scal = L.scal.ensure(8, s);
    int* err = L.err.ensure(1, s);
    IDP_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    const bool haveC = L.ncb > 0;
    IDP_LAUNCH(k_block_jacobi, grid_for(nV, 256), 256, 0, s, nV, m.e_diag.p, L.evals.p,
               haveC ? L.cdiag.p : (const int*)nullptr, L.cvals.p, Pinv, err);
    IDP_LAUNCH(k_pcg_init, nblk, RB, 0, s, nV, b, Pinv, x, r, z, p, prz, prr);
    IDP_LAUNCH(k_pcg_init_scalars, 1, RB, 0, s, nblk, prz, prr, scal);
    double hs[5];
    IDP_HIP(hipMemcpyAsync(hs, scal, 5 * sizeof(double), hipMemcpyDeviceToHost, s));
    int herr = 0;
    IDP_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    if (herr) return -1;
    const double bn2 = hs[4];
    if (bn2 == 0.0) {
        *relres = 0.0;
        return 0;
    }
    const double stop2 = rtol * rtol * bn2;
    int it = 0;
    int par = 0;
    while (it < maxit) {
        IDP_LAUNCH(k_spmv_dot, nblk, RB, 0, s, nV, m.e_rowptr.p, m.e_col.p, L.evals.p,
                   haveC ? L.crp.p : (const int*)nullptr, L.ccol.p, L.cvals.p, p, q, ppq);
        IDP_LAUNCH(k_pcg_update, nblk, RB, 0, s, nV, nblk, ppq, scal, par, p, q, Pinv, x, r, z, prz, prr);
        IDP_LAUNCH(k_pcg_direction, nblk, RB, 0, s, nV, nblk, prz, prr, ppq, scal, par, z, p);
        par ^= 1;
        ++it;
        if (it % check_every == 0 || it == maxit) {
            IDP_HIP(hipMemcpyAsync(hs, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
            IDP_HIP(hipStreamSynchronize(s));
            if (hs[3] != 0.0) return -2;
            if (hs[2] <= stop2) break;
        }
    }
    IDP_HIP(hipMemcpyAsync(hs, scal, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    *relres = std::sqrt(hs[2] / bn2);
    return it;
}


int pcg_solve_persistent(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                         int maxit, double* relres)
{
    const int nV = m.nV;
    static int max_blocks = -1;
    if (max_blocks < 0) {
        int dev = 0, ncu = 0, per = 0;
        IDP_HIP(hipGetDevice(&dev));
        IDP_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        IDP_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pcg_persistent, RB, 0));
        max_blocks = std::max(1, ncu * std::max(per, 1));
    }
    const int nblk = std::max(1, std::min(max_blocks, grid_for(nV, RB)));
    double* Pinv = L.pinv.ensure(9 * (size_t)nV, s);
    int* err = L.err.ensure(1, s);
    IDP_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    const bool haveC = L.ncb > 0;
    IDP_LAUNCH(k_block_jacobi, grid_for(nV, 256), 256, 0, s, nV, m.e_diag.p, L.evals.p,
               haveC ? L.cdiag.p : (const int*)nullptr, L.cvals.p, Pinv, err);
    PcgArgs A;
    A.nV = nV;
    A.erp = m.e_rowptr.p;
    A.ecol = m.e_col.p;
    A.eval = L.evals.p;
    A.crp = haveC ? L.crp.p : nullptr;
    A.ccol = L.ccol.p;
    A.cval = L.cvals.p;
    A.b = b;
    A.Pinv = Pinv;
    A.x = x;
    A.r = L.r.ensure(3 * (size_t)nV, s);
    A.z = L.z.ensure(3 * (size_t)nV, s);
    A.p0 = L.p.ensure(3 * (size_t)nV, s);
    A.p1 = L.p1.ensure(3 * (size_t)nV, s);
    A.q = L.q.ensure(3 * (size_t)nV, s);
    A.part_a = L.part0.ensure(nblk, s);
    A.part_b = L.part1.ensure(nblk, s);
    A.part_c = L.part2.ensure(nblk, s);
    A.out = L.scal.ensure(8, s);
    A.rtol2 = rtol * rtol;
    A.maxit = maxit;
    void* args[] = {&A};
    const int tok = L.prof ? L.prof->begin(IDP_K_PCG) : -1;
    IDP_HIP(hipLaunchCooperativeKernel((const void*)k_pcg_persistent, dim3(nblk), dim3(RB), args, 0, s));
    if (L.prof) L.prof->end(IDP_K_PCG, tok);
    double hs[4];
    int herr = 0;
    IDP_HIP(hipMemcpyAsync(hs, A.out, 4 * sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    if (herr) return -1;
    if (hs[3] != 0.0) return -2;
    *relres = hs[2] > 0 ? std::sqrt(hs[1] / hs[2]) : 0.0;
    // algorithmic bytes per PCG iteration: BSR SpMV (76 nnzb + 4 (V+1) + 48 V),
    // block-Jacobi apply (72 V + 48 V), two dots and three axpys (8 * 24 V)
    if (L.prof) {
        const double nnzb = (double)m.nnzb_e + (double)L.ncb;
        L.prof->add_work(IDP_K_PCG, hs[0] * (76.0 * nnzb + 364.0 * nV + 4.0));
    }
    return (int)hs[0];
}

double dev_dot(LinWork& L, hipStream_t s, long n, const double* a, const double* b)
{
    int nblk = grid_for(n, RB);
    double* part = L.part3.ensure(nblk + 1, s);
    IDP_LAUNCH(k_dot_parts, nblk, RB, 0, s, n, a, b, part);
    IDP_LAUNCH(k_sum_partials, 1, RB, 0, s, nblk, part, part + nblk);
    double v;
    IDP_HIP(hipMemcpyAsync(&v, part + nblk, sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    return v;
}

void dev_axpy(hipStream_t s, long n, const double* a, double sc, const double* b, double* x)
{
    IDP_LAUNCH(k_axpy3, grid_for(n, 256), 256, 0, s, n, a, sc, b, x);
}

void dev_velocity(hipStream_t s, long n, const double* x, const double* xn, double h, double* v)
{
    IDP_LAUNCH(k_velocity, grid_for(n, 256), 256, 0, s, n, x, xn, h, v);
}

double dev_sum_partials(LinWork& L, hipStream_t s, int npart, const double* part)
{
    double* o = L.scal2.ensure(2, s);
    IDP_LAUNCH(k_sum_partials, 1, RB, 0, s, npart, part, o);
    double v;
    IDP_HIP(hipMemcpyAsync(&v, o, sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace idp
