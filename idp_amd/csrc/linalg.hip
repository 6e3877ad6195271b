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
#include "prim.hpp"
#include "kernels.hpp"
#include "linalg.hpp"


namespace idp {

// ---------------------------------------------------------------------------
// elastic BSR values and gradient gather
// ---------------------------------------------------------------------------
__global__ void k_elastic_values(long nnzb, const int* __restrict__ rowof, const int* __restrict__ col,
                                 const int* __restrict__ sptr, const int* __restrict__ src,
                                 const double* __restrict__ Hl, const double* __restrict__ mass,
                                 const int* __restrict__ vinst, const unsigned char* __restrict__ act,
                                 double* __restrict__ val)
{
    long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nnzb) return;
    if (act && !act[vinst[rowof[k]]]) return;  // instance not iterating
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
#pragma unroll
    for (int i = 0; i < 9; ++i) val[i * (size_t)nnzb + k] = acc[i];  // component planes
}

// g_v = elastic (sum over incident tets) + M (x - x~) + barrier (sum over the
// contact diagonal run of v); DBC rows zeroed; rhs = -g
__global__ void k_gradient_gather(int nV, const int* __restrict__ ediag, const int* __restrict__ sptr,
                                  const int* __restrict__ src, const double* __restrict__ gl,
                                  const double* __restrict__ X, const double* __restrict__ Xt,
                                  const double* __restrict__ mass, const int* __restrict__ cdiag,
                                  const unsigned* __restrict__ cseg, const unsigned* __restrict__ cval,
                                  const double* __restrict__ gb, const unsigned char* __restrict__ dbc,
                                  const int* __restrict__ vinst, const unsigned char* __restrict__ act,
                                  double* __restrict__ g, double* __restrict__ rhs)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    if (act && !act[vinst[v]]) return;  // instance not iterating
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
                                 const double* __restrict__ Hb, const int* __restrict__ crow,
                                 const int* __restrict__ vinst, const unsigned char* __restrict__ act,
                                 double* __restrict__ cval)
{
    unsigned b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    if (act && !act[vinst[crow[b]]]) return;  // instance not iterating
    double acc[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) acc[i] = 0.0;
    for (unsigned j = seg[b]; j < seg[b + 1]; ++j) {
        const double* h = Hb + 9 * (size_t)(vals[j] & 0x0FFFFFFFu);
#pragma unroll
        for (int i = 0; i < 9; ++i) acc[i] += h[i];
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) cval[i * (size_t)nb + b] = acc[i];  // component planes
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
    for (int i = 0; i < 9; ++i) val[i * (size_t)nnzb + k] = (i % 4 == 0) ? d : 0.0;
}

// ---------------------------------------------------------------------------
// PCG kernels
// ---------------------------------------------------------------------------
// block-Jacobi: inverse of E_vv + C_vv
__global__ void k_block_jacobi(int nV, const int* __restrict__ ediag, const double* __restrict__ evals, long eld,
                               const int* __restrict__ cdiag, const double* __restrict__ cvals, long cld,
                               const int* __restrict__ vinst, double* __restrict__ Pinv, int* __restrict__ err)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV) return;
    double a[9];
    const long ed = ediag[v];
#pragma unroll
    for (int i = 0; i < 9; ++i) a[i] = evals[i * eld + ed];
    if (cdiag) {
        int c = cdiag[v];
        if (c >= 0)
#pragma unroll
            for (int i = 0; i < 9; ++i) a[i] += cvals[i * cld + c];
    }
    double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8], c2 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
    double* o = Pinv + 9 * (size_t)v;
    if (!(det > 0)) {
        atomicOr(err + vinst[v], 1);
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

// ---------------------------------------------------------------------------
// Cluster block-Jacobi preconditioner.  Stiff barrier terms couple the
// vertices of a contact pair far more strongly than mass or elasticity do,
// and a 3x3 block-Jacobi leaves those couplings (and the elastic ones) to the
// iteration.  Vertices are grouped by two rounds of mutual heavy-edge
// matching over the off-diagonal blocks, elastic and contact (contact only
// with precond 2), into clusters of at most 4 vertices, never across
// instances; each cluster's (3k)^2 diagonal block of E + C is inverted
// exactly.  Vertex v stores the 3 x 3k rows of its cluster inverse as k 3x3
// blocks (plane b for member b), z_v = sum_b Z_vb r_mb, and the member list
// (ascending, -1 padded) as one int4, so the PCG reads both without a
// dependent load.
// ---------------------------------------------------------------------------
constexpr int CL_MAX = 4;
constexpr int CL_ZS = 9 * CL_MAX;  // doubles of preconditioner rows per vertex

__device__ inline double cblock_norm2(const double* __restrict__ val, long ld, int k)
{
    double w = 0.0;
#pragma unroll
    for (int e = 0; e < 9; ++e) w += val[e * ld + k] * val[e * ld + k];
    return w;
}

// vertices of the instances that solve (act null: all); every cluster lies
// in one instance, so a skipped vertex is never read by a kept one
struct ClMask {
    const int* vinst;
    const unsigned char* act;
    __device__ bool skip(int v) const { return act && !act[vinst[v]]; }
};

// elastic blocks taken into the matching as well (precond 0); rp null: contact blocks only
struct ClElastic {
    const int* rp;
    const int* col;
    const double* val;
    long ld;
};

// strongest contact neighbour of every vertex (ties: smaller id); -1 if none
__global__ void k_cl_best(int nV, const int* __restrict__ crp, const int* __restrict__ ccol,
                          const double* __restrict__ cval, long cld, ClElastic el, ClMask cm,
                          int* __restrict__ best)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || cm.skip(v)) return;
    int b = -1;
    double bw = 0.0;
    auto consider = [&](int u, double w) {
        if (u != v && (w > bw || (w == bw && b >= 0 && u < b))) {
            bw = w;
            b = u;
        }
    };
    if (crp)
        for (int k = crp[v]; k < crp[v + 1]; ++k) consider(ccol[k], cblock_norm2(cval, cld, k));
    if (el.rp)
        for (int k = el.rp[v]; k < el.rp[v + 1]; ++k) consider(el.col[k], cblock_norm2(el.val, el.ld, k));
    best[v] = b;
}

__global__ void k_cl_pair(int nV, const int* __restrict__ best, ClMask cm, int* __restrict__ root,
                          int* __restrict__ size)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || cm.skip(v)) return;
    const int b = best[v];
    const bool pair = b >= 0 && best[b] == v;
    root[v] = pair ? min(v, b) : v;
    size[v] = pair ? 2 : 1;  // meaningful at roots
}

// second matching pass: strongest neighbouring cluster of every cluster (max
// block norm over member-to-member blocks, contact and, with precond 0,
// elastic), packed (float weight, -root) so that a 64-bit max picks the
// heaviest, then the lowest root
__global__ void k_cl_best2(int nV, const int* __restrict__ crp, const int* __restrict__ ccol,
                           const double* __restrict__ cval, long cld, const int* __restrict__ root,
                           const int* __restrict__ size, ClElastic el, ClMask cm,
                           unsigned long long* __restrict__ pack)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || (!crp && !el.rp) || cm.skip(v)) return;
    const int r = root[v];
    unsigned long long bestp = 0;
    auto consider = [&](int u, double w) {
        const int ru = root[u];
        if (ru == r || size[r] + size[ru] > CL_MAX) return;
        const float wf = (float)w;
        const unsigned long long p =
            ((unsigned long long)__float_as_uint(wf) << 32) | (unsigned long long)(0xFFFFFFFFu - (unsigned)ru);
        if (p > bestp) bestp = p;
    };
    if (crp)
        for (int k = crp[v]; k < crp[v + 1]; ++k) consider(ccol[k], cblock_norm2(cval, cld, k));
    if (el.rp)
        for (int k = el.rp[v]; k < el.rp[v + 1]; ++k) consider(el.col[k], cblock_norm2(el.val, el.ld, k));
    if (bestp) atomicMax(pack + r, bestp);
}

__global__ void k_cl_merge(int nV, int* __restrict__ root, const unsigned long long* __restrict__ pack,
                           int* __restrict__ size, ClMask cm, int* __restrict__ root2)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || cm.skip(v)) return;
    const int r = root[v];
    int nr = r;
    const unsigned long long p = pack[r];
    if (p) {
        const int b = (int)(0xFFFFFFFFu - (unsigned)(p & 0xFFFFFFFFull));
        const unsigned long long q = pack[b];
        if (q && (int)(0xFFFFFFFFu - (unsigned)(q & 0xFFFFFFFFull)) == r) nr = min(r, b);
    }
    root2[v] = nr;
}

__global__ void k_cl_members(int nV, const int* __restrict__ root, ClMask cm, int* __restrict__ cnt,
                             int* __restrict__ mem)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || cm.skip(v)) return;
    const int r = root[v];
    const int slot = atomicAdd(cnt + r, 1);
    if (slot < CL_MAX) mem[CL_MAX * r + slot] = v;
}

// members of each cluster in ascending order
__global__ void k_cl_sort(int nV, const int* __restrict__ root, const int* __restrict__ cnt, ClMask cm,
                          int* __restrict__ mem)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || cm.skip(v)) return;
    const int r = root[v];
    const int k = min(cnt[r], CL_MAX);
    if (v == r) {
        int* m = mem + CL_MAX * r;
        for (int i = 1; i < k; ++i)
            for (int j = i; j > 0 && m[j - 1] > m[j]; --j) {
                const int t = m[j - 1];
                m[j - 1] = m[j];
                m[j] = t;
            }
    }
}

// locate block (a, b) in a BSR row; returns -1 when absent
__device__ inline int bsr_find(const int* __restrict__ rp, const int* __restrict__ col, int a, int b)
{
    for (int k = rp[a]; k < rp[a + 1]; ++k)
        if (col[k] == b) return k;
    return -1;
}

// one thread per cluster root: gather the (3k)^2 block of E + C, Cholesky,
// inverse, and scatter the member rows to Z.  Singletons keep the 3x3 inverse
// of the block-Jacobi preconditioner (same arithmetic).
__global__ void k_cl_build(int nV, const int* __restrict__ root, const int* __restrict__ cnt,
                           const int* __restrict__ mem, const int* __restrict__ ediag,
                           const int* __restrict__ erp, const int* __restrict__ ecol, const double* __restrict__ evals,
                           long eld, const int* __restrict__ cdiag, const int* __restrict__ crp,
                           const int* __restrict__ ccol, const double* __restrict__ cvals, long cld,
                           const int* __restrict__ vinst, ClMask cmask, double* __restrict__ Z,
                           int4* __restrict__ cm4, double* __restrict__ work, int* __restrict__ err)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= nV || cmask.skip(r) || root[r] != r) return;
    const int k = min(cnt[r], CL_MAX);
    const int* m = mem + CL_MAX * r;
    const int4 list = make_int4(m[0], k > 1 ? m[1] : -1, k > 2 ? m[2] : -1, k > 3 ? m[3] : -1);
    for (int a = 0; a < k; ++a) cm4[m[a]] = list;
    if (k == 1) {
        double a[9];
        const long ed = ediag[r];
#pragma unroll
        for (int i = 0; i < 9; ++i) a[i] = evals[i * eld + ed];
        if (cdiag) {
            const int c = cdiag[r];
            if (c >= 0)
#pragma unroll
                for (int i = 0; i < 9; ++i) a[i] += cvals[i * cld + c];
        }
        double c0 = a[4] * a[8] - a[5] * a[7], c1 = a[5] * a[6] - a[3] * a[8], c2 = a[3] * a[7] - a[4] * a[6];
        double det = a[0] * c0 + a[1] * c1 + a[2] * c2;
        double* o = Z + 9 * (size_t)r;  // plane 0
        if (!(det > 0)) {
            atomicOr(err + vinst[r], 1);
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
        return;
    }
    const int n = 3 * k;
    double* A = work + (size_t)2 * 144 * r;  // n x n, row-major
    double* Ainv = A + 144;
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            double blk[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const int ke = bsr_find(erp, ecol, m[a], m[b]);
            if (ke >= 0)
                for (int i = 0; i < 9; ++i) blk[i] += evals[i * eld + ke];
            if (crp) {
                const int kc = bsr_find(crp, ccol, m[a], m[b]);
                if (kc >= 0)
                    for (int i = 0; i < 9; ++i) blk[i] += cvals[i * cld + kc];
            }
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) A[(3 * a + i) * n + 3 * b + j] = blk[3 * i + j];
        }
    // symmetrise (the assembled blocks are symmetric up to rounding) and factor A = L L^T in place
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < i; ++j) {
            const double sym = 0.5 * (A[i * n + j] + A[j * n + i]);
            A[i * n + j] = sym;
            A[j * n + i] = sym;
        }
    bool ok = true;
    for (int j = 0; j < n && ok; ++j) {
        double d = A[j * n + j];
        for (int q = 0; q < j; ++q) d -= A[j * n + q] * A[j * n + q];
        if (!(d > 0)) {
            ok = false;
            break;
        }
        const double ljj = sqrt(d);
        A[j * n + j] = ljj;
        for (int i = j + 1; i < n; ++i) {
            double t = A[i * n + j];
            for (int q = 0; q < j; ++q) t -= A[i * n + q] * A[j * n + q];
            A[i * n + j] = t / ljj;
        }
    }
    if (!ok) {
        atomicOr(err + vinst[r], 1);
        for (int i = 0; i < n * n; ++i) Ainv[i] = (i % (n + 1)) == 0 ? 1.0 : 0.0;
    } else {
        // inverse column by column: solve L y = e_c, L^T x = y
        for (int c = 0; c < n; ++c) {
            double y[12];
            for (int i = 0; i < n; ++i) {
                double t = (i == c) ? 1.0 : 0.0;
                for (int q = 0; q < i; ++q) t -= A[i * n + q] * y[q];
                y[i] = t / A[i * n + i];
            }
            for (int i = n - 1; i >= 0; --i) {
                double t = y[i];
                for (int q = i + 1; q < n; ++q) t -= A[q * n + i] * Ainv[q * n + c];
                Ainv[i * n + c] = t / A[i * n + i];
            }
        }
    }
    // member a's rows, as k 3x3 blocks: block (a, b) in plane b at row m[a]
    for (int a = 0; a < k; ++a)
        for (int b = 0; b < k; ++b) {
            double* o = Z + 9 * ((size_t)b * nV + m[a]);
            for (int i = 0; i < 3; ++i)
                for (int j = 0; j < 3; ++j) o[3 * i + j] = Ainv[(3 * a + i) * n + 3 * b + j];
        }
}

// z_v = Z_v r_members (members' r given by rfun).  Z holds CL_MAX planes of
// 9 nV doubles: the 3x3 block of vertex v for its cluster's b-th member is
// at plane b, row v, so singletons (most vertices) read plane 0 only and a
// wavefront's reads of a plane are contiguous.
template <class RF>
__device__ inline void cl_apply(int4 cm, const double* __restrict__ Z, int nV, int v, const double* own, RF&& rfun,
                                double* zv)
{
    const double* o = Z + 9 * (size_t)v;
    if (cm.y < 0) {  // singleton: the 3x3 inverse of the diagonal block
#pragma unroll
        for (int i = 0; i < 3; ++i) zv[i] = o[3 * i] * own[0] + o[3 * i + 1] * own[1] + o[3 * i + 2] * own[2];
        return;
    }
    const int mb[CL_MAX] = {cm.x, cm.y, cm.z, cm.w};
    zv[0] = zv[1] = zv[2] = 0.0;
#pragma unroll
    for (int b = 0; b < CL_MAX; ++b) {
        if (mb[b] < 0) break;
        double rb[3];
        rfun(mb[b], rb);
        const double* ob = o + 9 * (size_t)b * nV;
#pragma unroll
        for (int i = 0; i < 3; ++i) zv[i] += ob[3 * i] * rb[0] + ob[3 * i + 1] * rb[1] + ob[3 * i + 2] * rb[2];
    }
}

// ---------------------------------------------------------------------------
// Block-Jacobi PCG as two kernels per iteration, enqueued in chunks with no
// host synchronisation inside a chunk.  A grid-wide barrier inside one
// persistent kernel costs a device-scope release/acquire per block (the
// eight XCD L2s are not coherent for ordinary device memory), which grows
// with the block count; a kernel boundary does that once.
//  * k_pcg_spmv(it): every block re-reduces the previous update's partials
//    (r.z, r.r) in a fixed order, decides convergence identically, forms
//    p_it = z + beta p_(it-1) on the fly and computes q = A p_it with a group
//    of PCG_G lanes per block row (xor-butterfly row reduction), plus
//    partial p.q.
//  * k_pcg_update(it): alpha = r.z / p.q; x += alpha p; r -= alpha q; z = P r;
//    partials of the new r.z and r.r.
// p and the r.z / r.r partials ping-pong on the iteration parity.  When the
// residual test holds (or on breakdown), the deciding kernel sets a flag and
// every later kernel returns at once, so the host may enqueue past the end.
// ---------------------------------------------------------------------------
struct PcgArgs {
    int nV;
    const int* erp;
    const int* ecol;
    const double* eval;  // 9 component planes of eld blocks
    const int* crp;      // null when there are no contact blocks
    const int* ccol;
    const double* cval;  // 9 component planes of cld blocks
    long eld, cld;
    const double* b;
    const double* Z;      // preconditioner rows: CL_MAX planes of 9 nV doubles (cl_apply)
    const int4* cm4;      // per vertex: its cluster's members, ascending, -1 padded
    double* x;
    double* r;        // 2 x 3V: the update of iteration it reads half it&1, writes the other
    double* z;
    double* pbuf;     // 2 x 3V
    double* q;
    double* part_pq;  // nbA
    double* part_bx;  // nbA: b.x of the warm-start pass
    const unsigned char* warm;  // per instance: start from gamma x (x holds the previous direction); null = none
    double* part_rz;  // 2 x nbB
    double* part_rr;  // 2 x nbB
    int nbA, nbB;
    // launch tables: one int4 {first row, end row, instance, block id} per
    // launched block, SpMV blocks of PCG_ROWS rows (atab) and update blocks of
    // 64 rows (btab), each aligned to its instance's first row.  Only the
    // blocks of instances still iterating are launched; partials are indexed
    // by the block id, so an instance's partial range [apbeg, bpbeg) is the
    // same whatever else is launched.
    const int4 *atab, *btab;
    const int *aoff, *boff;  // per launched block: iterations its instance did in earlier calls
    int ident;               // one instance, all blocks launched: entries follow from blockIdx (no table load)
    int nbl_upd;             // update blocks launched (upd_waves per workgroup)
    const int *apbeg, *bpbeg;
    const unsigned char* solve;  // per instance: 0 not solving, 1 new solve, 2 continued from an earlier call
    int* flag;        // per instance: 0 running, 1 converged, 2 breakdown, 3 not solving, 4 maxit reached
    double* scal;     // per instance, 3 each: |b|^2, final |r|^2, iterations
    double rtol2;
    int maxit;
    // one instance with many blocks: the sums of the partials are formed once
    // per kernel boundary by k_pcg_red_spmv / k_pcg_red_upd (r.z, r.r, the
    // previous r.z or |b|^2; p.q, r.z) instead of by every block (null)
    double* red;
    // coarse level (one instance): partials of r_c . e_c, cnsb per parity,
    // added to r.z wherever it is formed (null: none)
    const double* cpart;
    int cnsb;
};

constexpr int PCG_ROW_THREADS = 256;         // row threads of a SpMV block (4 wavefronts), 256 / G rows
constexpr int PCG_PREREDUCE_BLOCKS = 1024;   // update blocks of one instance from which partials are pre-reduced
constexpr int PCG_UPD_BLOCK = 64;            // rows of an update block: one wavefront, reductions need no barrier
constexpr int PCG_UPD_WAVES_MAX = 8;         // update blocks per workgroup (independent wavefronts), at most
// SpMV launch shapes: G lanes per block row, scalars in wavefront 0 (before
// its rows) or in a fifth wavefront of their own
struct SpmvShape {
    int g;
    bool sw;
};
constexpr SpmvShape SPMV_SHAPES[] = {{4, false}, {8, false}, {16, false}, {32, false},
                                     {4, true},  {8, true},  {16, true},  {32, true}};
inline int spmv_rows(int g) { return PCG_ROW_THREADS / g; }

// Launch-table entry of this block: {first row, end row, instance, block id}.
// A one-instance launch derives it from blockIdx, saving a dependent load.
template <int ROWS>
__device__ inline int4 launch_entry(const PcgArgs& A, const int4* __restrict__ tab, int blk)
{
    if (A.ident) {
        const int lo = blk * ROWS;
        return make_int4(lo, min(lo + ROWS, A.nV), 0, blk);
    }
    return tab[blk];
}
template <int ROWS>
__device__ inline int4 launch_entry(const PcgArgs& A, const int4* __restrict__ tab)
{
    return launch_entry<ROWS>(A, tab, (int)blockIdx.x);
}
__device__ inline int launch_off(const PcgArgs& A, const int* __restrict__ off, int blk)
{
    return A.ident ? A.ident - 1 : off[blk];
}
__device__ inline int launch_off(const PcgArgs& A, const int* __restrict__ off)
{
    return launch_off(A, off, (int)blockIdx.x);
}

// Sum over a wavefront by a fixed xor butterfly (every lane gets the total).
__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sum of n partials by one wavefront in a fixed order.  The loads of a chunk
// of NL per lane are issued together (independent), then added in order, so
// the latency is paid once per 64 NL partials rather than once per 64.
template <int NL = 16>
__device__ inline double wave_sum_partials(const double* __restrict__ part, int n)
{
    const int lane = threadIdx.x & 63;
    double acc = 0.0;
    for (int base = 0; base < n; base += 64 * NL) {
        double v[NL];
#pragma unroll
        for (int k = 0; k < NL; ++k) {
            const int i = base + lane + 64 * k;
            v[k] = i < n ? part[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < NL; ++k) acc += v[k];
    }
    return wave_sum(acc);
}

// Three sums of n partials each, every load issued before any add (one
// memory round trip for up to 512 partials per array); each sum is formed in
// the order wave_sum_partials uses, so the results are the same.
__device__ inline void wave_sum_partials3(const double* __restrict__ p0, const double* __restrict__ p1,
                                          const double* __restrict__ p2, int n, double& s0, double& s1,
                                          double& s2)
{
    const int lane = threadIdx.x & 63;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int base = 0; base < n; base += 64 * 8) {
        double v0[8], v1[8], v2[8];
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const int i = base + lane + 64 * k;
            const bool in = i < n;
            v0[k] = in ? p0[i] : 0.0;
            v1[k] = in ? p1[i] : 0.0;
            v2[k] = in ? p2[i] : 0.0;
        }
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            a0 += v0[k];
            a1 += v1[k];
            a2 += v2[k];
        }
    }
    s0 = wave_sum(a0);
    s1 = wave_sum(a1);
    s2 = wave_sum(a2);
}

// Two sums of partials (na and nb of them) with the first NA and NB loads
// per lane of both issued before any add; the per-lane order of the adds is
// that of wave_sum_partials, so the sums are the same.
template <int NA, int NB>
__device__ inline void wave_sum_partials2(const double* __restrict__ pa, int na, const double* __restrict__ pb,
                                          int nb, double& sa, double& sb)
{
    const int lane = threadIdx.x & 63;
    double va[NA], vb[NB];
#pragma unroll
    for (int k = 0; k < NA; ++k) va[k] = lane + 64 * k < na ? pa[lane + 64 * k] : 0.0;
#pragma unroll
    for (int k = 0; k < NB; ++k) vb[k] = lane + 64 * k < nb ? pb[lane + 64 * k] : 0.0;
    double a = 0.0, b = 0.0;
#pragma unroll
    for (int k = 0; k < NA; ++k) a += va[k];
#pragma unroll
    for (int k = 0; k < NB; ++k) b += vb[k];
    for (int i = lane + 64 * NA; i < na; i += 64) a += pa[i];
    for (int i = lane + 64 * NB; i < nb; i += 64) b += pb[i];
    sa = wave_sum(a);
    sb = wave_sum(b);
}

// y += A_row z (and w += A_row po when po is given) over this lane's blocks of
// row v: the row's elastic blocks followed by its contact blocks, lanes of
// a group taking slots lane, lane + G, ...  Rows of at most G blocks
// (most of them) take one slot per lane, so every load of a lane is issued at
// once.  Values are 9 component planes of ld doubles: the lanes of a group
// read consecutive addresses.
template <int G>
__device__ inline void row_spmv(int lane, int e0, int ne, int c0, int nc, const PcgArgs& A,
                                const double* __restrict__ z, const double* __restrict__ po, double* y, double* w)
{
    for (int j = lane; j < ne + nc; j += G) {
        const bool el = j < ne;
        const double* __restrict__ val = el ? A.eval : A.cval;
        const long ld = el ? A.eld : A.cld;
        const int k = el ? e0 + j : c0 + (j - ne);
        const int col = el ? A.ecol[k] : A.ccol[k];
        const double bk[9] = {val[k], val[ld + k], val[2 * ld + k], val[3 * ld + k], val[4 * ld + k],
                              val[5 * ld + k], val[6 * ld + k], val[7 * ld + k], val[8 * ld + k]};
        const size_t c = 3 * (size_t)col;
        const double x0 = z[c], x1 = z[c + 1], x2 = z[c + 2];
        y[0] += bk[0] * x0 + bk[1] * x1 + bk[2] * x2;
        y[1] += bk[3] * x0 + bk[4] * x1 + bk[5] * x2;
        y[2] += bk[6] * x0 + bk[7] * x1 + bk[8] * x2;
        if (po) {
            const double p0 = po[c], p1 = po[c + 1], p2 = po[c + 2];
            w[0] += bk[0] * p0 + bk[1] * p1 + bk[2] * p2;
            w[1] += bk[3] * p0 + bk[4] * p1 + bk[5] * p2;
            w[2] += bk[6] * p0 + bk[7] * p1 + bk[8] * p2;
        }
    }
}

// row extents of row v in both block sets
__device__ inline void row_extents(const PcgArgs& A, int v, int& e0, int& ne, int& c0, int& nc)
{
    e0 = A.erp[v];
    ne = A.erp[v + 1] - e0;
    c0 = A.crp ? A.crp[v] : 0;
    nc = A.crp ? A.crp[v + 1] - c0 : 0;
}

template <int G>
__device__ inline double row_butterfly(double* y, int lane)
{
#pragma unroll
    for (int o = G / 2; o > 0; o >>= 1) {
        y[0] += __shfl_xor(y[0], o, G);
        y[1] += __shfl_xor(y[1], o, G);
        y[2] += __shfl_xor(y[2], o, G);
    }
    // select by value (an index into y would put the array in scratch memory)
    const double y0 = y[0], y1 = y[1], y2 = y[2];
    double r = y2;
    r = lane == 1 ? y1 : r;
    r = lane == 0 ? y0 : r;
    return r;
}

// sum of one double per thread of the row threads -> one partial per block
__device__ inline double block_row_sum(double v, double* wsum)
{
    v = wave_sum(v);
    if (threadIdx.x < PCG_ROW_THREADS && (threadIdx.x & 63) == 0) wsum[threadIdx.x >> 6] = v;
    __syncthreads();
    double a = 0.0;
#pragma unroll
    for (int w2 = 0; w2 < PCG_ROW_THREADS / 64; ++w2) a += wsum[w2];
    return a;
}

// x0 = gamma x_prev (gamma = b.x / x.Ax, the A-norm-closest multiple of the
// previous direction) for warm instances, else 0; r0 = b - gamma A x_prev;
// z0 = P r0; partials of r.z, r.r and b.b (the last in the second half of
// the r.r partials, read by the first SpMV)
__global__ void __launch_bounds__(PCG_UPD_BLOCK) k_pcg_init(PcgArgs A)
{
    const int4 t = launch_entry<PCG_UPD_BLOCK>(A, A.btab);
    const int inst = t.z, bid = t.w;
    if (A.flag[inst] != 0 || (A.solve && A.solve[inst] != 1)) return;  // not solving, or continued
    const int v = t.x + threadIdx.x;
    double gamma = 0.0;
    if (A.warm && A.warm[inst]) {
        const int ab = A.apbeg[inst], na = A.apbeg[inst + 1] - ab;
        const double xq = wave_sum_partials<32>(A.part_pq + ab, na);
        const double bx = wave_sum_partials<32>(A.part_bx + ab, na);
        gamma = xq > 0.0 ? bx / xq : 0.0;
    }
    const bool w = gamma != 0.0;
    auto res = [&](int u, double* o) {
#pragma unroll
        for (int i = 0; i < 3; ++i) o[i] = w ? A.b[3 * u + i] - gamma * A.q[3 * u + i] : A.b[3 * u + i];
    };
    double drz = 0.0, drr = 0.0, dbb = 0.0;
    if (v < t.y) {
        double rv[3], zv[3];
        res(v, rv);
        cl_apply(A.cm4[v], A.Z, A.nV, v, rv, res, zv);
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double bi = A.b[3 * v + i];
            A.x[3 * v + i] = w ? gamma * A.x[3 * v + i] : 0.0;
            A.r[3 * v + i] = rv[i];
            A.z[3 * v + i] = zv[i];
            drz += rv[i] * zv[i];
            drr += rv[i] * rv[i];
            dbb += bi * bi;
        }
    }
    drz = wave_sum(drz);
    drr = wave_sum(drr);
    dbb = wave_sum(dbb);
    if (threadIdx.x == 0) {
        A.part_rz[bid] = drz;
        A.part_rr[bid] = drr;
        A.part_rr[A.nbB + bid] = dbb;
    }
}

// warm start: q = A x for the instances that start from their previous
// direction, with per-block partials of x.q and b.x
template <int G>
__global__ void __launch_bounds__(PCG_ROW_THREADS) k_pcg_warm(PcgArgs A)
{
    __shared__ double wsum[2][PCG_ROW_THREADS / 64];
    const int4 t = launch_entry<PCG_ROW_THREADS / G>(A, A.atab);
    const int inst = t.z;
    const bool on = A.warm[inst] && A.flag[inst] == 0;  // uniform over the block
    const int lane = threadIdx.x & (G - 1);
    const int v = t.x + threadIdx.x / G;
    const bool row = on && threadIdx.x < PCG_ROW_THREADS && v < t.y;
    double y[3] = {0.0, 0.0, 0.0}, w[3] = {0.0, 0.0, 0.0};
    double dxq = 0.0, dbx = 0.0;
    if (row) {
        int e0, ne, c0, nc;
        row_extents(A, v, e0, ne, c0, nc);
        row_spmv<G>(lane, e0, ne, c0, nc, A, A.x, nullptr, y, w);
        const double yl = row_butterfly<G>(y, lane);
        if (lane < 3) {
            const size_t i = 3 * (size_t)v + lane;
            const double xi = A.x[i];
            A.q[i] = yl;
            dxq = xi * yl;
            dbx = A.b[i] * xi;
        }
    }
    dxq = wave_sum(dxq);
    dbx = wave_sum(dbx);
    if (threadIdx.x < PCG_ROW_THREADS && (threadIdx.x & 63) == 0) {
        wsum[0][threadIdx.x >> 6] = dxq;
        wsum[1][threadIdx.x >> 6] = dbx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, c = 0.0;
#pragma unroll
        for (int w2 = 0; w2 < PCG_ROW_THREADS / 64; ++w2) {
            a += wsum[0][w2];
            c += wsum[1][w2];
        }
        A.part_pq[t.w] = a;
        A.part_bx[t.w] = c;
    }
}

// SpMV step of iteration it.  The scalar wavefront re-reduces the instance's
// partials of the previous update and decides convergence while the row
// wavefronts stream their blocks; p_it = z + beta p_(it-1) is never read:
// the rows accumulate A z and A p_(it-1) separately and combine them once
// beta is known, q = A z + beta A p_(it-1).
template <int G, bool SW>
__global__ void __launch_bounds__(PCG_ROW_THREADS + (SW ? 64 : 0)) k_pcg_spmv(PcgArgs A, int it_call)
{
    __shared__ double sbeta;
    __shared__ double wsum[PCG_ROW_THREADS / 64];
    __shared__ int sflag;
    const int it = it_call + launch_off(A, A.aoff);  // the instance's own iteration
    const int par = it & 1;
    const int4 t = launch_entry<PCG_ROW_THREADS / G>(A, A.atab);
    const int inst = t.z, bid = t.w;
    const double* po = it > 0 ? A.pbuf + (size_t)(par ^ 1) * 3 * A.nV : nullptr;
    double y[3] = {0.0, 0.0, 0.0}, w[3] = {0.0, 0.0, 0.0};
    double zo = 0.0, poo = 0.0;
    const int lane = threadIdx.x & (G - 1);
    const int v = t.x + threadIdx.x / G;
    bool row = false;
    auto scalars = [&]() {
        const int f = *(volatile int*)(A.flag + inst);
        const int pb = A.bpbeg[inst], np = A.bpbeg[inst + 1] - pb;
        // r.z and r.r of this iteration, and r.z of the previous one (|b|^2 at it = 0)
        double rz, rr, third;
        if (A.red) {
            rz = A.red[0];
            rr = A.red[1];
            third = A.red[2];
        } else {
            wave_sum_partials3(A.part_rz + par * A.nbB + pb, A.part_rr + par * A.nbB + pb,
                               it > 0 ? A.part_rz + (par ^ 1) * A.nbB + pb : A.part_rr + A.nbB + pb, np, rz, rr,
                               third);
        }
        if (A.cpart) {  // r . (z1 + P e_c) = r . z1 + r_c . e_c
            rz += wave_sum_partials<8>(A.cpart + par * A.cnsb, A.cnsb);
            if (it > 0) third += wave_sum_partials<8>(A.cpart + (par ^ 1) * A.cnsb, A.cnsb);
        }
        const double rzo = it > 0 ? third : 1.0;
        const double bbp = it == 0 ? third : 0.0;
        if ((threadIdx.x & 63) == 0) {
            double* sci = A.scal + 3 * inst;
            const double bb = it == 0 ? bbp : sci[0];
            int stop = f != 0 ? 2 : 0;
            const bool conv = !(bb > 0.0) || (it > 0 && rr <= A.rtol2 * bb);
            if (!stop && (conv || it >= A.maxit)) {
                stop = 1;
                if (bid == A.apbeg[inst]) {
                    A.flag[inst] = conv ? 1 : 4;  // 4: the iteration limit ended the solve
                    sci[0] = bb;
                    sci[1] = rr;
                    sci[2] = (double)it;
                }
            }
            if (!stop && it == 0 && bid == A.apbeg[inst]) sci[0] = bb;
            sflag = stop;
            sbeta = it > 0 ? rz / rzo : 0.0;
        }
    };
    const bool scalar_wave = SW ? threadIdx.x >= PCG_ROW_THREADS : threadIdx.x < 64;
    if (!SW && scalar_wave) scalars();
    if (threadIdx.x < PCG_ROW_THREADS && v < t.y) {
        // the flag is loaded with the row extents; a stopped instance skips its blocks
        const int f = *(volatile int*)(A.flag + inst);
        int e0, ne, c0, nc;
        row_extents(A, v, e0, ne, c0, nc);
        if (lane < 3) {
            zo = A.z[3 * v + lane];
            poo = po ? po[3 * v + lane] : 0.0;
        }
        row = f == 0;
        if (row) row_spmv<G>(lane, e0, ne, c0, nc, A, A.z, po, y, w);
    }
    if (SW && scalar_wave) scalars();
    __syncthreads();
    if (sflag != 0) return;  // block-uniform
    const double beta = sbeta;
    double dpq = 0.0;
    if (row) {
#pragma unroll
        for (int d = 0; d < 3; ++d) y[d] += beta * w[d];
        const double yl = row_butterfly<G>(y, lane);
        if (lane < 3) {
            const size_t i = 3 * (size_t)v + lane;
            const double pv = po ? zo + beta * poo : zo;
            A.pbuf[(size_t)par * 3 * A.nV + i] = pv;
            A.q[i] = yl;
            dpq = pv * yl;
        }
    }
    const double a = block_row_sum(dpq, wsum);
    if (threadIdx.x == 0) A.part_pq[bid] = a;
}

// Sums of n partials by one workgroup of PCG_RED threads in a fixed order:
// thread t adds i = t, t + PCG_RED, ... (loads issued PCG_RED_NL at a time),
// then the workgroup's sums are added in thread order.
constexpr int PCG_RED = 256;
constexpr int PCG_RED_NL = 8;

template <int K>
__device__ inline void block_sums_partials(const double* const* __restrict__ p, int n, double* out)
{
    __shared__ double sh[K][PCG_RED];
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int base = 0; base < n; base += PCG_RED * PCG_RED_NL) {
        double v[K][PCG_RED_NL];
#pragma unroll
        for (int j = 0; j < PCG_RED_NL; ++j) {
            const int i = base + (int)threadIdx.x + PCG_RED * j;
#pragma unroll
            for (int k = 0; k < K; ++k) v[k][j] = i < n ? p[k][i] : 0.0;
        }
#pragma unroll
        for (int j = 0; j < PCG_RED_NL; ++j)
#pragma unroll
            for (int k = 0; k < K; ++k) acc[k] += v[k][j];
    }
#pragma unroll
    for (int k = 0; k < K; ++k) sh[k][threadIdx.x] = acc[k];
    __syncthreads();
    if (threadIdx.x < K) {
        double a = 0.0;
        for (int t = 0; t < PCG_RED; ++t) a += sh[threadIdx.x][t];
        out[threadIdx.x] = a;
    }
}

// the sums the SpMV of iteration it needs (one instance spanning the launch)
__global__ void __launch_bounds__(PCG_RED) k_pcg_red_spmv(PcgArgs A, int it_call)
{
    const int it = it_call + A.ident - 1;
    const int par = it & 1;
    const int pb = A.bpbeg[0], np = A.bpbeg[1] - pb;
    const double* p[3] = {A.part_rz + par * A.nbB + pb, A.part_rr + par * A.nbB + pb,
                          it > 0 ? A.part_rz + (par ^ 1) * A.nbB + pb : A.part_rr + A.nbB + pb};
    block_sums_partials<3>(p, np, A.red);
}

// the sums the update of iteration it needs
__global__ void __launch_bounds__(PCG_RED) k_pcg_red_upd(PcgArgs A, int it_call)
{
    const int it = it_call + A.ident - 1;
    const int par = it & 1;
    const int pb = A.bpbeg[0];
    __shared__ double o[2];
    const double* pa[1] = {A.part_pq + A.apbeg[0]};
    const double* pr[1] = {A.part_rz + par * A.nbB + pb};
    block_sums_partials<1>(pa, A.apbeg[1] - A.apbeg[0], o);
    __syncthreads();
    block_sums_partials<1>(pr, A.bpbeg[1] - pb, o + 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        A.red[3] = o[0];
        A.red[4] = o[1];
    }
}

// Update step of iteration it: alpha = r.z / p.q; x += alpha p; r -= alpha q;
// z = P r (cluster members' residuals recomputed from the unwritten half of
// r); partials of the new r.z and r.r.  Every load that does not depend on
// alpha, the cluster's member list and the preconditioner rows included, is
// issued before the partials are reduced.
__global__ void __launch_bounds__(PCG_UPD_BLOCK * PCG_UPD_WAVES_MAX) k_pcg_update(PcgArgs A, int it_call)
{
    // each wavefront is one update block of 64 rows, independent of the
    // others in its workgroup
    const int blk = blockIdx.x * (int)(blockDim.x >> 6) + (int)(threadIdx.x >> 6);
    if (blk >= A.nbl_upd) return;  // uniform over the wavefront
    const int lane = threadIdx.x & 63;
    const int it = it_call + launch_off(A, A.boff, blk);  // the instance's own iteration
    const int4 t = launch_entry<PCG_UPD_BLOCK>(A, A.btab, blk);
    const int inst = t.z, bid = t.w;
    const int par = it & 1;
    const double* p = A.pbuf + (size_t)par * 3 * A.nV;
    const double* r_in = A.r + (size_t)par * 3 * A.nV;  // not written by this kernel: members read it safely
    double* r_out = A.r + (size_t)(par ^ 1) * 3 * A.nV;
    const int v = t.x + lane;
    const bool own = v < t.y;
    const int vc = own ? v : t.x;
    const int f = *(volatile int*)(A.flag + inst);
    double xv[3], pv[3], rv[3], qv[3], o0[9];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
        xv[i] = A.x[3 * vc + i];
        pv[i] = p[3 * vc + i];
        rv[i] = r_in[3 * vc + i];
        qv[i] = A.q[3 * vc + i];
    }
    const int4 cm = A.cm4[vc];
    const double* o = A.Z + 9 * (size_t)vc;  // plane 0
#pragma unroll
    for (int i = 0; i < 9; ++i) o0[i] = o[i];
    // the partials are read in the same round trip as the flag and the rows
    // (a stopped instance's sums are formed and dropped)
    const int pb = A.bpbeg[inst];
    double pq, rz;
    if (A.red) {
        pq = A.red[3];
        rz = A.red[4];
    } else {
        wave_sum_partials2<8, 4>(A.part_pq + A.apbeg[inst], A.apbeg[inst + 1] - A.apbeg[inst],
                                  A.part_rz + par * A.nbB + pb, A.bpbeg[inst + 1] - pb, pq, rz);
    }
    if (A.cpart) rz += wave_sum_partials<8>(A.cpart + par * A.cnsb, A.cnsb);
    if (f != 0) return;  // uniform over the wavefront
    if (!(pq > 0.0)) {
        if (bid == pb && lane == 0) {
            A.flag[inst] = 2;
            A.scal[3 * inst + 2] = (double)it;
        }
        return;
    }
    const double alpha = rz / pq;
    double drz = 0.0, drr = 0.0;
    if (own) {
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A.x[3 * v + i] = xv[i] + alpha * pv[i];
            rv[i] = rv[i] - alpha * qv[i];
            r_out[3 * v + i] = rv[i];
        }
        double zv[3];
        if (cm.y < 0) {
#pragma unroll
            for (int i = 0; i < 3; ++i) zv[i] = o0[3 * i] * rv[0] + o0[3 * i + 1] * rv[1] + o0[3 * i + 2] * rv[2];
        } else {
            // z = M r with the members' updated residuals (the own one as above)
            cl_apply(cm, A.Z, A.nV, v, rv,
                     [&](int u, double* out) {
                         if (u == v) {
                             out[0] = rv[0];
                             out[1] = rv[1];
                             out[2] = rv[2];
                         } else {
#pragma unroll
                             for (int i = 0; i < 3; ++i) out[i] = r_in[3 * u + i] - alpha * A.q[3 * u + i];
                         }
                     },
                     zv);
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A.z[3 * v + i] = zv[i];
            drz += rv[i] * zv[i];
            drr += rv[i] * rv[i];
        }
    }
    drz = wave_sum(drz);
    drr = wave_sum(drr);
    if (lane == 0) {
        A.part_rz[(par ^ 1) * A.nbB + bid] = drz;
        A.part_rr[(par ^ 1) * A.nbB + bid] = drr;
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

void contact_structure(LinWork& L, hipStream_t s, int nV, long ncs0, const int* cs, long nf, const int* fcs)
{
    L.ncb = 0;
    L.nrec = 0;
    const long ncs = ncs0 + nf;
    if (ncs == 0) return;
    if (ncs >= (1L << 24)) throw std::length_error("contact structure: more than 2^24 stencils");
    int* vid = L.vid.ensure(4 * (size_t)ncs, s);
    int* nv = L.nv.ensure(ncs, s);
    if (ncs0) launch_constraint_vids(s, ncs0, cs, vid, nv);
    if (nf) launch_constraint_vids(s, nf, fcs, vid + 4 * ncs0, nv + ncs0);
    unsigned* cnt = L.u0.ensure(ncs + 1, s);
    unsigned* off = L.u1.ensure(ncs + 1, s);
    IDP_LAUNCH(k_contact_nv2, grid_for(ncs, 256), 256, 0, s, ncs, nv, cnt);
    IDP_HIP(hipMemsetAsync(cnt + ncs, 0, sizeof(unsigned), s));
    excl_scan<unsigned>(s, L.tmp, cnt, off, (int)ncs + 1);
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
    excl_scan<unsigned>(s, L.tmp, head, hs, (int)nrec + 1);
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
                             const double* gl, const double* gb, double* g, double* rhs, const unsigned char* act)
{
    bool haveC = L.ncb > 0 && gb;
    IDP_LAUNCH(k_gradient_gather, grid_for(m.nV, 256), 256, 0, s, m.nV, m.e_diag.p, m.e_src_ptr.p, m.e_src.p, gl, X,
               Xt, m.mass.p, haveC ? L.cdiag.p : (const int*)nullptr, L.seg.p, L.rv1.p, gb, m.proj_mask(),
               m.inst.vinst.p, act, g, rhs);
}

void assemble(LinWork& L, hipStream_t s, const MeshDev& m, const double* Hl, const double* Hb, bool any_dbc,
              const unsigned char* act)
{
    // algorithmic bytes: 80 per contribution (9 doubles + key), 76 per block, rowptr
    const double ncontrib = 16.0 * m.nT + 9.0 * m.nE + (double)L.nrec + m.nV;
    const int tok = L.prof ? L.prof->begin(IDP_K_ASSEMBLY) : -1;
    double* ev = L.evals.ensure(9 * (size_t)m.nnzb_e, s);
    IDP_LAUNCH(k_elastic_values, grid_for(m.nnzb_e, 256), 256, 0, s, m.nnzb_e, L.erow.p, m.e_col.p, m.e_src_ptr.p,
               m.e_src.p, Hl, m.mass.p, m.inst.vinst.p, act, ev);
    if (L.ncb) {
        double* cv = L.cvals.ensure(9 * (size_t)L.ncb, s);
        IDP_LAUNCH(k_contact_values, grid_for(L.ncb, 256), 256, 0, s, (unsigned)L.ncb, L.seg.p, L.rv1.p, Hb, L.crow.p,
                   m.inst.vinst.p, act, cv);
    }
    if (any_dbc) {
        IDP_LAUNCH(k_project_dbc, grid_for(m.nnzb_e, 256), 256, 0, s, m.nnzb_e, L.erow.p, m.e_col.p, m.proj_mask(), 1,
                   ev);
        if (L.ncb)
            IDP_LAUNCH(k_project_dbc, grid_for(L.ncb, 256), 256, 0, s, (long)L.ncb, L.crow.p, L.ccol.p, m.proj_mask(), 0,
                       L.cvals.p);
    }
    if (L.prof)
        L.prof->end(IDP_K_ASSEMBLY, tok, 80.0 * ncontrib + 76.0 * ((double)m.nnzb_e + L.ncb) + 4.0 * (m.nV + 1));
}

void build_inst_layout(MeshDev& m, const std::vector<long>& vbeg, const std::vector<long>& tbeg,
                       const std::vector<long>& nbeg, const std::vector<long>& ebeg, const std::vector<long>& fbeg,
                       hipStream_t s)
{
    InstDev& I = m.inst;
    I.B = (int)vbeg.size() - 1;
    I.vbeg = vbeg;
    I.tbeg = tbeg;
    I.nbeg = nbeg;
    I.ebeg = ebeg;
    I.fbeg = fbeg;
    std::vector<int> vi(std::max<long>(vbeg.back(), 1), 0);
    for (int i = 0; i < I.B; ++i)
        for (long v = vbeg[i]; v < vbeg[i + 1]; ++v) vi[v] = i;
    I.vinst.upload(vi.data(), vi.size(), s);
    auto up = [&](DBuf<int>& d, const std::vector<long>& h) {
        std::vector<int> t(h.begin(), h.end());
        d.upload(t.data(), t.size(), s);
    };
    up(I.d_vbeg, vbeg);
    up(I.d_nbeg, nbeg);
    up(I.d_ebeg, ebeg);
    up(I.d_fbeg, fbeg);
    I.echunk.build(ebeg, 256, s);
    I.vblk.build(vbeg, RB, s);
    I.nblk.build(nbeg, RB, s);
    I.tblk.build(tbeg, RB, s);
    for (int gi = 0; gi < 4; ++gi) I.pcg_rows[gi].build(vbeg, spmv_rows(4 << gi), s);
    I.pcg_upd.build(vbeg, 64, s);
    std::vector<long> v3(vbeg.size());
    for (size_t i = 0; i < vbeg.size(); ++i) v3[i] = 3 * vbeg[i];
    I.dblk.build(v3, RB, s);
    IDP_HIP(hipStreamSynchronize(s));
}

void init_erow(LinWork& L, hipStream_t s, const MeshDev& m, const std::vector<int>& rowof)
{
    L.erow.upload(rowof.data(), rowof.size(), s);
}

__global__ void k_pcg_start(int B, const unsigned char* __restrict__ solve, int* __restrict__ flag,
                            int* __restrict__ err)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= B) return;
    flag[i] = (solve && !solve[i]) ? 3 : 0;  // new and continued solves iterate
    err[i] = 0;
}

// Preconditioned CG per instance on the assembled system (E + C); b, x: 3V.
// L.precond picks cluster (0: elastic and contact matching, 2: contact only) or 3x3 (1) block Jacobi.
// solve (device, B bytes; null = all) selects the instances to solve.
// Per instance on return: iters[i] >= 0, or -1 (singular diagonal block) /
// -2 (breakdown, p.q <= 0); relres[i] = |r|/|b|.
void pcg_solve_batch(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                     int maxit, const unsigned char* solve_h, int* iters, double* relres, const unsigned char* warm,
                     bool handover)
{
    const InstDev& I = m.inst;
    const int B = I.B;
    const int nV = m.nV;
    // per instance: 0 not solving, 1 new solve, 2 continued (its state is on the device)
    std::vector<unsigned char> sv(B, 1);
    if (solve_h) sv.assign(solve_h, solve_h + B);
    L.h_it0.resize(B, 0);
    L.h_maxit.assign(B, 0);
    for (int i = 0; i < B; ++i)
        if (sv[i] != 2) L.h_it0[i] = 0;
    const unsigned char* solve = L.solve_d.upload_get(sv.data(), B, s);
    const double nnzb = (double)m.nnzb_e + (double)L.ncb;
    int* err = L.err.ensure(2 * (size_t)B + 2, s);
    int* flag = err + B;
    IDP_LAUNCH(k_pcg_start, grid_for(B, 64), 64, 0, s, B, solve, flag, err);
    const bool haveC = L.ncb > 0;
    Profiler* P = L.prof;
    const int tok = P ? P->begin(IDP_K_PCG) : -1;
    // cluster block-Jacobi preconditioner
    const int vg = grid_for(nV, 256);
    int* best = L.cl_best.ensure(nV, s);
    int* root = L.cl_root.ensure(nV, s);
    int* size = L.cl_size.ensure(nV, s);
    int* root2 = L.cl_root2.ensure(nV, s);
    int* cnt = L.cl_cnt.ensure(nV, s);
    int* mem = L.cl_mem.ensure(CL_MAX * (size_t)nV, s);
    unsigned long long* pack = L.cl_pack.ensure(nV, s);
    double* Z = L.pinv.ensure(CL_ZS * (size_t)nV, s);
    int4* cm4 = (int4*)L.cl_cm4.ensure(4 * (size_t)nV, s);
    const int* crp = haveC ? L.crp.p : nullptr;
    const int* crp_match = L.precond != 1 ? crp : nullptr;  // no contact rows: singletons (3x3 block Jacobi)
    // precond 0: the matching weighs elastic couplings too, so clusters form
    // everywhere; precond 2: contact couplings only
    const ClElastic elm{L.precond == 0 ? m.e_rowptr.p : nullptr, m.e_col.p, L.evals.p, m.nnzb_e};
    // only the instances that solve (new or continued) build their clusters
    const ClMask cmk{I.vinst.p, solve};
    IDP_LAUNCH(k_cl_best, vg, 256, 0, s, nV, crp_match, L.ccol.p, L.cvals.p, L.ncb, elm, cmk, best);
    IDP_LAUNCH(k_cl_pair, vg, 256, 0, s, nV, best, cmk, root, size);
    IDP_HIP(hipMemsetAsync(pack, 0, nV * sizeof(unsigned long long), s));
    IDP_LAUNCH(k_cl_best2, vg, 256, 0, s, nV, crp_match, L.ccol.p, L.cvals.p, L.ncb, root, size, elm, cmk, pack);
    IDP_LAUNCH(k_cl_merge, vg, 256, 0, s, nV, root, pack, size, cmk, root2);
    IDP_HIP(hipMemsetAsync(cnt, 0, nV * sizeof(int), s));
    IDP_LAUNCH(k_cl_members, vg, 256, 0, s, nV, root2, cmk, cnt, mem);
    IDP_LAUNCH(k_cl_sort, vg, 256, 0, s, nV, root2, cnt, cmk, mem);
    double* work = L.cl_work.ensure(2 * 144 * (size_t)nV, s);
    IDP_LAUNCH(k_cl_build, vg, 256, 0, s, nV, root2, cnt, mem, m.e_diag.p, m.e_rowptr.p, m.e_col.p, L.evals.p,
               m.nnzb_e, haveC ? L.cdiag.p : (const int*)nullptr, crp, L.ccol.p, L.cvals.p, L.ncb, I.vinst.p, cmk, Z,
               cm4, work, err);
    // coarse level: one instance with the positions of its system given
    const bool coarse = B == 1 && L.xpos && L.coarse.verts_per_agg > 0 && sv[0] == 1 && coarse_setup(L, s, m, L.xpos);
    PcgArgs A;
    A.nV = nV;
    A.cpart = coarse ? L.coarse.part.p : nullptr;
    A.cnsb = coarse ? L.coarse.nsb : 0;
    A.erp = m.e_rowptr.p;
    A.ecol = m.e_col.p;
    A.eval = L.evals.p;
    A.crp = haveC ? L.crp.p : nullptr;
    A.ccol = L.ccol.p;
    A.cval = L.cvals.p;
    A.eld = m.nnzb_e;
    A.cld = L.ncb;
    A.b = b;
    A.Z = Z;
    A.cm4 = cm4;
    A.x = x;
    A.r = L.r.ensure(6 * (size_t)nV, s);
    A.z = L.z.ensure(3 * (size_t)nV, s);
    A.pbuf = L.p.ensure(6 * (size_t)nV, s);
    A.q = L.q.ensure(3 * (size_t)nV, s);
    const SpmvShape shape = SPMV_SHAPES[L.spmv_shape];
    const int uw = std::min(std::max(L.upd_waves, 1), PCG_UPD_WAVES_MAX);
    const int gi = shape.g == 4 ? 0 : shape.g == 8 ? 1 : shape.g == 16 ? 2 : 3;
    const SegDev& rows = I.pcg_rows[gi];
    A.nbA = rows.nblk;
    A.nbB = I.pcg_upd.nblk;
    A.apbeg = rows.pbeg.p;
    A.bpbeg = I.pcg_upd.pbeg.p;
    // launch tables over the instances that solve (re-made as instances stop)
    std::vector<int> hf(2 * B);
    IDP_HIP(hipMemcpyAsync(hf.data(), err, 2 * B * sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    std::vector<char> run(B);
    for (int i = 0; i < B; ++i) run[i] = hf[B + i] == 0;
    int nA = 0, nB = 0;
    double frac = 1.0;  // launched share of the rows (exact for one instance, row-proportional otherwise)
    auto tables = [&]() {
        auto fill = [&](const SegDev& sd, std::vector<int4>& out, std::vector<int>& off) {
            out.clear();
            off.clear();
            for (int i = 0; i < B; ++i)
                if (run[i])
                    for (int k = sd.h_pbeg[i]; k < sd.h_pbeg[i + 1]; ++k) {
                        out.push_back(make_int4(sd.h_lo[k], sd.h_hi[k], i, k));
                        off.push_back(L.h_it0[i]);
                    }
        };
        fill(rows, L.h_atab, L.h_aoff);
        fill(I.pcg_upd, L.h_btab, L.h_boff);
        nA = (int)L.h_atab.size();
        nB = (int)L.h_btab.size();
        A.nbl_upd = nB;
        long rows = 0;
        for (const int4& t : L.h_btab) rows += t.y - t.x;
        frac = nV ? (double)rows / nV : 0.0;
        A.atab = L.atab.upload_get(L.h_atab.data(), L.h_atab.size(), s);
        A.btab = L.btab.upload_get(L.h_btab.data(), L.h_btab.size(), s);
        A.aoff = L.aoff.upload_get(L.h_aoff.data(), L.h_aoff.size(), s);
        A.boff = L.boff.upload_get(L.h_boff.data(), L.h_boff.size(), s);
        // iteration offset + 1 when one instance spans the launch (0: use the tables)
        A.ident = (B == 1 && nA == A.nbA && nB == A.nbB) ? L.h_it0[0] + 1 : 0;
    };
    A.solve = solve;
    tables();
    A.part_pq = L.part0.ensure(A.nbA + 1, s);
    A.part_bx = L.part_bx.ensure(A.nbA + 1, s);
    A.warm = warm;
    A.part_rz = L.part1.ensure(2 * (size_t)A.nbB + 2, s);
    A.part_rr = L.part2.ensure(2 * (size_t)A.nbB + 2, s);
    A.flag = flag;
    A.scal = L.scal.ensure(3 * (size_t)B + 3, s);
    A.rtol2 = rtol * rtol;
    A.maxit = maxit;
    // one instance with many update blocks: reduce the partials once per
    // kernel boundary instead of in every block (IDP_PCG_PREREDUCE = 0 / 1
    // overrides the size rule)
    {
        const char* e = std::getenv("IDP_PCG_PREREDUCE");
        const int force = e ? std::atoi(e) : -1;
        const bool big = A.nbB >= PCG_PREREDUCE_BLOCKS;
        A.red = (B == 1 && A.ident && !coarse && (force < 0 ? big : force > 0)) ? L.red.ensure(8, s) : nullptr;
    }
    if (warm && nA) {
        if (shape.g == 4) IDP_LAUNCH(k_pcg_warm<4>, nA, PCG_ROW_THREADS, 0, s, A);
        else if (shape.g == 8) IDP_LAUNCH(k_pcg_warm<8>, nA, PCG_ROW_THREADS, 0, s, A);
        else if (shape.g == 16) IDP_LAUNCH(k_pcg_warm<16>, nA, PCG_ROW_THREADS, 0, s, A);
        else IDP_LAUNCH(k_pcg_warm<32>, nA, PCG_ROW_THREADS, 0, s, A);
    }
    auto spmv = [&](int it, hipEvent_t e0, hipEvent_t e1) {
        const int nt = PCG_ROW_THREADS + (shape.sw ? 64 : 0);
        switch (L.spmv_shape) {
        case 0: IDP_LAUNCH_TIMED((k_pcg_spmv<4, false>), nA, nt, 0, s, e0, e1, A, it); break;
        case 1: IDP_LAUNCH_TIMED((k_pcg_spmv<8, false>), nA, nt, 0, s, e0, e1, A, it); break;
        case 2: IDP_LAUNCH_TIMED((k_pcg_spmv<16, false>), nA, nt, 0, s, e0, e1, A, it); break;
        case 3: IDP_LAUNCH_TIMED((k_pcg_spmv<32, false>), nA, nt, 0, s, e0, e1, A, it); break;
        case 4: IDP_LAUNCH_TIMED((k_pcg_spmv<4, true>), nA, nt, 0, s, e0, e1, A, it); break;
        case 5: IDP_LAUNCH_TIMED((k_pcg_spmv<8, true>), nA, nt, 0, s, e0, e1, A, it); break;
        case 6: IDP_LAUNCH_TIMED((k_pcg_spmv<16, true>), nA, nt, 0, s, e0, e1, A, it); break;
        default: IDP_LAUNCH_TIMED((k_pcg_spmv<32, true>), nA, nt, 0, s, e0, e1, A, it); break;
        }
    };
    if (nB) IDP_LAUNCH(k_pcg_init, nB, PCG_UPD_BLOCK, 0, s, A);
    if (coarse) coarse_apply(L, s, nV, A.r, A.z, flag, 0);
    // algorithmic bytes of one SpMV launch (blocks + columns, two row-pointer
    // arrays, z and p_old read, p_new and q written) and of one update launch
    // (x, p, r, q, P^-1 read; x, r, z written)
    const double spmv_bytes = 76.0 * nnzb + 8.0 * (nV + 1) + 96.0 * nV;
    const double upd_bytes = 240.0 * nV;
    std::vector<double> hs(3 * B);
    // Hand-over (batched engine): once half of the new and continued solves
    // have stopped, the others get as many iterations again and then keep
    // their state for the next call.  One solve alone is never handed over.
    int nstart = 0;
    for (int i = 0; i < B; ++i) nstart += run[i];
    int it_half = -1, it_end = INT_MAX;
    int it = 0, chunk = 8;
    for (; nA > 0 && it < it_end;) {
        for (int j = 0; j < chunk && it <= maxit && it < it_end; ++j, ++it) {
            // kernel times (events recorded by the dispatch itself) on one
            // iteration in 8
            const bool sample = P && (it & 7) == 3;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            int t = sample ? P->pair(e0, e1) : -1;
            if (A.red) IDP_LAUNCH(k_pcg_red_spmv, 1, PCG_RED, 0, s, A, it);
            spmv(it, e0, e1);
            if (sample) P->end_pair(IDP_K_PCG_SPMV, t, spmv_bytes * frac, it);
            if (A.red) IDP_LAUNCH(k_pcg_red_upd, 1, PCG_RED, 0, s, A, it);
            t = sample ? P->pair(e0, e1) : -1;
            IDP_LAUNCH_TIMED(k_pcg_update, (nB + uw - 1) / uw, PCG_UPD_BLOCK * uw, 0, s, e0, e1, A, it);
            if (sample) P->end_pair(IDP_K_PCG_UPDATE, t, upd_bytes * frac, it);
            if (coarse) coarse_apply(L, s, nV, A.r + (size_t)((it + 1) & 1) * 3 * nV, A.z, flag, (it + 1) & 1);
        }
        IDP_HIP(hipMemcpyAsync(hf.data(), err, 2 * B * sizeof(int), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        bool all = true, fewer = false;
        int nrun = 0;
        for (int i = 0; i < B; ++i) {
            all &= hf[B + i] != 0;
            if (run[i] && hf[B + i] != 0) {
                run[i] = 0;
                fewer = true;
            }
            nrun += run[i];
        }
        if (all || it > maxit) break;
        if (handover && nstart > 1 && it_half < 0 && 2 * nrun <= nstart) {
            it_half = it;
            it_end = 2 * it;
        }
        if (it >= it_end) break;
        if (fewer) tables();  // the stream is idle here: the tables may be rewritten
        chunk = std::min(chunk * 2, 256);
    }
    if (P) P->end(IDP_K_PCG, tok);
    IDP_HIP(hipMemcpyAsync(hs.data(), A.scal, 3 * B * sizeof(double), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    int maxdone = 0;
    double done_iters = 0;
    for (int i = 0; i < B; ++i) {
        const int f = hf[B + i];
        iters[i] = 0;
        relres[i] = 0;
        if (f == 0) {
            if (it < it_end) throw HipError(IDP_E_HIP, "PCG did not stop");
            iters[i] = PCG_HANDED_OVER;  // continues in the next call
            L.h_it0[i] += it;
            maxdone = std::max(maxdone, it);
            continue;
        }
        if (f == 3) continue;
        const int k = (int)hs[3 * i + 2];
        maxdone = std::max(maxdone, k - L.h_it0[i]);
        if (hf[i]) iters[i] = -1;
        else if (f == 2) iters[i] = -2;
        else {
            iters[i] = k;
            relres[i] = hs[3 * i] > 0 ? std::sqrt(hs[3 * i + 1] / hs[3 * i]) : 0.0;
            done_iters += k;
            if (f == 4) L.h_maxit[i] = 1;
        }
    }
    if (P) {
        // launches at or after the last stopping iteration did no work
        P->discard_from(IDP_K_PCG_SPMV, maxdone);
        P->discard_from(IDP_K_PCG_UPDATE, maxdone);
        P->checkpoint();
        P->add_work(IDP_K_PCG, B == 1 ? done_iters * (spmv_bytes + upd_bytes) : 0.0);
    }
}

int pcg_solve_persistent(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                         int maxit, double* relres, const unsigned char* warm)
{
    int it = 0;
    double rel = 0;
    pcg_solve_batch(L, s, m, b, x, rtol, maxit, nullptr, &it, &rel, warm, false);
    *relres = rel;
    return it;
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

void launch_sum_partials(hipStream_t s, int npart, const double* part, double* out)
{
    IDP_LAUNCH(k_sum_partials, 1, RB, 0, s, npart, part, out);
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
