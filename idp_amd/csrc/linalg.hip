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

// ---------------------------------------------------------------------------
// Block-Jacobi PCG as two kernels per iteration, enqueued in chunks with no
// host synchronisation inside a chunk.  A grid-wide barrier inside one
// persistent kernel costs a device-scope release/acquire per block (the
// eight XCD L2s are not coherent for ordinary device memory), which grows
// with the block count; a kernel boundary does that once.
//  * k_pcg_spmv(it): every block re-reduces the previous update's partials
//    (r.z, r.r) in a fixed order, decides convergence identically, forms
//    p_it = z + beta p_(it-1) on the fly and computes q = A p_it with a group
//    of G lanes per block row (xor-butterfly row reduction), plus partial p.q.
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
    const double* eval;
    const int* crp;  // null when there are no contact blocks
    const int* ccol;
    const double* cval;
    const double* b;
    const double* Pinv;
    double* x;
    double* r;
    double* z;
    double* pbuf;     // 2 x 3V
    double* q;
    double* part_pq;  // nbA
    double* part_rz;  // 2 x nbB
    double* part_rr;  // 2 x nbB
    int nbA, nbB;
    int* flag;        // 0 running, 1 stopped (converged or maxit), 2 breakdown
    double* scal;     // [0] |b|^2, [1] final |r|^2, [2] iterations
    double rtol2;
    int maxit;
};

constexpr int PCG_SPMV_BLOCK = 1024;

template <int G>
__device__ inline void group_row_spmv(int v, int lane, const int* __restrict__ rp, const int* __restrict__ col,
                                      const double* __restrict__ val, const double* __restrict__ z,
                                      const double* __restrict__ po, double beta, double& y0, double& y1, double& y2)
{
    const int k1 = rp[v + 1];
    for (int k = rp[v] + lane; k < k1; k += G) {
        const double* bk = val + 9 * (size_t)k;
        const size_t c = 3 * (size_t)col[k];
        double x0 = z[c], x1 = z[c + 1], x2 = z[c + 2];
        if (po) {
            x0 += beta * po[c];
            x1 += beta * po[c + 1];
            x2 += beta * po[c + 2];
        }
        y0 += bk[0] * x0 + bk[1] * x1 + bk[2] * x2;
        y1 += bk[3] * x0 + bk[4] * x1 + bk[5] * x2;
        y2 += bk[6] * x0 + bk[7] * x1 + bk[8] * x2;
    }
}

// x = 0, r = b, z = P b, partials of r.z and r.r (parity 0)
__global__ void __launch_bounds__(RB) k_pcg_init(PcgArgs A)
{
    __shared__ double sh[RB];
    const int v = blockIdx.x * RB + threadIdx.x;
    double drz = 0.0, drr = 0.0;
    if (v < A.nV) {
        const double rv[3] = {A.b[3 * v], A.b[3 * v + 1], A.b[3 * v + 2]};
        const double* P = A.Pinv + 9 * (size_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A.x[3 * v + i] = 0.0;
            A.r[3 * v + i] = rv[i];
            const double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
            A.z[3 * v + i] = zi;
            drz += rv[i] * zi;
            drr += rv[i] * rv[i];
        }
    }
    const double a = block_sum<RB>(drz, sh), c = block_sum<RB>(drr, sh);
    if (threadIdx.x == 0) {
        A.part_rz[blockIdx.x] = a;
        A.part_rr[blockIdx.x] = c;
        if (blockIdx.x == 0) A.flag[0] = 0;
    }
}

template <int G>
__global__ void __launch_bounds__(PCG_SPMV_BLOCK) k_pcg_spmv(PcgArgs A, int it)
{
    __shared__ double sh[PCG_SPMV_BLOCK];
    __shared__ int sflag;
    if (threadIdx.x == 0) sflag = *(volatile int*)A.flag;
    __syncthreads();
    if (sflag != 0) return;  // block-uniform
    const int par = it & 1;
    const double rz = sum_partials_block_n<PCG_SPMV_BLOCK>(A.part_rz + par * A.nbB, A.nbB, sh);
    const double rr = sum_partials_block_n<PCG_SPMV_BLOCK>(A.part_rr + par * A.nbB, A.nbB, sh);
    const double bb = it == 0 ? rr : A.scal[0];
    if (!(bb > 0.0) || (it > 0 && rr <= A.rtol2 * bb) || it >= A.maxit) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            A.flag[0] = 1;
            A.scal[0] = bb;
            A.scal[1] = rr;
            A.scal[2] = (double)it;
        }
        return;
    }
    if (it == 0 && blockIdx.x == 0 && threadIdx.x == 0) A.scal[0] = bb;
    double beta = 0.0;
    const double* po = nullptr;
    if (it > 0) {
        beta = rz / sum_partials_block_n<PCG_SPMV_BLOCK>(A.part_rz + (par ^ 1) * A.nbB, A.nbB, sh);
        po = A.pbuf + (size_t)(par ^ 1) * 3 * A.nV;
    }
    double* pn = A.pbuf + (size_t)par * 3 * A.nV;
    const int lane = threadIdx.x & (G - 1);
    const int v = (blockIdx.x * PCG_SPMV_BLOCK + threadIdx.x) / G;
    double dpq = 0.0;
    if (v < A.nV) {
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        group_row_spmv<G>(v, lane, A.erp, A.ecol, A.eval, A.z, po, beta, y0, y1, y2);
        if (A.crp) group_row_spmv<G>(v, lane, A.crp, A.ccol, A.cval, A.z, po, beta, y0, y1, y2);
#pragma unroll
        for (int o = G / 2; o > 0; o >>= 1) {
            y0 += __shfl_xor(y0, o, G);
            y1 += __shfl_xor(y1, o, G);
            y2 += __shfl_xor(y2, o, G);
        }
        if (lane < 3) {
            const double yl = lane == 0 ? y0 : (lane == 1 ? y1 : y2);
            const size_t i = 3 * (size_t)v + lane;
            const double pv = po ? A.z[i] + beta * po[i] : A.z[i];
            pn[i] = pv;
            A.q[i] = yl;
            dpq = pv * yl;
        }
    }
    const double a = block_sum<PCG_SPMV_BLOCK>(dpq, sh);
    if (threadIdx.x == 0) A.part_pq[blockIdx.x] = a;
}

__global__ void __launch_bounds__(RB) k_pcg_update(PcgArgs A, int it)
{
    __shared__ double sh[RB];
    __shared__ int sflag;
    if (threadIdx.x == 0) sflag = *(volatile int*)A.flag;
    __syncthreads();
    if (sflag != 0) return;  // block-uniform
    const int par = it & 1;
    const double pq = sum_partials_block(A.part_pq, A.nbA, sh);
    const double rz = sum_partials_block(A.part_rz + par * A.nbB, A.nbB, sh);
    if (!(pq > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            A.flag[0] = 2;
            A.scal[2] = (double)it;
        }
        return;
    }
    const double alpha = rz / pq;
    const double* p = A.pbuf + (size_t)par * 3 * A.nV;
    const int v = blockIdx.x * RB + threadIdx.x;
    double drz = 0.0, drr = 0.0;
    if (v < A.nV) {
        double rv[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            A.x[3 * v + i] += alpha * p[3 * v + i];
            rv[i] = A.r[3 * v + i] - alpha * A.q[3 * v + i];
            A.r[3 * v + i] = rv[i];
        }
        const double* P = A.Pinv + 9 * (size_t)v;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            const double zi = P[3 * i] * rv[0] + P[3 * i + 1] * rv[1] + P[3 * i + 2] * rv[2];
            A.z[3 * v + i] = zi;
            drz += rv[i] * zi;
            drr += rv[i] * rv[i];
        }
    }
    const double a = block_sum<RB>(drz, sh), c = block_sum<RB>(drr, sh);
    if (threadIdx.x == 0) {
        A.part_rz[(par ^ 1) * A.nbB + blockIdx.x] = a;
        A.part_rr[(par ^ 1) * A.nbB + blockIdx.x] = c;
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

int pcg_solve_persistent(LinWork& L, hipStream_t s, const MeshDev& m, const double* b, double* x, double rtol,
                         int maxit, double* relres)
{
    const int nV = m.nV;
    const double nnzb = (double)m.nnzb_e + (double)L.ncb;
    // lanes per block row from the mean row length (a tet-mesh row holds ~15 blocks)
    const double avg = nnzb / std::max(nV, 1);
    const int G = avg <= 8.0 ? 4 : (avg <= 24.0 ? 8 : 16);
    double* Pinv = L.pinv.ensure(9 * (size_t)nV, s);
    int* err = L.err.ensure(4, s);
    IDP_HIP(hipMemsetAsync(err, 0, sizeof(int), s));
    const bool haveC = L.ncb > 0;
    const int tok = L.prof ? L.prof->begin(IDP_K_PCG) : -1;
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
    A.pbuf = L.p.ensure(6 * (size_t)nV, s);
    A.q = L.q.ensure(3 * (size_t)nV, s);
    A.nbA = std::max(1, grid_for((long)nV * G, PCG_SPMV_BLOCK));
    A.nbB = std::max(1, grid_for(nV, RB));
    A.part_pq = L.part0.ensure(A.nbA, s);
    A.part_rz = L.part1.ensure(2 * (size_t)A.nbB, s);
    A.part_rr = L.part2.ensure(2 * (size_t)A.nbB, s);
    A.flag = err + 1;
    A.scal = L.scal.ensure(8, s);
    A.rtol2 = rtol * rtol;
    A.maxit = maxit;
    IDP_LAUNCH(k_pcg_init, A.nbB, RB, 0, s, A);
    struct {
        int err, flag;
    } hf{0, 0};
    double hs[3] = {0, 0, 0};
    // algorithmic bytes of one SpMV launch (blocks + columns, two row-pointer
    // arrays, z and p_old read, p_new and q written) and of one update launch
    // (x, p, r, q, P^-1 read; x, r, z written)
    const double spmv_bytes = 76.0 * nnzb + 8.0 * (nV + 1) + 96.0 * nV;
    const double upd_bytes = 240.0 * nV;
    Profiler* P = L.prof;
    int it = 0, chunk = 8;
    for (;;) {
        for (int j = 0; j < chunk && it <= maxit; ++j, ++it) {
            // per-kernel events on one iteration in 8 (an event pair between
            // every launch would itself widen the launch gaps)
            const bool sample = P && (it & 7) == 3;
            int t = sample ? P->begin(IDP_K_PCG_SPMV) : -1;
            if (G == 4) IDP_LAUNCH(k_pcg_spmv<4>, A.nbA, PCG_SPMV_BLOCK, 0, s, A, it);
            else if (G == 8) IDP_LAUNCH(k_pcg_spmv<8>, A.nbA, PCG_SPMV_BLOCK, 0, s, A, it);
            else IDP_LAUNCH(k_pcg_spmv<16>, A.nbA, PCG_SPMV_BLOCK, 0, s, A, it);
            if (sample) P->end(IDP_K_PCG_SPMV, t, spmv_bytes, it);
            t = sample ? P->begin(IDP_K_PCG_UPDATE) : -1;
            IDP_LAUNCH(k_pcg_update, A.nbB, RB, 0, s, A, it);
            if (sample) P->end(IDP_K_PCG_UPDATE, t, upd_bytes, it);
        }
        IDP_HIP(hipMemcpyAsync(&hf, err, sizeof(hf), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipMemcpyAsync(hs, A.scal, sizeof(hs), hipMemcpyDeviceToHost, s));
        IDP_HIP(hipStreamSynchronize(s));
        if (hf.err || hf.flag || it > maxit) break;
        chunk = std::min(chunk * 2, 256);
    }
    if (P) P->end(IDP_K_PCG, tok);
    if (hf.err) return -1;
    if (hf.flag == 0) throw HipError(IDP_E_HIP, "PCG did not stop");
    const int iters = (int)hs[2];
    if (P) {
        // launches at or after the stopping iteration did no SpMV/update work
        P->discard_from(IDP_K_PCG_SPMV, iters);
        P->discard_from(IDP_K_PCG_UPDATE, iters);
        P->checkpoint();
    }
    if (hf.flag == 2) return -2;
    *relres = hs[0] > 0 ? std::sqrt(hs[1] / hs[0]) : 0.0;
    // algorithmic bytes per PCG iteration (each array touched once):
    //   SpMV: blocks + columns 76 nnzb, two row pointers 8 (V+1),
    //         read z, p_old, write p_new, q: 96 V
    //   update: read x, p_new, r, q, P^-1 (96 V + 72 V), write x, r, z (72 V)
    if (P) P->add_work(IDP_K_PCG, iters * (spmv_bytes + upd_bytes));
    return iters;
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
