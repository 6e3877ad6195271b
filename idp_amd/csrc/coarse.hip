// Coarse level of the PCG preconditioner for one mesh: an additive
// two-level preconditioner M = M1 + P A_c^-1 P^T, with M1 the cluster block
// Jacobi and P the rigid-body modes (3 translations, 3 rotations about the
// centroid, orthonormal per aggregate) of aggregates of free vertices that
// are cells of a voxel grid over the current positions.  The PCG solves of
// the Newton iterations are ill-conditioned in the smooth, near-rigid motions
// of large pieces of the mesh (one iteration of block Jacobi moves
// information one block row); the coarse space removes those directly.
//
// Every few Newton iterations (coarse_setup): the positions are read back, the
// aggregates and P are built on the host (6 columns per aggregate, a
// rotation direction dropped when its inertia is tiny), A_c = P^T A P is
// assembled on the device with a fixed summation order (a stable radix sort
// of the blocks by aggregate pair, one wavefront per pair), and replaced by
// its inverse (Cholesky, rocSOLVER potrf + potri).  In between, the level of
// the last build is reused: with an earlier system's A_c^-1 and P the
// preconditioner is still symmetric positive definite, only less exact.  Per PCG iteration (two
// kernels after the update): r_c = P^T r (one workgroup per aggregate), then
// per aggregate its rows of e_c = A_c^-1 r_c, their part of r_c . e_c and
// z += P e_c on its vertices (one workgroup per aggregate).  Since r . P e_c = r_c . e_c, the preconditioned dot of the
// iteration is the update's r . z1 plus those partials.
// Dirichlet vertices belong to no aggregate, so the coarse correction never
// moves them.
#include <cmath>
#include <unordered_map>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "common.hpp"
#include "kernels.hpp"
#include "linalg.hpp"
#include "prim.hpp"

namespace idp {

namespace {

// key of each block by the aggregate pair of its row and column; K * K (a
// run of its own, sorted last, then skipped) when either vertex is in no
// aggregate
__global__ void k_coarse_keys(long ne, const int* __restrict__ erow, const int* __restrict__ ecol, long nc,
                              const int* __restrict__ crow, const int* __restrict__ ccol, const int* __restrict__ agg,
                              int K, unsigned long long* __restrict__ key, unsigned* __restrict__ val)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne + nc) return;
    const int r = i < ne ? erow[i] : crow[i - ne];
    const int c = i < ne ? ecol[i] : ccol[i - ne];
    const int ar = agg[r], ac = agg[c];
    key[i] = (ar < 0 || ac < 0) ? (unsigned long long)K * K : (unsigned long long)ar * K + ac;
    val[i] = (unsigned)i;
}

__global__ void k_coarse_heads(long n, const unsigned long long* __restrict__ key, unsigned* __restrict__ head)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    head[i] = (i < n && (i == 0 || key[i] != key[i - 1])) ? 1u : 0u;
}

// segment starts: seg[hs[i]] = i for each run head
__global__ void k_coarse_segs(long n, const unsigned* __restrict__ head, const unsigned* __restrict__ hs,
                              unsigned* __restrict__ seg)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !head[i]) return;
    seg[hs[i]] = (unsigned)i;
}

__global__ void k_coarse_fill(long n, double* __restrict__ a, double v)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

// one wavefront per aggregate pair: sum of P_r^T A_rc P_c over its blocks
// (lanes take every 64th block, then a fixed butterfly), into the dense
// column-major A_c
__global__ void __launch_bounds__(256) k_coarse_ac(unsigned nseg, const unsigned* __restrict__ seg,
                                                   unsigned nrun, const unsigned long long* __restrict__ key,
                                                   const unsigned* __restrict__ val, long ne,
                                                   const int* __restrict__ erow, const int* __restrict__ ecol,
                                                   const double* __restrict__ evals, const int* __restrict__ crow,
                                                   const int* __restrict__ ccol, const double* __restrict__ cvals,
                                                   long nc, const double* __restrict__ P, int K, int ncd,
                                                   double* __restrict__ Ac)
{
    const unsigned sgi = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (sgi >= nseg) return;
    const int lane = threadIdx.x & 63;
    const unsigned b0 = seg[sgi], b1 = sgi + 1 < nseg ? seg[sgi + 1] : nrun;
    const unsigned long long kk = key[b0];
    if (kk == (unsigned long long)K * K) return;  // blocks outside the aggregates
    double acc[36];
#pragma unroll
    for (int e = 0; e < 36; ++e) acc[e] = 0.0;
    for (unsigned j = b0 + lane; j < b1; j += 64) {
        const long i = val[j];
        const bool el = i < ne;
        const long k = el ? i : i - ne;
        const int r = el ? erow[k] : crow[k];
        const int c = el ? ecol[k] : ccol[k];
        const double* vals = el ? evals : cvals;
        const long ld = el ? ne : nc;
        double a[9];
#pragma unroll
        for (int e = 0; e < 9; ++e) a[e] = vals[e * ld + k];
        const double* Pr = P + 18 * (size_t)r;
        const double* Pc = P + 18 * (size_t)c;
        double ap[18];  // A P_c, 3 x 6
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int m = 0; m < 6; ++m) ap[6 * q + m] = a[3 * q] * Pc[m] + a[3 * q + 1] * Pc[6 + m] + a[3 * q + 2] * Pc[12 + m];
#pragma unroll
        for (int l = 0; l < 6; ++l)
#pragma unroll
            for (int m = 0; m < 6; ++m) acc[6 * l + m] += Pr[l] * ap[m] + Pr[6 + l] * ap[6 + m] + Pr[12 + l] * ap[12 + m];
    }
#pragma unroll
    for (int e = 0; e < 36; ++e) {
        double v = acc[e];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[e] = v;
    }
    if (lane == 0) {
        const int ar = (int)(kk / (unsigned long long)K), ac = (int)(kk % (unsigned long long)K);
#pragma unroll
        for (int l = 0; l < 6; ++l)
#pragma unroll
            for (int m = 0; m < 6; ++m) Ac[(size_t)(6 * ar + l) + (size_t)(6 * ac + m) * ncd] = acc[6 * l + m];
    }
}

// dofs of dropped modes (zero rows): 1 on the diagonal
__global__ void k_coarse_diag(int ncd, double* __restrict__ Ac)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < ncd && Ac[(size_t)i * ncd + i] == 0.0) Ac[(size_t)i * ncd + i] = 1.0;
}

// the inverse from potri is in the lower triangle: mirror it
__global__ void k_coarse_sym(int ncd, double* __restrict__ Ac)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)ncd * ncd) return;
    const int i = (int)(t % ncd), j = (int)(t / ncd);  // column-major (i row, j column)
    if (i < j) Ac[(size_t)i + (size_t)j * ncd] = Ac[(size_t)j + (size_t)i * ncd];
}

}  // namespace

// ---------------------------------------------------------------------------
// per PCG iteration
// ---------------------------------------------------------------------------
constexpr int COARSE_BLOCK = 256;
constexpr int COARSE_MAX_AGG = 320;  // aggregates at most (A_c of order 1920, 29 MB)

// fixed-order sum of N values over a workgroup of COARSE_BLOCK threads:
// xor butterfly per wavefront, then the wavefronts in order (all threads
// get the totals)
template <int N>
__device__ inline void coarse_block_sum(double* acc, double (*sh)[N])
{
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int m = 0; m < N; ++m) {
        double v = acc[m];
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
        acc[m] = v;
    }
    if (lane == 0)
#pragma unroll
        for (int m = 0; m < N; ++m) sh[w][m] = acc[m];
    __syncthreads();
#pragma unroll
    for (int m = 0; m < N; ++m) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < COARSE_BLOCK / 64; ++k) v += sh[k][m];
        acc[m] = v;
    }
}

// r_c = P^T r: one workgroup per aggregate over its vertex list
__global__ void __launch_bounds__(COARSE_BLOCK) k_coarse_restrict(int K, const int* __restrict__ aptr,
                                                                  const int* __restrict__ avert,
                                                                  const double* __restrict__ Ps,
                                                                  const double* __restrict__ r,
                                                                  const int* __restrict__ flag,
                                                                  double* __restrict__ rc)
{
    __shared__ double sh[COARSE_BLOCK / 64][6];
    const int a = blockIdx.x;
    if (*(volatile const int*)flag != 0) return;  // uniform over the grid
    double acc[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = aptr[a] + (int)threadIdx.x; j < aptr[a + 1]; j += COARSE_BLOCK) {
        const int v = avert[j];
        const double* Pv = Ps + 18 * (size_t)j;
        const double r0 = r[3 * v], r1 = r[3 * v + 1], r2 = r[3 * v + 2];
#pragma unroll
        for (int m = 0; m < 6; ++m) acc[m] += Pv[m] * r0 + Pv[6 + m] * r1 + Pv[12 + m] * r2;
    }
    coarse_block_sum<6>(acc, sh);
    if (threadIdx.x < 6) {
        double v = acc[0];
#pragma unroll
        for (int m = 1; m < 6; ++m) v = (int)threadIdx.x == m ? acc[m] : v;
        rc[6 * a + threadIdx.x] = v;
    }
}

// one workgroup per aggregate a: e = rows 6a..6a+5 of A_c^-1 times r_c, the
// partial r_c . e_c of those rows, and z += P_v e for the vertices of a
__global__ void __launch_bounds__(COARSE_BLOCK) k_coarse_solve(int K, int ncd, const int* __restrict__ aptr,
                                                               const int* __restrict__ avert,
                                                               const double* __restrict__ Ps,
                                                               const double* __restrict__ Ainv,
                                                               const double* __restrict__ rc,
                                                               const int* __restrict__ flag,
                                                               double* __restrict__ part, double* __restrict__ z)
{
    __shared__ double sh[COARSE_BLOCK / 64][6];
    const int a = blockIdx.x;
    if (*(volatile const int*)flag != 0) return;  // uniform over the grid
    double e[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    const double* A0 = Ainv + (size_t)(6 * a) * ncd;  // symmetric: rows = columns
    // eight columns' loads in flight at once; each thread still adds its
    // columns in increasing order
    constexpr int U = 8;
    int j = threadIdx.x;
    for (; j + (U - 1) * COARSE_BLOCK < ncd; j += U * COARSE_BLOCK) {
        double rj[U], av[U][6];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            rj[u] = rc[j + u * COARSE_BLOCK];
#pragma unroll
            for (int m = 0; m < 6; ++m) av[u][m] = A0[(size_t)m * ncd + j + u * COARSE_BLOCK];
        }
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int m = 0; m < 6; ++m) e[m] += av[u][m] * rj[u];
    }
    for (; j < ncd; j += COARSE_BLOCK) {
        const double rj = rc[j];
#pragma unroll
        for (int m = 0; m < 6; ++m) e[m] += A0[(size_t)m * ncd + j] * rj;
    }
    coarse_block_sum<6>(e, sh);
    if (threadIdx.x == 0) {
        double d = 0.0;
#pragma unroll
        for (int m = 0; m < 6; ++m) d += rc[6 * a + m] * e[m];
        part[a] = d;
    }
    for (int j = aptr[a] + (int)threadIdx.x; j < aptr[a + 1]; j += COARSE_BLOCK) {
        const int v = avert[j];
        const double* Pv = Ps + 18 * (size_t)j;
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 6; ++m) s += Pv[6 * q + m] * e[m];
            z[3 * v + q] += s;
        }
    }
}

// ---------------------------------------------------------------------------
// host
// ---------------------------------------------------------------------------
namespace {

// symmetric 3x3 eigen-decomposition by cyclic Jacobi (columns of V)
void eig3(double* a, double* w, double* V)
{
    for (int i = 0; i < 9; ++i) V[i] = (i % 4 == 0) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 50; ++sweep) {
        const double off = a[1] * a[1] + a[2] * a[2] + a[5] * a[5];
        if (off < 1e-300) break;
        for (int p = 0; p < 2; ++p)
            for (int q = p + 1; q < 3; ++q) {
                const double apq = a[3 * p + q];
                if (apq == 0.0) continue;
                const double th = (a[3 * q + q] - a[3 * p + p]) / (2 * apq);
                const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1));
                const double c = 1 / std::sqrt(t * t + 1), s = t * c;
                for (int k = 0; k < 3; ++k) {
                    const double x = a[3 * k + p], y = a[3 * k + q];
                    a[3 * k + p] = c * x - s * y;
                    a[3 * k + q] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = a[3 * p + k], y = a[3 * q + k];
                    a[3 * p + k] = c * x - s * y;
                    a[3 * q + k] = s * x + c * y;
                }
                for (int k = 0; k < 3; ++k) {
                    const double x = V[3 * k + p], y = V[3 * k + q];
                    V[3 * k + p] = c * x - s * y;
                    V[3 * k + q] = s * x + c * y;
                }
            }
    }
    for (int i = 0; i < 3; ++i) w[i] = a[4 * i];
}

}  // namespace

CoarseWork::~CoarseWork()
{
    if (handle) rocblas_destroy_handle((rocblas_handle)handle);
}

bool coarse_setup(LinWork& L, hipStream_t s, const MeshDev& m, const double* X)
{
    CoarseWork& C = L.coarse;
    const int nV = m.nV;
    if (C.verts_per_agg <= 0 || m.inst.B != 1 || !X || nV < 2 * C.verts_per_agg) return false;
    if (C.ready && C.age < C.rebuild_every && C.nV == nV && C.mask == m.proj_mask()) {
        ++C.age;  // the level of an earlier system of this mesh
        return true;
    }
    C.ready = false;
    // positions and Dirichlet marks on the host
    C.hx.resize(3 * (size_t)nV);
    IDP_HIP(hipMemcpyAsync(C.hx.data(), X, 3 * (size_t)nV * sizeof(double), hipMemcpyDeviceToHost, s));
    std::vector<unsigned char> dbc(nV, 0);
    if (m.proj_mask()) IDP_HIP(hipMemcpyAsync(dbc.data(), m.proj_mask(), nV, hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    const double* x = C.hx.data();
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    int nfree = 0;
    for (int v = 0; v < nV; ++v) {
        if (dbc[v]) continue;
        ++nfree;
        for (int d = 0; d < 3; ++d) {
            lo[d] = std::min(lo[d], x[3 * v + d]);
            hi[d] = std::max(hi[d], x[3 * v + d]);
        }
    }
    if (nfree < 2 * C.verts_per_agg) return false;
    // voxel edge for about nfree / verts_per_agg cells over the box
    double vol = 1.0;
    int nd = 0;
    for (int d = 0; d < 3; ++d)
        if (hi[d] > lo[d]) {
            vol *= hi[d] - lo[d];
            ++nd;
        }
    const double target = std::min((double)nfree / C.verts_per_agg, (double)COARSE_MAX_AGG);
    double hcell = std::pow(vol / target, 1.0 / std::max(nd, 1));
    // aggregates: voxel cells in first-seen vertex order; coarser cells while
    // there are more than COARSE_MAX_AGG of them (A_c^-1 is read whole by
    // every PCG iteration)
    std::unordered_map<long long, int> cell;
    std::vector<int> agg(nV, -1);
    for (int pass = 0; pass < 8; ++pass) {
        cell.clear();
        for (int v = 0; v < nV; ++v) {
            if (dbc[v]) continue;
            long long k = 0;
            for (int d = 0; d < 3; ++d) k = k * 2097152 + (long long)std::floor((x[3 * v + d] - lo[d]) / hcell);
            auto it = cell.emplace(k, (int)cell.size()).first;
            agg[v] = it->second;
        }
        if ((int)cell.size() <= COARSE_MAX_AGG) break;
        hcell *= std::cbrt((double)cell.size() / COARSE_MAX_AGG) * 1.05;
    }
    const int K = (int)cell.size();
    if (K > COARSE_MAX_AGG || K < 2) return false;
    // per aggregate: vertex lists, centroids, inertia
    std::vector<int> aptr(K + 1, 0), avert(nfree);
    for (int v = 0; v < nV; ++v)
        if (agg[v] >= 0) ++aptr[agg[v] + 1];
    for (int a = 0; a < K; ++a) aptr[a + 1] += aptr[a];
    {
        std::vector<int> f(aptr.begin(), aptr.end() - 1);
        for (int v = 0; v < nV; ++v)
            if (agg[v] >= 0) avert[f[agg[v]]++] = v;
    }
    std::vector<double> P(18 * (size_t)nV, 0.0);
    for (int a = 0; a < K; ++a) {
        const int n = aptr[a + 1] - aptr[a];
        double c[3] = {0, 0, 0};
        for (int j = aptr[a]; j < aptr[a + 1]; ++j)
            for (int d = 0; d < 3; ++d) c[d] += x[3 * avert[j] + d];
        for (int d = 0; d < 3; ++d) c[d] /= n;
        double J[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int j = aptr[a]; j < aptr[a + 1]; ++j) {
            const int v = avert[j];
            const double dv[3] = {x[3 * v] - c[0], x[3 * v + 1] - c[1], x[3 * v + 2] - c[2]};
            const double dd = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
            for (int p = 0; p < 3; ++p)
                for (int q = 0; q < 3; ++q) J[3 * p + q] += (p == q ? dd : 0.0) - dv[p] * dv[q];
        }
        double w[3], V[9];
        eig3(J, w, V);
        const double wmax = std::max(w[0], std::max(w[1], w[2]));
        // rotation directions V_k scaled by w_k^(-1/2); tiny inertia: dropped
        double Rs[9];
        for (int k = 0; k < 3; ++k) {
            const double sc = (wmax > 0 && w[k] > 1e-10 * wmax) ? 1.0 / std::sqrt(w[k]) : 0.0;
            for (int p = 0; p < 3; ++p) Rs[3 * p + k] = V[3 * p + k] * sc;
        }
        const double tn = 1.0 / std::sqrt((double)n);
        for (int j = aptr[a]; j < aptr[a + 1]; ++j) {
            const int v = avert[j];
            const double d[3] = {x[3 * v] - c[0], x[3 * v + 1] - c[1], x[3 * v + 2] - c[2]};
            double* Pv = P.data() + 18 * (size_t)v;
            // translations
            for (int q = 0; q < 3; ++q) Pv[6 * q + q] = tn;
            // rotation about axis u: u x d = -[d]x u; column k uses u = Rs(:, k)
            for (int k = 0; k < 3; ++k) {
                const double u[3] = {Rs[k], Rs[3 + k], Rs[6 + k]};
                const double cr[3] = {u[1] * d[2] - u[2] * d[1], u[2] * d[0] - u[0] * d[2], u[0] * d[1] - u[1] * d[0]};
                for (int q = 0; q < 3; ++q) Pv[6 * q + 3 + k] = cr[q];
            }
        }
    }
    const int ncd = 6 * K;
    C.K = K;
    C.ncd = ncd;
    int* d_agg = C.agg.upload_get(agg.data(), nV, s);
    C.aptr.upload(aptr.data(), aptr.size(), s);
    C.avert.upload(avert.data(), avert.size(), s);
    double* d_P = C.P.upload_get(P.data(), P.size(), s);
    {
        // the apply kernels read P by slot, so its loads do not wait for avert's
        std::vector<double> Ps(18 * (size_t)nfree);
        for (int j = 0; j < nfree; ++j)
            for (int q = 0; q < 18; ++q) Ps[18 * (size_t)j + q] = P[18 * (size_t)avert[j] + q];
        C.Ps.upload(Ps.data(), Ps.size(), s);
    }
    // A_c = P^T A P, summed per aggregate pair in block order
    const long ne = m.nnzb_e, nc = L.ncb, nb = ne + nc;
    unsigned long long* k0 = C.key0.ensure(nb + 1, s);
    unsigned* v0 = C.val0.ensure(nb + 1, s);
    IDP_LAUNCH(k_coarse_keys, grid_for(nb, 256), 256, 0, s, ne, L.erow.p, m.e_col.p, nc, L.crow.p, L.ccol.p, d_agg, K,
               k0, v0);
    unsigned long long* k1 = C.key1.ensure(nb + 1, s);
    unsigned* v1 = C.val1.ensure(nb + 1, s);
    int kbits = 1;
    while (kbits < 63 && ((unsigned long long)K * K >> kbits)) ++kbits;
    sort_pairs(s, C.tmp, k0, k1, v0, v1, nb, kbits);
    unsigned* head = C.head.ensure(nb + 1, s);
    unsigned* hs = C.hs.ensure(nb + 1, s);
    IDP_LAUNCH(k_coarse_heads, grid_for(nb + 1, 256), 256, 0, s, nb, k1, head);
    excl_scan<unsigned>(s, C.tmp, head, hs, nb + 1);
    unsigned nseg = 0;
    IDP_HIP(hipMemcpyAsync(&nseg, hs + nb, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    const unsigned nrun = (unsigned)nb;
    unsigned* seg = C.seg.ensure(nseg + 1, s);
    IDP_LAUNCH(k_coarse_segs, grid_for(nb, 256), 256, 0, s, nb, head, hs, seg);
    double* Ac = C.Ainv.ensure((size_t)ncd * ncd, s);
    IDP_LAUNCH(k_coarse_fill, grid_for((long)ncd * ncd, 256), 256, 0, s, (long)ncd * ncd, Ac, 0.0);
    IDP_LAUNCH(k_coarse_ac, grid_for((long)nseg * 64, 256), 256, 0, s, nseg, seg, nrun, k1, v1, ne, L.erow.p,
               m.e_col.p, L.evals.p, L.crow.p, L.ccol.p, L.cvals.p, nc, d_P, K, ncd, Ac);
    IDP_LAUNCH(k_coarse_diag, grid_for(ncd, 256), 256, 0, s, ncd, Ac);
    // inverse by Cholesky
    if (!C.handle) {
        rocblas_handle h;
        if (rocblas_create_handle(&h) != rocblas_status_success) throw HipError(IDP_E_HIP, "rocblas_create_handle");
        C.handle = h;
    }
    rocblas_handle h = (rocblas_handle)C.handle;
    rocblas_set_stream(h, s);
    int* info = C.info.ensure(2, s);
    if (rocsolver_dpotrf(h, rocblas_fill_lower, ncd, Ac, ncd, info) != rocblas_status_success)
        throw HipError(IDP_E_HIP, "rocsolver_dpotrf");
    if (rocsolver_dpotri(h, rocblas_fill_lower, ncd, Ac, ncd, info + 1) != rocblas_status_success)
        throw HipError(IDP_E_HIP, "rocsolver_dpotri");
    int hinfo[2] = {0, 0};
    IDP_HIP(hipMemcpyAsync(hinfo, info, 2 * sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    if (hinfo[0] != 0 || hinfo[1] != 0) return false;  // not positive definite: no coarse level this time
    IDP_LAUNCH(k_coarse_sym, grid_for((long)ncd * ncd, 256), 256, 0, s, ncd, Ac);
    C.rc.ensure(ncd, s);
    C.nsb = K;  // one partial of r_c . e_c per aggregate
    C.part.ensure(2 * (size_t)C.nsb + 2, s);
    C.ready = true;
    C.age = 1;
    C.nV = nV;
    C.mask = m.proj_mask();
    return true;
}

void coarse_apply(LinWork& L, hipStream_t s, int nV, const double* r, double* z, const int* flag, int parity)
{
    (void)nV;
    CoarseWork& C = L.coarse;
    IDP_LAUNCH(k_coarse_restrict, C.K, COARSE_BLOCK, 0, s, C.K, C.aptr.p, C.avert.p, C.Ps.p, r, flag, C.rc.p);
    IDP_LAUNCH(k_coarse_solve, C.K, COARSE_BLOCK, 0, s, C.K, C.ncd, C.aptr.p, C.avert.p, C.Ps.p, C.Ainv.p, C.rc.p,
               flag, C.part.p + (size_t)parity * C.nsb, z);
}

}  // namespace idp
