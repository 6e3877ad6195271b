// Boundary primitives of a tet mesh on the device, in the reference's order:
//   triangles  Find_Surface_TriMesh<T,false> (Library/Utils/MESHIO.h:486-545):
//              the tet faces (0,2,1) (0,3,2) (0,1,3) (1,2,3) that have no face
//              of the opposite orientation over the same three vertices,
//              listed in lexicographic order of the stored vertex triple
//              (std::map<Triplet> order, MESHIO.h:450-484);
//   edges      Find_Surface_Primitives_And_Compute_Area (MESHIO.h:769-835):
//              each undirected edge once, in the orientation of its first
//              occurrence when the triangles are walked in order with their
//              edges (0,1) (1,2) (2,0), listed in lexicographic order of that
//              oriented pair (std::map<VECTOR<int,2>> order);
//   nodes      vertices of a triangle of nonzero area at X, ascending.
// Keys: a face is the 64-bit word a<<42 | b<<21 | c (21 bits per vertex id);
// its vertex set with the orientation parity of the stored order is
// (sorted triple)<<1 | parity, so that both orientations of one set sort
// next to each other.  Edges sort on (undirected pair, occurrence index).
#include "common.hpp"
#include "kernels.hpp"
#include "prim.hpp"

namespace idp {

namespace {

constexpr int VB = 21;  // bits per vertex id in a face key
constexpr unsigned long long VM = (1ull << VB) - 1;

__device__ inline unsigned long long fkey(int a, int b, int c)
{
    return ((unsigned long long)a << (2 * VB)) | ((unsigned long long)b << VB) | (unsigned long long)c;
}

// every tet face: key of its vertex set with parity, and its oriented key
__global__ void k_tet_faces(int nT, const int4* __restrict__ tets, unsigned long long* __restrict__ setkey,
                            unsigned long long* __restrict__ okey)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 4L * nT) return;
    const int4 t = tets[i >> 2];
    const int f = (int)(i & 3);
    int a, b, c;
    if (f == 0) { a = t.x; b = t.z; c = t.y; }
    else if (f == 1) { a = t.x; b = t.w; c = t.z; }
    else if (f == 2) { a = t.x; b = t.y; c = t.w; }
    else { a = t.y; b = t.z; c = t.w; }
    okey[i] = fkey(a, b, c);
    // sort (a, b, c) with a parity count of the swaps
    int u = a, v = b, w = c, par = 0;
    if (u > v) { int q = u; u = v; v = q; par ^= 1; }
    if (v > w) { int q = v; v = w; w = q; par ^= 1; }
    if (u > v) { int q = u; u = v; v = q; par ^= 1; }
    setkey[i] = (fkey(u, v, w) << 1) | (unsigned long long)par;
}

// a face is on the boundary when no face over the same vertex set has the
// other parity; faces with equal oriented keys are listed once
__global__ void k_boundary_faces(long n, const unsigned long long* __restrict__ setkey,
                                 const unsigned long long* __restrict__ okey, unsigned* __restrict__ cnt,
                                 unsigned long long* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long s = setkey[i] >> 1;
    const unsigned p = (unsigned)(setkey[i] & 1);
    bool keep = true;
    for (long j = i - 1; j >= 0 && (setkey[j] >> 1) == s; --j)
        if ((unsigned)(setkey[j] & 1) != p || okey[j] == okey[i]) keep = false;
    for (long j = i + 1; j < n && (setkey[j] >> 1) == s; ++j)
        if ((unsigned)(setkey[j] & 1) != p) keep = false;
    if (keep) out[atomicAdd(cnt, 1u)] = okey[i];
}

// triangles from their sorted keys; per triangle edge: (undirected pair,
// occurrence index 3 f + j) as one sort key; nodes of nonzero-area triangles
__global__ void k_tri_out(int nf, const unsigned long long* __restrict__ tkey, const double* __restrict__ X,
                          int3* __restrict__ tri, unsigned long long* __restrict__ ekey,
                          unsigned char* __restrict__ node)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= nf) return;
    const unsigned long long k = tkey[f];
    const int a = (int)(k >> (2 * VB)), b = (int)((k >> VB) & VM), c = (int)(k & VM);
    tri[f] = make_int3(a, b, c);
    const int e0[3] = {a, b, c}, e1[3] = {b, c, a};
    for (int j = 0; j < 3; ++j) {
        const unsigned lo = (unsigned)min(e0[j], e1[j]), hi = (unsigned)max(e0[j], e1[j]);
        ekey[3L * f + j] = ((unsigned long long)lo << 43) | ((unsigned long long)hi << 22) | (unsigned long long)(3 * f + j);
    }
    // BTArea = 0.5 |(v1 - v0) x (v2 - v0)|; a node counts when its area sum is nonzero
    const double ux = X[3 * b] - X[3 * a], uy = X[3 * b + 1] - X[3 * a + 1], uz = X[3 * b + 2] - X[3 * a + 2];
    const double vx = X[3 * c] - X[3 * a], vy = X[3 * c + 1] - X[3 * a + 1], vz = X[3 * c + 2] - X[3 * a + 2];
    const double nx = uy * vz - uz * vy, ny = uz * vx - ux * vz, nz = ux * vy - uy * vx;
    if ((0.5 * sqrt(nx * nx + ny * ny + nz * nz)) / 3 != 0.0) {
        node[a] = 1;
        node[b] = 1;
        node[c] = 1;
    }
}

// the first occurrence of each undirected edge gives its orientation
__global__ void k_edge_first(long n, const unsigned long long* __restrict__ ekey, const int3* __restrict__ tri,
                             unsigned* __restrict__ cnt, unsigned long long* __restrict__ out)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const unsigned long long pair = ekey[i] >> 22;
    if (i > 0 && (ekey[i - 1] >> 22) == pair) return;
    const unsigned occ = (unsigned)(ekey[i] & ((1ull << 22) - 1));
    const int3 t = tri[occ / 3];
    const int j = (int)(occ % 3);
    const int a = j == 0 ? t.x : (j == 1 ? t.y : t.z);
    const int b = j == 0 ? t.y : (j == 1 ? t.z : t.x);
    out[atomicAdd(cnt, 1u)] = ((unsigned long long)a << VB) | (unsigned long long)b;
}

__global__ void k_edge_out(int ne, const unsigned long long* __restrict__ k, int2* __restrict__ e)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ne) return;
    e[i] = make_int2((int)(k[i] >> VB), (int)(k[i] & VM));
}

__global__ void k_node_out(int nV, const unsigned char* __restrict__ node, const int* __restrict__ pos,
                           int* __restrict__ bn)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nV || !node[v]) return;
    bn[pos[v]] = v;
}

__global__ void k_u8_to_int(int n, const unsigned char* __restrict__ a, int* __restrict__ o)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) o[i] = (i < n && a[i]) ? 1 : 0;
}

unsigned read_count(hipStream_t s, const unsigned* p)
{
    unsigned v = 0;
    IDP_HIP(hipMemcpyAsync(&v, p, sizeof(v), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    return v;
}

}  // namespace

bool surface_device(hipStream_t s, SurfWork& w, int nV, const double* X, int nT, const int* tets,
                    std::vector<int>& bn, std::vector<std::array<int, 2>>& be, std::vector<std::array<int, 3>>& bt)
{
    // 21-bit vertex ids in the face keys
    if (nV <= 0 || nT <= 0 || nV > (1 << VB)) return false;
    const long nfc = 4L * nT;
    unsigned* cnt = w.cnt.ensure(4, s);
    IDP_HIP(hipMemsetAsync(cnt, 0, 4 * sizeof(unsigned), s));
    unsigned long long* sk0 = w.k0.ensure(nfc, s);
    unsigned long long* sk1 = w.k1.ensure(nfc, s);
    unsigned long long* ok0 = w.v0.ensure(nfc, s);
    unsigned long long* ok1 = w.v1.ensure(nfc, s);
    IDP_LAUNCH(k_tet_faces, grid_for(nfc, 256), 256, 0, s, nT, (const int4*)tets, sk0, ok0);
    sort_pairs(s, w.tmp, sk0, sk1, ok0, ok1, nfc, 64);
    // boundary faces (oriented keys, any order), then in key order
    IDP_LAUNCH(k_boundary_faces, grid_for(nfc, 256), 256, 0, s, nfc, sk1, ok1, cnt, sk0);
    const int nf = (int)read_count(s, cnt);
    if (3L * nf >= (1L << 22)) return false;  // 22-bit edge occurrence indices
    sort_keys(s, w.tmp, sk0, ok0, nf, 63);
    int3* tri = (int3*)w.tri.ensure(3 * (size_t)std::max(nf, 1), s);
    unsigned long long* ek0 = sk0;  // reuse: 3 nf <= 4 nT
    unsigned char* node = w.node.ensure(nV, s);
    IDP_HIP(hipMemsetAsync(node, 0, nV, s));
    if (nf) IDP_LAUNCH(k_tri_out, grid_for(nf, 256), 256, 0, s, nf, ok0, X, tri, ek0, node);
    const long nec = 3L * nf;
    sort_keys(s, w.tmp, ek0, sk1, nec, 64);
    if (nec) IDP_LAUNCH(k_edge_first, grid_for(nec, 256), 256, 0, s, nec, sk1, tri, cnt + 1, ok1);
    const int ne = (int)read_count(s, cnt + 1);
    sort_keys(s, w.tmp, ok1, sk0, ne, 2 * VB);
    int2* edge = (int2*)w.edge.ensure(2 * (size_t)std::max(ne, 1), s);
    if (ne) IDP_LAUNCH(k_edge_out, grid_for(ne, 256), 256, 0, s, ne, sk0, edge);
    // nodes: ascending ids of the flagged vertices
    int* flag = w.flag.ensure(nV + 1, s);
    int* pos = w.pos.ensure(nV + 1, s);
    IDP_LAUNCH(k_u8_to_int, grid_for(nV + 1, 256), 256, 0, s, nV, node, flag);
    excl_scan(s, w.tmp, flag, pos, nV + 1);
    int nb = 0;
    IDP_HIP(hipMemcpyAsync(&nb, pos + nV, sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    int* bnd = w.bn.ensure(std::max(nb, 1), s);
    IDP_LAUNCH(k_node_out, grid_for(nV, 256), 256, 0, s, nV, node, pos, bnd);
    bn.resize(nb);
    be.resize(ne);
    bt.resize(nf);
    if (nb) IDP_HIP(hipMemcpyAsync(bn.data(), bnd, nb * sizeof(int), hipMemcpyDeviceToHost, s));
    if (ne) IDP_HIP(hipMemcpyAsync(be.data(), edge, 2 * (size_t)ne * sizeof(int), hipMemcpyDeviceToHost, s));
    if (nf) IDP_HIP(hipMemcpyAsync(bt.data(), tri, 3 * (size_t)nf * sizeof(int), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    return true;
}

}  // namespace idp
