// Lagged IPC friction on the device (Library/FEM/FRICTION.h:16-660, the
// smoothing functions, closest points, tangent bases and TT maps of
// Library/FEM/FRICTION_UTILS.h:9-287).
//
// Basis update (Compute_Friction_Basis): the friction constraints are the
// barrier constraints that are not mollified, in the set's order (a stable
// compaction: flag, exclusive scan, scatter).  Each keeps its int4 entry and
// a 10-double record: closest-point parameters (2), tangent basis (column 0,
// column 1), normal force -b'(d - t^2) 2 sqrt(d) (times the per-component
// coefficient when those are given), one pad.
//
// Energy, gradient and Hessian read the record and the displacement
// x - x_n of the constraint's vertices, one constraint per lane.  The
// relative displacement is sum_a c_a dx_a with the TT coefficients c of the
// kind, so every Hessian block is c_a c_b B W B^T with B the 3 x 2 tangent
// basis and W the 2 x 2 inner matrix (rank one above the static bound,
// f1 I at rest, the clamped (f2/|u|) u u^T + f1 I between): the kernel forms
// the 3 x 3 product once and scales it per block.  The reference's
// point-point relative displacement reads the second point's position, not
// its displacement (FRICTION.h:212-214, 318-320, 479-481), and so does this.
#include "common.hpp"
#include "constraint_dev.hpp"
#include "geom_dev.hpp"
#include "kernels.hpp"
#include "prim.hpp"

namespace idp {

namespace {

// Eigen's normalized(): v / sqrt(|v|^2) when |v|^2 > 0, else v
__device__ inline d3 nrmz(d3 v)
{
    const double n = sqn(v);
    if (n > 0) {
        const double r = sqrt(n);
        return {v.x / r, v.y / r, v.z / r};
    }
    return v;
}

// FRICTION_UTILS.h:9-36 (C1 clamping)
__device__ inline double f0_sf(double x2, double epsvh)
{
    if (x2 >= epsvh * epsvh) return sqrt(x2);
    return x2 * (-sqrt(x2) / 3.0 + epsvh) / (epsvh * epsvh) + epsvh / 3.0;
}
__device__ inline double f1_div(double x2, double epsvh)
{
    if (x2 >= epsvh * epsvh) return 1 / sqrt(x2);
    return (-sqrt(x2) + 2.0 * epsvh) / (epsvh * epsvh);
}

// TT coefficients per kind (FRICTION_UTILS.h:92-102, 176-186, 235-244, 281-287)
__device__ inline void tt_coef(int kind, double b1, double b2, double* c)
{
    if (kind == K_PT) {
        c[0] = 1; c[1] = -1 + b1 + b2; c[2] = -b1; c[3] = -b2;
    } else if (kind == K_EE) {
        c[0] = 1.0 - b1; c[1] = b1; c[2] = b2 - 1.0; c[3] = -b2;
    } else if (kind == K_PE) {
        c[0] = 1; c[1] = b1 - 1.0; c[2] = -b1; c[3] = 0;
    } else {
        c[0] = 1; c[1] = -1; c[2] = 0; c[3] = 0;
    }
}

struct FricLoc {
    Decoded k;
    double cp0, cp1;
    d3 b0, b1;  // tangent basis columns
    double N;
    double u0, u1;  // tangential relative displacement
};

__device__ inline FricLoc fric_load(long i, const int4* __restrict__ fcs, const double* __restrict__ fdat,
                                    const double* __restrict__ X, const double* __restrict__ Xn)
{
    FricLoc f;
    f.k = decode(fcs[i]);
    const double2* r = reinterpret_cast<const double2*>(fdat + 10 * i);
    const double2 q0 = r[0], q1 = r[1], q2 = r[2], q3 = r[3], q4 = r[4];
    f.cp0 = q0.x;
    f.cp1 = q0.y;
    f.b0 = {q1.x, q1.y, q2.x};
    f.b1 = {q2.y, q3.x, q3.y};
    f.N = q4.x;
    const int* v = f.k.v;
    auto dx = [&](int a) { return ld3(X, v[a]) - ld3(Xn, v[a]); };
    d3 rel;
    if (f.k.kind == K_PT) {
        const d3 d0 = dx(0), d1 = dx(1), d2 = dx(2), d3v = dx(3);
        rel = d0 - ((d1 + f.cp0 * (d2 - d1)) + f.cp1 * (d3v - d1));
    } else if (f.k.kind == K_EE) {
        const d3 d0 = dx(0), d1 = dx(1), d2 = dx(2), d3v = dx(3);
        rel = (d0 + f.cp0 * (d1 - d0)) - (d2 + f.cp1 * (d3v - d2));
    } else if (f.k.kind == K_PE) {
        const d3 d0 = dx(0), d1 = dx(1), d2 = dx(2);
        rel = d0 - (d1 + f.cp0 * (d2 - d1));
    } else {
        rel = dx(0) - ld3(X, v[1]);  // the reference reads x_q here
    }
    f.u0 = rel.x * f.b0.x + rel.y * f.b0.y + rel.z * f.b0.z;
    f.u1 = rel.x * f.b1.x + rel.y * f.b1.y + rel.z * f.b1.z;
    return f;
}

// 2 x 2 spectral PSD clamp (Library/Math/UTILS.h:9-27 on a 2 x 2): unchanged
// when the smaller eigenvalue is >= 0
__device__ inline void make_pd2(double& p, double& q, double& r)
{
    const double m = 0.5 * (p + r), h = 0.5 * (p - r);
    const double d = sqrt(h * h + q * q);
    const double l1 = m - d, l2 = m + d;
    if (l1 >= 0) return;
    if (l2 <= 0) {
        p = q = r = 0.0;
        return;
    }
    // eigenvector of l2: (q, l2 - p) or (l2 - r, q), whichever is longer
    double ex = q, ey = l2 - p;
    const double fx = l2 - r, fy = q;
    if (fx * fx + fy * fy > ex * ex + ey * ey) {
        ex = fx;
        ey = fy;
    }
    const double n2 = ex * ex + ey * ey;
    if (!(n2 > 0)) {  // q == 0 and p == l2: already diagonal
        p = p > 0 ? p : 0.0;
        r = r > 0 ? r : 0.0;
        return;
    }
    p = l2 * ex * ex / n2;
    q = l2 * ex * ey / n2;
    r = l2 * ey * ey / n2;
}

}  // namespace

// flag[i] = 1 for a constraint that carries friction (not mollified:
// FRICTION.h:38-42); flag[n] = 0 so the exclusive scan's last entry counts them
__global__ void k_fric_flag(long n, const int4* __restrict__ cs, unsigned* __restrict__ flag)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i > n) return;
    if (i == n) {
        flag[n] = 0;
        return;
    }
    const int4 c = cs[i];
    flag[i] = (c.x >= 0 && (c.z < 0 || c.w < 0)) ? 0u : 1u;
}

// Compute_Friction_Basis (FRICTION.h:16-117) and Compute_Friction_Coef
// (FRICTION.h:119-169) for the flagged constraints, written at their scan
// position
__global__ void k_fric_basis(long n, const int4* __restrict__ cs, const unsigned* __restrict__ flag,
                             const unsigned* __restrict__ pos, const double* __restrict__ X, double dh, double th2,
                             double kappa, int ncomp, const int* __restrict__ comp_range,
                             const double* __restrict__ mu_comp, int4* __restrict__ fcs, double* __restrict__ fdat,
                             int* __restrict__ err)
{
    long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    const int4 c = cs[i];
    const Decoded k = decode(c);
    const int* v = k.v;
    double q0 = 0.0, q1 = 0.0;
    d3 c0, c1;
    if (k.kind == K_EE) {
        const d3 a0 = ld3(X, v[0]), a1 = ld3(X, v[1]), b0 = ld3(X, v[2]), b1 = ld3(X, v[3]);
        const d3 e20 = a0 - b0, e01 = a1 - a0, e23 = b1 - b0;
        ldlt2(dot(e01, e01), -dot(e23, e01), dot(e23, e23), -dot(e20, e01), dot(e20, e23), q0, q1);
        c0 = nrmz(e01);
        c1 = nrmz(cross(cross(e01, e23), e01));
    } else if (k.kind == K_PT) {
        const d3 p = ld3(X, v[0]), t0 = ld3(X, v[1]), t1 = ld3(X, v[2]), t2 = ld3(X, v[3]);
        const d3 r0 = t1 - t0, r1 = t2 - t0, pd = p - t0;
        ldlt2(dot(r0, r0), dot(r0, r1), dot(r1, r1), dot(r0, pd), dot(r1, pd), q0, q1);
        c0 = nrmz(r0);
        c1 = nrmz(cross(cross(r0, r1), r0));
    } else if (k.kind == K_PE) {
        const d3 p = ld3(X, v[0]), e0 = ld3(X, v[1]), e1 = ld3(X, v[2]);
        const d3 e = e1 - e0;
        q0 = dot(p - e0, e) / sqn(e);
        c0 = nrmz(e);
        c1 = nrmz(cross(e, p - e0));
    } else {
        const d3 p = ld3(X, v[0]), p1 = ld3(X, v[1]);
        const d3 w = p1 - p;
        const d3 xc{0.0, -w.z, w.y}, yc{w.z, 0.0, -w.x};  // UnitX x w, UnitY x w
        const d3 r = sqn(xc) > sqn(yc) ? xc : yc;
        c0 = nrmz(r);
        c1 = nrmz(cross(w, r));
    }
    const double d2 = constraint_d2(k, X);
    double N = -barrier_db(d2 - th2, dh, kappa) * 2 * sqrt(d2);
    if (ncomp > 0) {
        const int va = v[0], vb = (k.kind == K_EE) ? v[2] : v[1];
        int ca = -1, cb = -1;
        for (int q = ncomp - 1; q >= 0; --q) {
            if (va < comp_range[q]) ca = q;
            if (vb < comp_range[q]) cb = q;
        }
        if (ca < 0 || cb < 0) {
            err[0] = 1;
            N = 0.0;
        } else {
            N *= mu_comp[ca + cb * ncomp];
        }
    }
    const long j = pos[i];
    fcs[j] = c;
    double2* o = reinterpret_cast<double2*>(fdat + 10 * j);
    o[0] = make_double2(q0, q1);
    o[1] = make_double2(c0.x, c0.y);
    o[2] = make_double2(c0.z, c1.x);
    o[3] = make_double2(c1.y, c1.z);
    o[4] = make_double2(N, 0.0);
}

// Compute_Friction_Potential (FRICTION.h:171-243): per-block partial sums of
// N f0(|u|) (times the multiplicity above 1); mu is applied by the caller
__global__ void __launch_bounds__(RB) k_fric_energy(long nf, const int4* __restrict__ fcs,
                                                    const double* __restrict__ fdat, const double* __restrict__ X,
                                                    const double* __restrict__ Xn, double epsvh,
                                                    double* __restrict__ part)
{
    __shared__ double sh[RB];
    const long i = (long)blockIdx.x * RB + threadIdx.x;
    double e = 0.0;
    if (i < nf) {
        const FricLoc f = fric_load(i, fcs, fdat, X, Xn);
        e = f0_sf(f.u0 * f.u0 + f.u1 * f.u1, epsvh);
        e *= f.N;
        if (f.k.mult > 1.0) e *= f.k.mult;
    }
    const double s = block_sum<RB>(e, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// Compute_Friction_Gradient (FRICTION.h:245-383): 12 doubles per constraint,
// vertex slots in decode order
__global__ void k_fric_grad(long nf, const int4* __restrict__ fcs, const double* __restrict__ fdat,
                            const double* __restrict__ X, const double* __restrict__ Xn, double epsvh, double mu,
                            double* __restrict__ gout)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nf) return;
    const FricLoc f = fric_load(i, fcs, fdat, X, Xn);
    const double f1 = f1_div(f.u0 * f.u0 + f.u1 * f.u1, epsvh);
    const bool pm = f.k.kind == K_PE || f.k.kind == K_PP;
    const double sc = pm ? f1 * f.k.mult * mu * f.N : f1 * mu * f.N;
    const double u0 = f.u0 * sc, u1 = f.u1 * sc;
    const d3 t{f.b0.x * u0 + f.b1.x * u1, f.b0.y * u0 + f.b1.y * u1, f.b0.z * u0 + f.b1.z * u1};
    double c[4];
    tt_coef(f.k.kind, f.cp0, f.cp1, c);
    double* o = gout + 12 * i;
    for (int a = 0; a < f.k.nv; ++a) {
        const d3 ta = (f.k.kind == K_EE || a > 0) ? c[a] * t : t;
        o[3 * a] = ta.x;
        o[3 * a + 1] = ta.y;
        o[3 * a + 2] = ta.z;
    }
}

// Compute_Friction_Hessian (FRICTION.h:385-660) with projectSPD: the blocks
// c_a c_b B W B^T.  mode 1: 3 x 3 blocks (a, b) at out + 144 i + 9 (a nv + b)
// (the contact assembly's layout); mode 0: the dense (3 nv)^2 row-major
// matrix at out + 144 i, zero padded.  A team of FT lanes per constraint:
// each lane forms W and M = B W B^T from the (shared, cached) record and
// writes every FT-th double of the constraint's 144, so a wavefront's
// stores cover whole contiguous runs.
constexpr int FT = 16;

__global__ void __launch_bounds__(256) k_fric_hess(long nf, const int4* __restrict__ fcs,
                                                   const double* __restrict__ fdat, const double* __restrict__ X,
                                                   const double* __restrict__ Xn, double epsvh, double mu,
                                                   int project, int mode, double* __restrict__ out)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long i = t / FT;
    const int tl = (int)(t % FT);
    if (i >= nf) return;
    const FricLoc f = fric_load(i, fcs, fdat, X, Xn);
    const double x2 = f.u0 * f.u0 + f.u1 * f.u1, xn = sqrt(x2);
    const double f1 = f1_div(x2, epsvh), f2 = -1 / (epsvh * epsvh);
    const bool pm = f.k.kind == K_PE || f.k.kind == K_PP;
    const double s = pm ? f.k.mult * mu * f.N : mu * f.N;
    double w00, w01, w11;  // W, symmetric
    if (x2 >= epsvh * epsvh) {  // rank one along ubar = (-u1, u0)
        const double sc = s * f1 / x2;
        w00 = sc * f.u1 * f.u1;
        w01 = -sc * f.u1 * f.u0;
        w11 = sc * f.u0 * f.u0;
    } else if (xn == 0) {
        w00 = w11 = s * f1;
        w01 = 0.0;
    } else {
        const double g = f2 / xn;
        double p = g * f.u0 * f.u0 + f1, q = g * f.u0 * f.u1, r = g * f.u1 * f.u1 + f1;
        if (project) make_pd2(p, q, r);
        w00 = s * p;
        w01 = s * q;
        w11 = s * r;
    }
    // M = B W B^T
    const double bc[2][3] = {{f.b0.x, f.b0.y, f.b0.z}, {f.b1.x, f.b1.y, f.b1.z}};
    double M[9];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        const double l0 = bc[0][r] * w00 + bc[1][r] * w01;
        const double l1 = bc[0][r] * w01 + bc[1][r] * w11;
#pragma unroll
        for (int cc = 0; cc < 3; ++cc) M[3 * r + cc] = l0 * bc[0][cc] + l1 * bc[1][cc];
    }
    double c[4];
    tt_coef(f.k.kind, f.cp0, f.cp1, c);
    const int nv = f.k.nv, n3 = 3 * nv, nval = 9 * nv * nv;
    double* o = out + 144 * i;
    for (int e = tl; e < 144; e += FT) {
        double val = 0.0;
        if (e < nval) {
            int a, b, r, cc;
            if (mode == 1) {
                const int bk = e / 9, q = e - 9 * bk;
                a = bk / nv;
                b = bk - a * nv;
                r = q / 3;
                cc = q - 3 * r;
            } else {
                const int row = e / n3, col = e - row * n3;
                a = row / 3;
                r = row - 3 * a;
                b = col / 3;
                cc = col - 3 * b;
            }
            double ca = c[0], cb = c[0];
#pragma unroll
            for (int k = 1; k < 4; ++k) {
                ca = (a == k) ? c[k] : ca;
                cb = (b == k) ? c[k] : cb;
            }
            double m = M[0];
#pragma unroll
            for (int k = 1; k < 9; ++k) m = (3 * r + cc == k) ? M[k] : m;
            val = (ca * cb) * m;
        } else if (mode == 1) {
            break;  // the assembly reads only the nv^2 blocks
        }
        o[e] = val;
    }
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
long fric_basis_update(hipStream_t s, DBuf<unsigned char>& tmp, long n, const int* cs, const double* X, double dh,
                       double th2, double kappa, int ncomp, const int* comp_range, const double* mu_comp,
                       DBuf<unsigned>& flag, DBuf<unsigned>& pos, DBuf<int4>& fcs, DBuf<double>& fdat, int* err)
{
    if (n == 0) return 0;
    unsigned* fl = flag.ensure(n + 1, s);
    unsigned* ps = pos.ensure(n + 1, s);
    IDP_LAUNCH(k_fric_flag, grid_for(n + 1, 256), 256, 0, s, n, (const int4*)cs, fl);
    excl_scan<unsigned>(s, tmp, fl, ps, n + 1);
    unsigned nf = 0;
    IDP_HIP(hipMemcpyAsync(&nf, ps + n, sizeof(unsigned), hipMemcpyDeviceToHost, s));
    IDP_HIP(hipStreamSynchronize(s));
    int4* fc = fcs.ensure(nf + 1, s);
    double* fd = fdat.ensure(10 * (size_t)nf + 10, s);
    IDP_LAUNCH(k_fric_basis, grid_for(n, 256), 256, 0, s, n, (const int4*)cs, fl, ps, X, dh, th2, kappa, ncomp,
               comp_range, mu_comp, fc, fd, err);
    return (long)nf;
}

void launch_fric_energy(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                        const double* Xn, double epsvh, double* part)
{
    IDP_LAUNCH(k_fric_energy, grid_for(nf, RB), RB, 0, s, nf, fcs, fdat, X, Xn, epsvh, part);
}

void launch_fric_grad(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                      const double* Xn, double epsvh, double mu, double* gout)
{
    IDP_LAUNCH(k_fric_grad, grid_for(nf, 256), 256, 0, s, nf, fcs, fdat, X, Xn, epsvh, mu, gout);
}

void launch_fric_hess(hipStream_t s, long nf, const int4* fcs, const double* fdat, const double* X,
                      const double* Xn, double epsvh, double mu, int project, int mode, double* out)
{
    IDP_LAUNCH(k_fric_hess, grid_for(nf * FT, 256), 256, 0, s, nf, fcs, fdat, X, Xn, epsvh, mu, project, mode, out);
}

}  // namespace idp
