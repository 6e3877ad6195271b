// Oracle (test infrastructure only): CPU restatement of the IDP geometric
// predicates, squared distances, IPC barrier, edge-edge mollifier and the
// spectral PSD clamp.  Nothing in the product links this file; it is the
// checker the HIP path is compared against.
//
// Floating-point policy: compiled with -ffp-contract=off; every sum is
// evaluated left to right exactly as written.  The HIP kernels use the same
// order for every *decision* (type classification, AABB tests, thresholds,
// ACCD), so pair sets and CCD step sizes agree bit for bit.  Where the
// reference's rounding depends on Eigen's expression evaluation (pivoted
// 2x2 LDLT, dot-product association, GCC -mfma contraction) the oracle fixes
// one order; ties at the last ulp are "parity unpinned" against the
// reference binary (which cannot be built here).
#pragma once
#include <cmath>
#include <cfloat>
#include <cstring>
#include <algorithm>

namespace orc {

struct V3 {
    double x, y, z;
};
static inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
static inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double sqn(V3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
static inline V3 cross(V3 a, V3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}

// ---------------------------------------------------------------------------
// Distance-type classification
// ---------------------------------------------------------------------------

// Library/Math/Distance/DISTANCE_TYPE.h:11-27
static inline int pe_type(V3 p, V3 e0, V3 e1)
{
    V3 e = e1 - e0;
    double ratio = dot(e, p - e0) / sqn(e);
    if (ratio < 0) return 0;
    if (ratio > 1) return 1;
    return 2;
}

// Solve the 2x2 symmetric system [[m00 m01][m01 m11]] x = r the way Eigen's
// pivoted LDLT (lower, diagonal pivoting on the largest |d|, first index on
// ties) does for a 2x2: this is what DISTANCE_TYPE.h:46,54,62 call.
static inline void ldlt2_solve(double m00, double m01, double m11, double r0, double r1,
                               double& x0, double& x1)
{
    const bool swap = std::fabs(m11) > std::fabs(m00);
    const double a = swap ? m11 : m00;
    const double c = swap ? m00 : m11;
    const double b0 = swap ? r1 : r0;
    const double b1 = swap ? r0 : r1;
    if (!(std::fabs(a) > 0.0)) {  // whole diagonal zero: Eigen leaves L = I, D = 0
        x0 = 0.0;
        x1 = 0.0;
        return;
    }
    const double l = m01 / a;
    const double d1 = c - l * (a * l);
    double y0 = b0;
    double y1 = b1 - l * y0;
    y0 = (std::fabs(a) > DBL_MIN) ? y0 / a : 0.0;
    y1 = (std::fabs(d1) > DBL_MIN) ? y1 / d1 : 0.0;
    const double z1 = y1;
    const double z0 = y0 - l * z1;
    x0 = swap ? z1 : z0;
    x1 = swap ? z0 : z1;
}

// Library/Math/Distance/DISTANCE_TYPE.h:30-82
// 0,1,2: PP with t0,t1,t2; 3,4,5: PE with t0t1, t1t2, t2t0; 6: PT
static inline int pt_type(V3 p, V3 t0, V3 t1, V3 t2)
{
    V3 b0 = t1 - t0;
    V3 b1 = t2 - t0;
    const V3 n = cross(b0, b1);
    double s0[3], s1[3];
    {
        V3 r = cross(b0, n);
        V3 q = p - t0;
        ldlt2_solve(dot(b0, b0), dot(b0, r), dot(r, r), dot(b0, q), dot(r, q), s0[0], s1[0]);
        if (s0[0] > 0.0 && s0[0] < 1.0 && s1[0] >= 0.0) return 3;
    }
    {
        V3 e = t2 - t1;
        V3 r = cross(e, n);
        V3 q = p - t1;
        ldlt2_solve(dot(e, e), dot(e, r), dot(r, r), dot(e, q), dot(r, q), s0[1], s1[1]);
        if (s0[1] > 0.0 && s0[1] < 1.0 && s1[1] >= 0.0) return 4;
    }
    {
        V3 e = t0 - t2;
        V3 r = cross(e, n);
        V3 q = p - t2;
        ldlt2_solve(dot(e, e), dot(e, r), dot(r, r), dot(e, q), dot(r, q), s0[2], s1[2]);
        if (s0[2] > 0.0 && s0[2] < 1.0 && s1[2] >= 0.0) return 5;
    }
    if (s0[0] <= 0.0 && s0[2] >= 1.0) return 0;
    if (s0[1] <= 0.0 && s0[0] >= 1.0) return 1;
    if (s0[2] <= 0.0 && s0[1] >= 1.0) return 2;
    return 6;
}

// Library/Math/Distance/DISTANCE_TYPE.h:84-165
// 0: a0-b0, 1: a0-b1, 2: a0-(b0b1), 3: a1-b0, 4: a1-b1, 5: a1-(b0b1),
// 6: b0-(a0a1), 7: b1-(a0a1), 8: line-line
static inline int ee_type(V3 ea0, V3 ea1, V3 eb0, V3 eb1)
{
    V3 u = ea1 - ea0, v = eb1 - eb0, w = ea0 - eb0;
    double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    double tD = D, sN, tN;
    int def = 8;
    sN = b * e - c * d;
    if (sN <= 0.0) {
        tN = e;
        tD = c;
        def = 2;
    } else if (sN >= D) {
        tN = e + b;
        tD = c;
        def = 5;
    } else {
        tN = a * e - b * d;
        if (tN > 0.0 && tN < tD) {
            V3 uv = cross(u, v);
            if (dot(uv, w) == 0.0 || sqn(uv) < 1.0e-20 * a * c) {
                if (sN < D / 2) {
                    tN = e;
                    tD = c;
                    def = 2;
                } else {
                    tN = e + b;
                    tD = c;
                    def = 5;
                }
            }
        }
    }
    if (tN <= 0.0) {
        if (-d <= 0.0) return 0;
        if (-d >= a) return 3;
        return 6;
    } else if (tN >= tD) {
        if ((-d + b) <= 0.0) return 1;
        if ((-d + b) >= a) return 4;
        return 7;
    }
    return def;
}

// ---------------------------------------------------------------------------
// Squared distances (POINT_POINT.h:12, POINT_EDGE.h:12-26, POINT_TRIANGLE.h:12-22,
// EDGE_EDGE.h:12-23)
// ---------------------------------------------------------------------------
static inline double d2_pp(V3 a, V3 b) { return sqn(a - b); }
static inline double d2_pe(V3 p, V3 e0, V3 e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); }
static inline double d2_pt(V3 p, V3 t0, V3 t1, V3 t2)
{
    V3 b = cross(t1 - t0, t2 - t0);
    double s = dot(p - t0, b);
    return s * s / sqn(b);
}
static inline double d2_ee(V3 a0, V3 a1, V3 b0, V3 b1)
{
    V3 b = cross(a1 - a0, b1 - b0);
    double s = dot(b0 - a0, b);
    return s * s / sqn(b);
}
static inline double ee_cross_sqn(V3 a0, V3 a1, V3 b0, V3 b1) { return sqn(cross(a1 - a0, b1 - b0)); }
// EDGE_EDGE_MOLLIFIER.h:583-592
static inline double mollifier_eps(V3 a0r, V3 a1r, V3 b0r, V3 b1r)
{
    return 1.0e-3 * sqn(a0r - a1r) * sqn(b0r - b1r);
}

// DISTANCE_UNCLASSIFIED.h:15-55
static inline double d2_pt_unclassified(V3 p, V3 t0, V3 t1, V3 t2)
{
    switch (pt_type(p, t0, t1, t2)) {
    case 0: return d2_pp(p, t0);
    case 1: return d2_pp(p, t1);
    case 2: return d2_pp(p, t2);
    case 3: return d2_pe(p, t0, t1);
    case 4: return d2_pe(p, t1, t2);
    case 5: return d2_pe(p, t2, t0);
    default: return d2_pt(p, t0, t1, t2);
    }
}
// DISTANCE_UNCLASSIFIED.h:57-120
static inline double d2_ee_unclassified(V3 a0, V3 a1, V3 b0, V3 b1)
{
    switch (ee_type(a0, a1, b0, b1)) {
    case 0: return d2_pp(a0, b0);
    case 1: return d2_pp(a0, b1);
    case 2: return d2_pe(a0, b0, b1);
    case 3: return d2_pp(a1, b0);
    case 4: return d2_pp(a1, b1);
    case 5: return d2_pe(a1, b0, b1);
    case 6: return d2_pe(b0, a0, a1);
    case 7: return d2_pe(b1, a0, a1);
    default: return d2_ee(a0, a1, b0, b1);
    }
}

// ---------------------------------------------------------------------------
// IPC log barrier on squared distance (BARRIER.h:10-62, non-elastic branch):
//   b(d) = -kappa (d - dh)^2 ln(d / dh)
// ---------------------------------------------------------------------------
static inline double barrier_b(double d, double dh, double k)
{
    double t = d - dh;
    return -k * t * t * std::log(d / dh);
}
static inline double barrier_db(double d, double dh, double k)
{
    double t = d - dh;
    return k * (-2.0 * t * std::log(d / dh) - t * t / d);
}
static inline double barrier_d2b(double d, double dh, double k)
{
    double t = d - dh;
    return k * (-2.0 * std::log(d / dh) - 4.0 * t / d + t * t / (d * d));
}
// Mollifier e(x) = (2 - x/eps) x/eps  (EDGE_EDGE_MOLLIFIER.h:441-476)
static inline double moll_e(double x, double eps)
{
    double r = x / eps;
    return (2.0 - r) * r;
}
static inline double moll_de(double x, double eps) { return 2.0 / eps * (1.0 - x / eps); }
static inline double moll_d2e(double, double eps) { return -2.0 / (eps * eps); }

// ---------------------------------------------------------------------------
// Second-order forward-mode jets: value, gradient and Hessian with respect to
// N input coordinates.  The oracle differentiates the closed-form squared
// distances with these instead of hand-written derivative formulas.
// ---------------------------------------------------------------------------
template <int N>
struct Jet {
    double v;
    double g[N];
    double H[N][N];
    Jet() { std::memset(this, 0, sizeof(*this)); }
    static Jet var(double val, int i)
    {
        Jet j;
        j.v = val;
        j.g[i] = 1.0;
        return j;
    }
    static Jet cst(double val)
    {
        Jet j;
        j.v = val;
        return j;
    }
};
template <int N>
Jet<N> operator+(const Jet<N>& a, const Jet<N>& b)
{
    Jet<N> r;
    r.v = a.v + b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] + b.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) r.H[i][j] = a.H[i][j] + b.H[i][j];
    return r;
}
template <int N>
Jet<N> operator-(const Jet<N>& a, const Jet<N>& b)
{
    Jet<N> r;
    r.v = a.v - b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] - b.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) r.H[i][j] = a.H[i][j] - b.H[i][j];
    return r;
}
template <int N>
Jet<N> operator*(const Jet<N>& a, const Jet<N>& b)
{
    Jet<N> r;
    r.v = a.v * b.v;
    for (int i = 0; i < N; ++i) r.g[i] = a.g[i] * b.v + a.v * b.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j)
            r.H[i][j] = a.H[i][j] * b.v + a.v * b.H[i][j] + a.g[i] * b.g[j] + a.g[j] * b.g[i];
    return r;
}
template <int N>
Jet<N> scale(double s, const Jet<N>& a)
{
    Jet<N> r;
    r.v = s * a.v;
    for (int i = 0; i < N; ++i) r.g[i] = s * a.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) r.H[i][j] = s * a.H[i][j];
    return r;
}
// scalar function composition: f(a) with f, f', f''
template <int N>
Jet<N> compose(const Jet<N>& a, double f, double df, double d2f)
{
    Jet<N> r;
    r.v = f;
    for (int i = 0; i < N; ++i) r.g[i] = df * a.g[i];
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) r.H[i][j] = df * a.H[i][j] + d2f * a.g[i] * a.g[j];
    return r;
}
template <int N>
Jet<N> recip(const Jet<N>& a)
{
    double iv = 1.0 / a.v;
    return compose(a, iv, -iv * iv, 2.0 * iv * iv * iv);
}
template <int N>
Jet<N> operator/(const Jet<N>& a, const Jet<N>& b) { return a * recip(b); }

template <int N>
struct JV3 {
    Jet<N> x, y, z;
};
template <int N>
JV3<N> jsub(const JV3<N>& a, const JV3<N>& b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
template <int N>
Jet<N> jdot(const JV3<N>& a, const JV3<N>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <int N>
JV3<N> jcross(const JV3<N>& a, const JV3<N>& b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
template <int N>
JV3<N> jvar(V3 p, int slot)
{
    return {Jet<N>::var(p.x, 3 * slot), Jet<N>::var(p.y, 3 * slot + 1), Jet<N>::var(p.z, 3 * slot + 2)};
}

// Jets of the squared distances, in the argument order of the reference
// functions (p,t0,t1,t2), (ea0,ea1,eb0,eb1), (p,e0,e1), (p0,p1).
static inline Jet<12> jd2_pt(V3 p, V3 t0, V3 t1, V3 t2)
{
    auto P = jvar<12>(p, 0), T0 = jvar<12>(t0, 1), T1 = jvar<12>(t1, 2), T2 = jvar<12>(t2, 3);
    auto b = jcross(jsub(T1, T0), jsub(T2, T0));
    auto s = jdot(jsub(P, T0), b);
    return (s * s) / jdot(b, b);
}
static inline Jet<12> jd2_ee(V3 a0, V3 a1, V3 b0, V3 b1)
{
    auto A0 = jvar<12>(a0, 0), A1 = jvar<12>(a1, 1), B0 = jvar<12>(b0, 2), B1 = jvar<12>(b1, 3);
    auto b = jcross(jsub(A1, A0), jsub(B1, B0));
    auto s = jdot(jsub(B0, A0), b);
    return (s * s) / jdot(b, b);
}
template <int N>
Jet<N> jd2_pe_slots(V3 p, V3 e0, V3 e1, int sp, int s0, int s1)
{
    auto P = jvar<N>(p, sp), E0 = jvar<N>(e0, s0), E1 = jvar<N>(e1, s1);
    auto c = jcross(jsub(E0, P), jsub(E1, P));
    auto e = jsub(E1, E0);
    return jdot(c, c) / jdot(e, e);
}
template <int N>
Jet<N> jd2_pp_slots(V3 a, V3 b, int sa, int sb)
{
    auto A = jvar<N>(a, sa), B = jvar<N>(b, sb);
    auto d = jsub(A, B);
    return jdot(d, d);
}
static inline Jet<12> jee_cross_sqn(V3 a0, V3 a1, V3 b0, V3 b1)
{
    auto A0 = jvar<12>(a0, 0), A1 = jvar<12>(a1, 1), B0 = jvar<12>(b0, 2), B1 = jvar<12>(b1, 3);
    auto c = jcross(jsub(A1, A0), jsub(B1, B0));
    return jdot(c, c);
}

// ---------------------------------------------------------------------------
// Symmetric eigen-decomposition (cyclic Jacobi with threshold, to machine
// precision) and the spectral PSD clamp of Library/Math/UTILS.h:9-27:
// negative eigenvalues are set to 0 and the matrix rebuilt; nothing is done
// when the smallest eigenvalue is >= 0.
// ---------------------------------------------------------------------------
static inline void jacobi_eig(int n, double* A /*n*n, destroyed*/, double* w, double* V)
{
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) V[i * n + j] = (i == j) ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int i = 0; i < n; ++i) {
            diag += A[i * n + i] * A[i * n + i];
            for (int j = i + 1; j < n; ++j) off += A[i * n + j] * A[i * n + j];
        }
        if (off <= 1e-36 * diag || off == 0.0) break;
        for (int p = 0; p < n - 1; ++p)
            for (int q = p + 1; q < n; ++q) {
                double apq = A[p * n + q];
                if (apq == 0.0) continue;
                double app = A[p * n + p], aqq = A[q * n + q];
                double theta = (aqq - app) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (std::fabs(theta) + std::sqrt(theta * theta + 1.0));
                if (std::fabs(theta) > 1e150) t = 0.5 / theta;
                double c = 1.0 / std::sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < n; ++k) {
                    double akp = A[k * n + p], akq = A[k * n + q];
                    A[k * n + p] = c * akp - s * akq;
                    A[k * n + q] = s * akp + c * akq;
                }
                for (int k = 0; k < n; ++k) {
                    double apk = A[p * n + k], aqk = A[q * n + k];
                    A[p * n + k] = c * apk - s * aqk;
                    A[q * n + k] = s * apk + c * aqk;
                }
                A[p * n + q] = A[q * n + p] = 0.0;
                for (int k = 0; k < n; ++k) {
                    double vkp = V[k * n + p], vkq = V[k * n + q];
                    V[k * n + p] = c * vkp - s * vkq;
                    V[k * n + q] = s * vkp + c * vkq;
                }
            }
    }
    for (int i = 0; i < n; ++i) w[i] = A[i * n + i];
}

static inline void make_pd(int n, double* M)
{
    double A[144], w[12], V[144];
    std::memcpy(A, M, sizeof(double) * n * n);
    jacobi_eig(n, A, w, V);
    double wmin = w[0];
    for (int i = 1; i < n; ++i) wmin = std::min(wmin, w[i]);
    if (wmin >= 0) return;
    for (int i = 0; i < n; ++i) w[i] = std::max(w[i], 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0.0;
            for (int k = 0; k < n; ++k) s += V[i * n + k] * w[k] * V[j * n + k];
            M[i * n + j] = s;
        }
}

}  // namespace orc
