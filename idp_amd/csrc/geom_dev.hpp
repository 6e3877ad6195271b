// Device geometry used in every *decision* of the contact path: distance-type
// classification, squared distances, AABB tests, the mollifier switch and the
// additive CCD loop.  The library is compiled with -ffp-contract=off so each
// expression rounds exactly as written (left-to-right sums, no fused
// multiply-add); pair sets and CCD step sizes are therefore bit-identical to
// the CPU restatement, which evaluates the same expressions in the same order.
#pragma once
#include <hip/hip_runtime.h>

namespace idp {

struct d3 {
    double x, y, z;
};
__host__ __device__ inline d3 operator-(d3 a, d3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
__host__ __device__ inline d3 operator+(d3 a, d3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
__host__ __device__ inline d3 operator*(double s, d3 a) { return {s * a.x, s * a.y, s * a.z}; }
__host__ __device__ inline double dot(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__host__ __device__ inline double sqn(d3 a) { return a.x * a.x + a.y * a.y + a.z * a.z; }
__host__ __device__ inline d3 cross(d3 a, d3 b)
{
    return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ inline d3 ld3(const double* X, int i) { return {X[3 * i], X[3 * i + 1], X[3 * i + 2]}; }

// Library/Math/Distance/DISTANCE_TYPE.h:11-27
__device__ inline int pe_type(d3 p, d3 e0, d3 e1)
{
    d3 e = e1 - e0;
    double ratio = dot(e, p - e0) / sqn(e);
    if (ratio < 0) return 0;
    if (ratio > 1) return 1;
    return 2;
}

// 2x2 symmetric solve with the arithmetic of a diagonally pivoted LDL^T
// (largest |diagonal| first, first index on ties), as DISTANCE_TYPE.h:46,54,62
// obtain the edge parameters.
__device__ inline void ldlt2(double m00, double m01, double m11, double r0, double r1, double& x0, double& x1)
{
    const bool sw = fabs(m11) > fabs(m00);
    const double a = sw ? m11 : m00;
    const double c = sw ? m00 : m11;
    const double b0 = sw ? r1 : r0;
    const double b1 = sw ? r0 : r1;
    if (!(fabs(a) > 0.0)) {
        x0 = 0.0;
        x1 = 0.0;
        return;
    }
    const double l = m01 / a;
    const double d1 = c - l * (a * l);
    double y0 = b0;
    double y1 = b1 - l * y0;
    y0 = (fabs(a) > 2.2250738585072014e-308) ? y0 / a : 0.0;
    y1 = (fabs(d1) > 2.2250738585072014e-308) ? y1 / d1 : 0.0;
    const double z1 = y1;
    const double z0 = y0 - l * z1;
    x0 = sw ? z1 : z0;
    x1 = sw ? z0 : z1;
}

// DISTANCE_TYPE.h:30-82: 0..2 PP(t0..t2), 3..5 PE(t0t1,t1t2,t2t0), 6 PT
__device__ inline int pt_type(d3 p, d3 t0, d3 t1, d3 t2)
{
    const d3 b0 = t1 - t0;
    const d3 n = cross(b0, t2 - t0);
    double s[3], u;
    {
        d3 r = cross(b0, n), q = p - t0;
        ldlt2(dot(b0, b0), dot(b0, r), dot(r, r), dot(b0, q), dot(r, q), s[0], u);
        if (s[0] > 0.0 && s[0] < 1.0 && u >= 0.0) return 3;
    }
    {
        d3 e = t2 - t1, r = cross(e, n), q = p - t1;
        ldlt2(dot(e, e), dot(e, r), dot(r, r), dot(e, q), dot(r, q), s[1], u);
        if (s[1] > 0.0 && s[1] < 1.0 && u >= 0.0) return 4;
    }
    {
        d3 e = t0 - t2, r = cross(e, n), q = p - t2;
        ldlt2(dot(e, e), dot(e, r), dot(r, r), dot(e, q), dot(r, q), s[2], u);
        if (s[2] > 0.0 && s[2] < 1.0 && u >= 0.0) return 5;
    }
    if (s[0] <= 0.0 && s[2] >= 1.0) return 0;
    if (s[1] <= 0.0 && s[0] >= 1.0) return 1;
    if (s[2] <= 0.0 && s[1] >= 1.0) return 2;
    return 6;
}

// DISTANCE_TYPE.h:84-165: 0 a0b0, 1 a0b1, 2 a0-b, 3 a1b0, 4 a1b1, 5 a1-b, 6 b0-a, 7 b1-a, 8 EE
__device__ inline int ee_type(d3 ea0, d3 ea1, d3 eb0, d3 eb1)
{
    d3 u = ea1 - ea0, v = eb1 - eb0, w = ea0 - eb0;
    double a = sqn(u), b = dot(u, v), c = sqn(v), d = dot(u, w), e = dot(v, w);
    double D = a * c - b * b;
    double tD = D, tN;
    int def = 8;
    double sN = b * e - c * d;
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
            d3 uv = cross(u, v);
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

__device__ inline double d2_pp(d3 a, d3 b) { return sqn(a - b); }
__device__ inline double d2_pe(d3 p, d3 e0, d3 e1) { return sqn(cross(e0 - p, e1 - p)) / sqn(e1 - e0); }
__device__ inline double d2_pt(d3 p, d3 t0, d3 t1, d3 t2)
{
    d3 b = cross(t1 - t0, t2 - t0);
    double s = dot(p - t0, b);
    return s * s / sqn(b);
}
__device__ inline double d2_ee(d3 a0, d3 a1, d3 b0, d3 b1)
{
    d3 b = cross(a1 - a0, b1 - b0);
    double s = dot(b0 - a0, b);
    return s * s / sqn(b);
}
__device__ inline double ee_cross_sqn(d3 a0, d3 a1, d3 b0, d3 b1) { return sqn(cross(a1 - a0, b1 - b0)); }
// EDGE_EDGE_MOLLIFIER.h:583-592
__device__ inline double mollifier_eps(d3 a0r, d3 a1r, d3 b0r, d3 b1r)
{
    return 1.0e-3 * sqn(a0r - a1r) * sqn(b0r - b1r);
}

__device__ inline double d2_pt_unclassified(d3 p, d3 t0, d3 t1, d3 t2)
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
__device__ inline double d2_ee_unclassified(d3 a0, d3 a1, d3 b0, d3 b1)
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

// CCD.h:150-185 (static AABB tests with radius)
__device__ inline bool pt_cd_broadphase(d3 p, d3 t0, d3 t1, d3 t2, double dist)
{
    if (p.x - fmax(t0.x, fmax(t1.x, t2.x)) > dist || fmin(t0.x, fmin(t1.x, t2.x)) - p.x > dist) return false;
    if (p.y - fmax(t0.y, fmax(t1.y, t2.y)) > dist || fmin(t0.y, fmin(t1.y, t2.y)) - p.y > dist) return false;
    if (p.z - fmax(t0.z, fmax(t1.z, t2.z)) > dist || fmin(t0.z, fmin(t1.z, t2.z)) - p.z > dist) return false;
    return true;
}
__device__ inline bool ee_cd_broadphase(d3 a0, d3 a1, d3 b0, d3 b1, double dist)
{
    if (fmin(a0.x, a1.x) - fmax(b0.x, b1.x) > dist || fmin(b0.x, b1.x) - fmax(a0.x, a1.x) > dist) return false;
    if (fmin(a0.y, a1.y) - fmax(b0.y, b1.y) > dist || fmin(b0.y, b1.y) - fmax(a0.y, a1.y) > dist) return false;
    if (fmin(a0.z, a1.z) - fmax(b0.z, b1.z) > dist || fmin(b0.z, b1.z) - fmax(a0.z, a1.z) > dist) return false;
    return true;
}

// CCD.h:188-235 (swept AABB tests over the full search direction)
__device__ inline bool pt_ccd_broadphase(d3 p, d3 t0, d3 t1, d3 t2, d3 dp, d3 dt0, d3 dt1, d3 dt2, double dist)
{
    d3 pe = p + dp, a = t0 + dt0, b = t1 + dt1, c = t2 + dt2;
#define IDP_AX(f)                                                                                        \
    {                                                                                                    \
        double mxp = fmax(p.f, pe.f), mnp = fmin(p.f, pe.f);                                             \
        double mxt = fmax(fmax(fmax(t0.f, t1.f), fmax(t2.f, a.f)), fmax(b.f, c.f));                     \
        double mnt = fmin(fmin(fmin(t0.f, t1.f), fmin(t2.f, a.f)), fmin(b.f, c.f));                     \
        if (mnp - mxt > dist || mnt - mxp > dist) return false;                                          \
    }
    IDP_AX(x) IDP_AX(y) IDP_AX(z)
#undef IDP_AX
    return true;
}
__device__ inline bool ee_ccd_broadphase(d3 a0, d3 a1, d3 b0, d3 b1, d3 da0, d3 da1, d3 db0, d3 db1, double dist)
{
    d3 a0e = a0 + da0, a1e = a1 + da1, b0e = b0 + db0, b1e = b1 + db1;
#define IDP_AX(f)                                                                                        \
    {                                                                                                    \
        double mxa = fmax(fmax(a0.f, a1.f), fmax(a0e.f, a1e.f)), mna = fmin(fmin(a0.f, a1.f), fmin(a0e.f, a1e.f)); \
        double mxb = fmax(fmax(b0.f, b1.f), fmax(b0e.f, b1e.f)), mnb = fmin(fmin(b0.f, b1.f), fmin(b0e.f, b1e.f)); \
        if (mna - mxb > dist || mnb - mxa > dist) return false;                                          \
    }
    IDP_AX(x) IDP_AX(y) IDP_AX(z)
#undef IDP_AX
    return true;
}

// IPC log barrier on squared distance (Library/Math/BARRIER.h:10-62)
__device__ inline double barrier_b(double d, double dh, double k)
{
    double t = d - dh;
    return -k * t * t * log(d / dh);
}
__device__ inline double barrier_db(double d, double dh, double k)
{
    double t = d - dh;
    return k * (-2.0 * t * log(d / dh) - t * t / d);
}
__device__ inline double barrier_d2b(double d, double dh, double k)
{
    double t = d - dh;
    return k * (-2.0 * log(d / dh) - 4.0 * t / d + t * t / (d * d));
}

}  // namespace idp
