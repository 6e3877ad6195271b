// Decoding of the barrier constraint set's int4 entries (the encoding of
// Library/FEM/IPC.h:589-657): the kind, the vertices in emission order and
// the PE / PP multiplicity, and the squared distance each kind is evaluated
// with.  Shared by the barrier and friction kernels.
#pragma once
#include "geom_dev.hpp"

namespace idp {

enum Kind { K_PT = 0, K_EE = 1, K_PE = 2, K_PP = 3, K_MEE = 4, K_MPE = 5, K_MPP = 6 };

struct Decoded {
    int kind;
    int nv;      // vertices carrying the local matrix
    int v[4];    // vertex ids in the reference's emission order
    double mult; // multiplicity (PE/PP merged duplicates)
};

__device__ inline Decoded decode(int4 c)
{
    Decoded d;
    d.mult = 1.0;
    if (c.x >= 0) {
        if (c.z >= 0 && c.w >= 0) {
            d.kind = K_EE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = c.w;
        } else if (c.w >= 0) {
            d.kind = K_MEE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = c.y; d.v[2] = -c.z - 1; d.v[3] = c.w;
        } else if (c.z >= 0) {
            d.kind = K_MPE;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = -c.w - 1; d.v[2] = c.y; d.v[3] = c.z;
        } else {
            d.kind = K_MPP;
            d.nv = 4;
            d.v[0] = c.x; d.v[1] = -c.z - 1; d.v[2] = c.y; d.v[3] = -c.w - 1;
        }
    } else {
        int p = -c.x - 1;
        if (c.w >= 0) {
            d.kind = K_PT;
            d.nv = 4;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = c.w;
        } else if (c.z >= 0) {
            d.kind = K_PE;
            d.nv = 3;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = c.z; d.v[3] = -1;
            d.mult = (double)(-c.w);
        } else {
            d.kind = K_PP;
            d.nv = 2;
            d.v[0] = p; d.v[1] = c.y; d.v[2] = -1; d.v[3] = -1;
            d.mult = (double)(-c.w);
        }
    }
    return d;
}

// squared distance exactly as Compute_Barrier / Compute_Min_Dist2 evaluate it
__device__ inline double constraint_d2(const Decoded& k, const double* X)
{
    switch (k.kind) {
    case K_PT: return d2_pt(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_EE:
    case K_MEE: return d2_ee(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_PE: return d2_pe(ld3(X, k.v[0]), ld3(X, k.v[1]), ld3(X, k.v[2]));
    case K_MPE: return d2_pe(ld3(X, k.v[0]), ld3(X, k.v[2]), ld3(X, k.v[3]));
    case K_PP: return d2_pp(ld3(X, k.v[0]), ld3(X, k.v[1]));
    default: return d2_pp(ld3(X, k.v[0]), ld3(X, k.v[2]));
    }
}

}  // namespace idp
