// A compiled caller of the C ABI (no Python in the loop): reads one
// animation-fix scene (tets and/or shell triangles) from a binary file, runs `steps` implicit-Euler steps
// through idp_amd::StepSession (hinge or flow variant, optional body force,
// optional moving DBC), calling set_mesh before every step as a driver that
// keeps its topology would, and writes the Newton counts, the mesh upload
// count, the final positions and velocities and kappa.  zero_v: velocities
// are zeroed between steps (MeshIO.Zero_Velocity in an animation-fix driver).
//
// input  (little endian): int32 nV nT nDBC flow has_b steps zero_v nE nH; float64 h
//        tol dHat2 kappa[3] thickness bend_k; float64 X[3nV] v[3nV] Xrest[3nV];
//        int32 tets[4nT]; float64 IB[9nT] vol[nT] lam[nT] mu[nT] mass[nV];
//        float64 b[3nV] (has_b); int32 dbc[nDBC]; float64 dbc_target[3nDBC];
//        int32 tris[3nE]; float64 tri_vol[nE] tri_lam[nE] tri_mu[nE] tri_IB[3nE];
//        int32 hinges[4nH]; float64 hinge_rest[3nH]
// output: int32 iters[steps] uploads; float64 X[3nV] v[3nV] kappa[3]
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "idp_amd_step.hpp"

template <class T>
static bool rd(FILE* f, T* p, size_t n)
{
    return n == 0 || std::fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char** argv)
{
    if (argc != 3) {
        std::fprintf(stderr, "usage: %s <input> <output>\n", argv[0]);
        return 2;
    }
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 2;
    int hdr[9];
    double par[8];
    if (!rd(f, hdr, 9) || !rd(f, par, 8)) return 2;
    const int nV = hdr[0], nT = hdr[1], nD = hdr[2], flow = hdr[3], has_b = hdr[4], steps = hdr[5], zero_v = hdr[6],
              nE = hdr[7], nH = hdr[8];
    std::vector<double> X(3 * nV), v(3 * nV), Xr(3 * nV), IB(9 * nT), vol(nT), lam(nT), mu(nT), mass(nV),
        b(has_b ? 3 * nV : 0), tgt(3 * nD);
    std::vector<int> tets(4 * nT), dbc(nD), tris(3 * nE), hinges(4 * nH);
    std::vector<double> tvol(nE), tlam(nE), tmu(nE), tIB(3 * nE), hrest(3 * nH);
    bool ok = rd(f, X.data(), X.size()) && rd(f, v.data(), v.size()) && rd(f, Xr.data(), Xr.size()) &&
              rd(f, tets.data(), tets.size()) && rd(f, IB.data(), IB.size()) && rd(f, vol.data(), vol.size()) &&
              rd(f, lam.data(), lam.size()) && rd(f, mu.data(), mu.size()) && rd(f, mass.data(), mass.size()) &&
              rd(f, b.data(), b.size()) && rd(f, dbc.data(), dbc.size()) && rd(f, tgt.data(), tgt.size()) &&
              rd(f, tris.data(), tris.size()) && rd(f, tvol.data(), tvol.size()) && rd(f, tlam.data(), tlam.size()) &&
              rd(f, tmu.data(), tmu.size()) && rd(f, tIB.data(), tIB.size()) && rd(f, hinges.data(), hinges.size()) &&
              rd(f, hrest.data(), hrest.size());
    std::fclose(f);
    if (!ok) return 2;

    idp_amd::StepSession sess(0);
    if (sess.status() != IDP_OK) {
        std::fprintf(stderr, "no device context: %d\n", sess.status());
        return 3;
    }
    idp_amd::TetMesh m;
    m.nV = nV;
    m.nT = nT;
    m.X_rest = Xr.data();
    m.tets = tets.data();
    m.IB = IB.data();
    m.vol = vol.data();
    m.lambda = lam.data();
    m.mu = mu.data();
    m.mass = mass.data();
    idp_amd::ShellPart sh;
    sh.nE = nE;
    sh.tris = tris.data();
    sh.vol = tvol.data();
    sh.lambda = tlam.data();
    sh.mu = tmu.data();
    sh.IB = tIB.data();
    sh.nH = nH;
    sh.hinges = hinges.data();
    sh.hinge_rest = hrest.data();
    sh.bend_k = par[7];
    double kappa[3] = {par[3], par[4], par[5]};
    idp_solver_opts o{};
    o.pcg_rtol = 1e-12;
    std::vector<int> iters;
    for (int k = 0; k < steps; ++k) {
        if (k > 0 && zero_v)
            for (double& q : v) q = 0.0;
        int rc = sess.set_mesh(m, nE > 0 ? &sh : nullptr);
        if (rc == IDP_OK && nD) rc = sess.set_dbc(nD, dbc.data(), tgt.data());
        if (rc != IDP_OK) {
            std::fprintf(stderr, "set-up failed: %d %s\n", rc, sess.error());
            return 4;
        }
        const int it = sess.step(X.data(), v.data(), has_b ? b.data() : nullptr, flow != 0, par[0], par[1], true,
                                 par[2], kappa, par[6], &o);
        std::printf("step %d: %d Newton iterations\n", k, it);
        if (it < 0) {
            std::fprintf(stderr, "step failed: %d %s\n", it, sess.error());
            return 5;
        }
        iters.push_back(it);
    }
    FILE* g = std::fopen(argv[2], "wb");
    if (!g) return 2;
    const int up = sess.mesh_uploads();
    std::fwrite(iters.data(), sizeof(int), iters.size(), g);
    std::fwrite(&up, sizeof(int), 1, g);
    std::fwrite(X.data(), sizeof(double), X.size(), g);
    std::fwrite(v.data(), sizeof(double), v.size(), g);
    std::fwrite(kappa, sizeof(double), 3, g);
    std::fclose(g);
    return 0;
}
