"""The allocation-free barrier energy (idp_barrier_energy_noalloc) through
the C ABI on MI355X: the same sum, bit for bit, as idp_barrier_energy on the
same constraints, over two calls in a row with the second one smaller (the
fixed scratch array keeps the first call's contents), plus its size limit."""
import numpy as np
import pytest
import torch

from idp_amd import capi, scenes

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _bend_state():
    sc = scenes.config1()
    # halfway to the self-intersecting target: both ends near contact
    X = sc.X + 0.45 * (sc.target - sc.X)
    return sc, X


def test_noalloc_energy_equals_barrier_energy_twice():
    sc, X = _bend_state()
    vol, IB, lam, mu, mass = capi.fem_init(sc.X, sc.tets, sc.density, sc.E, sc.nu)
    dHat2 = sc.dhat ** 2
    kappa = 1e5
    ctx = capi.Context(0)
    try:
        ctx.set_mesh(sc.X, sc.tets, IB, vol, lam, mu, mass)
        cs = ctx.constraint_set(X, dHat2 * 2500)  # a wide radius: many constraints of every kind
        say(f"[noalloc] {cs.shape[0]} constraints on the half-bent bar")
        assert cs.shape[0] > 2000
        Xd = torch.from_numpy(np.ascontiguousarray(X, dtype=np.float64)).to("cuda:0")
        for k, n in enumerate((cs.shape[0], cs.shape[0] // 3)):
            sub = np.ascontiguousarray(cs[:n], dtype=np.int32)
            csd = torch.from_numpy(sub).to("cuda:0")
            torch.cuda.synchronize()
            want = ctx.barrier_energy(X, sub, dHat2 * 2500, kappa)
            got = ctx.barrier_energy_noalloc(Xd.data_ptr(), csd.data_ptr(), n, dHat2 * 2500, kappa)
            say(f"[noalloc] call {k}: n {n}, E {got!r} (idp_barrier_energy {want!r})")
            assert got > 0
            assert got == want  # same blocks, same sums: bit for bit
        # no constraints: zero, like the existing entry point
        assert ctx.barrier_energy_noalloc(Xd.data_ptr(), 0, 0, dHat2, kappa) == 0.0
        # beyond the fixed scratch: a clear error, nothing launched
        with pytest.raises(capi.IdpError) as ei:
            ctx.barrier_energy_noalloc(Xd.data_ptr(), csd.data_ptr(), (1 << 24) + 1, dHat2, kappa)
        assert ei.value.code == capi.IDP_E_CAPACITY
    finally:
        ctx.close()
