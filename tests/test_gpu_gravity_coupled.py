"""Coupled MHD + multigrid self-gravity cycles (two MG solves per VL2 cycle,
main.cpp:459-466) on the GPU against the oracle."""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_agx
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("mode,thr,nit", [(0, 0.0, 10), (1, -1.0, 4)])
def test_selfgravity_cycles(mode, thr, nit):
    agx = load_agx()
    d = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, self_gravity=1,
                  four_pi_G=4.0, mg_mode=mode, mg_threshold=thr, mg_niter=nit)
    st = turbulent_state(d, seed=21, mach=1.0, drho=0.4)
    o = Oracle(d)
    m = agx.Mesh(d)
    load_state(o, st)
    load_state(m, st)
    o.initialize()
    m.initialize()
    for _ in range(3):
        o.cycle(1e9)
    m.cycle(1e9, 3)
    for n in ("u", "b1f", "phi"):
        a, b = m.get(n), o.get(n)
        err = np.abs(a - b).max()
        print(n, err, np.abs(b).max())
        assert err <= 1e-11 * np.abs(b).max(), n


def collapse_state(d, seed=4):
    """Isothermal sphere in the manner of pgen/collapse.cpp:157-159 on a uniform mesh:
    rho = (1 + r^2/(26/3 a^2))^-1.5 inside r_c (contrast ~14 at r_c = 6.45 a), uniform
    outside, x (1 + 0.1 U[-1, 1]) noise, at rest, uniform Bz at plasma beta 10 (c_s = 1)."""
    agx = load_agx()
    N = d.nx1
    L = d.x1max - d.x1min
    x = d.x1min + (np.arange(N) + 0.5) * L / N
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    a = 0.06 * L
    r = np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    rc = 6.45 * a
    prof = np.where(r < rc, (1.0 + r ** 2 / (26.0 / 3.0 * a * a)) ** -1.5, (1.0 + rc ** 2 / (26.0 / 3.0 * a * a)) ** -1.5)
    rng = np.random.default_rng(seed)
    rho = prof * (1.0 + 0.1 * rng.uniform(-1.0, 1.0, size=prof.shape))
    U = np.zeros((5, N, N, N))
    U[0] = rho
    bz = np.sqrt(2.0 * 1.0 / 10.0)
    return agx.scatter_cells(d, U), bz


def test_isothermal_multipole_collapse_cycles():
    """The non-SMR part of the self-gravitating isothermal collapse workload: isothermal
    HLLD MHD with multipole (order 4, centre of mass) multigrid gravity and FAS, three
    VL2 cycles on 8 MeshBlocks, GPU vs oracle (the hydro boundaries stay periodic)."""
    agx = load_agx()
    d = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((-0.5, 0.5),) * 3, cfl=0.3, self_gravity=1,
                  four_pi_G=20.0, mg_mode=0, mg_threshold=0.0, eos=1, iso_cs=1.0, mg_bc=1, mporder=4)
    u, bz = collapse_state(d)
    o = Oracle(d)
    m = agx.Mesh(d)
    o.set("u", u)
    m.set("u", u[:, :4])   # NHYDRO = 4 on the host side of the isothermal C ABI
    for h in (o, m):
        h.set("b3f", np.full_like(o.get("b3f"), bz))
    o.initialize()
    m.initialize()
    for _ in range(3):
        o.cycle(1e9)
    assert m.cycle(1e9, 3) == 3
    assert m.time()[0] == pytest.approx(o.time()[0], rel=1e-12)
    for n in ("u", "b3f", "phi"):
        a, b = m.get(n), o.get(n)
        if n == "u":
            b = b[:, :4]
        err = np.abs(a - b).max()
        print(n, err, np.abs(b).max(), flush=True)
        assert err <= 1e-11 * np.abs(b).max(), n
