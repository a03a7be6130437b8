"""Pins the oracle's isothermal path (eos = 1): HLLD-iso (hydro/rsolvers/mhd/hlld_iso.cpp),
the isothermal branch of HLLE (hlle_mhd.cpp), the isothermal fast speed and C2P
(eos/isothermal_mhd.cpp) and the VL2 update without an energy equation.

Checks (none retypes the oracle's formulas):
  - consistency: F(U,U) equals the physical isothermal MHD flux for both solvers;
  - mirror symmetry x -> -x;
  - HLLD-iso transmits a pure Alfven (rotational) discontinuity with no mass flux;
  - full scheme: second-order convergence of an isothermal sound wave (error ratio per
    doubling), conservation of mass and momentum to round-off, div B at round-off,
    uniform states stay uniform, bit-identical results for two decompositions, and the
    energy slot stays zero.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

CS = 1.3


def phys_flux_iso(w, bx, cs=CS):
    d, vx, vy, vz, _, by, bz = w
    b2 = bx * bx + by * by + bz * bz
    pt = cs * cs * d + 0.5 * b2
    return np.array([d * vx, d * vx * vx + pt - bx * bx, d * vx * vy - bx * by,
                     d * vx * vz - bx * bz, 0.0, by * vx - bx * vy, bz * vx - bx * vz])


STATES = [
    (np.array([1.0, 0.3, -0.2, 0.1, 0.0, 0.5, -0.3]), 0.7),
    (np.array([0.2, -1.5, 0.4, 0.0, 0.0, -1.1, 0.2]), -0.4),
    (np.array([3.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]), 0.0),
    (np.array([1.0, 2.0, 1.0, -1.0, 0.0, 2.0, 1.0]), 1.5),
]


@pytest.mark.parametrize("rs", [0, 1])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_iso_consistency(orc, rs, k):
    w, bx = STATES[k]
    f = orc.riemann_iso(rs, CS, w, w, bx)
    np.testing.assert_allclose(f, phys_flux_iso(w, bx), rtol=1e-13, atol=1e-13)


def mirror(w):
    d, vx, vy, vz, p, by, bz = w
    return np.array([d, -vx, vy, vz, p, by, bz])


@pytest.mark.parametrize("rs", [0, 1])
def test_iso_mirror_symmetry(orc, rs):
    rng = np.random.default_rng(9)
    sign = np.array([-1, 1, -1, -1, 1, -1, -1])
    for _ in range(50):
        wl = np.array([rng.uniform(0.2, 2), *rng.normal(size=3), 0.0, *rng.normal(size=2)])
        wr = np.array([rng.uniform(0.2, 2), *rng.normal(size=3), 0.0, *rng.normal(size=2)])
        bx = rng.normal()
        f = orc.riemann_iso(rs, CS, wl, wr, bx)
        g = orc.riemann_iso(rs, CS, mirror(wr), mirror(wl), -bx)
        np.testing.assert_allclose(g, sign * f, rtol=1e-9, atol=1e-11)


def test_hlld_iso_stationary_rotational_discontinuity(orc):
    """Equal density and |B_t|, transverse B rotated, transverse velocity jump
    dv = +dB/sqrt(rho) (bx > 0): an Alfven discontinuity moving at u - bx/sqrt(rho);
    with u = bx/sqrt(rho) it is stationary (the jump conditions of F_my and F_by give
    u^2 = bx^2/rho and [v_t] = u [B_t] / bx).  HLLD-iso resolves it: zero mass flux and the
    physical flux of either side."""
    rho, bx = 1.0, 0.8
    u = bx / np.sqrt(rho)
    bt_l = np.array([0.6, 0.0])
    bt_r = np.array([0.0, 0.6])
    v_l = np.array([0.1, -0.2])
    v_r = v_l + (bt_r - bt_l) / np.sqrt(rho)
    wl = np.array([rho, u, v_l[0], v_l[1], 0.0, bt_l[0], bt_l[1]])
    wr = np.array([rho, u, v_r[0], v_r[1], 0.0, bt_r[0], bt_r[1]])
    f = orc.riemann_iso(0, CS, wl, wr, bx)
    fl, fr = phys_flux_iso(wl, bx), phys_flux_iso(wr, bx)
    # the discontinuity is a wave of the exact solution: F(wl) and F(wr) coincide
    np.testing.assert_allclose(fl[[0, 1, 2, 3, 5, 6]], fr[[0, 1, 2, 3, 5, 6]], atol=1e-14)
    np.testing.assert_allclose(f, fl, rtol=1e-12, atol=1e-13)


def sound_wave(N, order=2, amp=1e-4):
    n2 = 4
    box = ((0.0, 1.0), (0.0, n2 / N), (0.0, n2 / N))
    d = make_desc(nx=(N, n2, n2), bnx=(N // 2, n2, n2), box=box, cfl=0.4, eos=1, iso_cs=CS,
                  xorder=order)
    o = Oracle(d)
    ng = d.nghost
    u = o.view("u")
    for g in range(u.shape[0]):
        x = o.coords(g, 0)[1]
        s = np.sin(2 * np.pi * x)
        rho = 1.0 + amp * s
        u[g, 0] = rho[None, None, :]
        u[g, 1] = (rho * amp * CS * s)[None, None, :]
        u[g, 2:] = 0.0
    o.view("b1f")[...] = 0.3
    o.view("b2f")[...] = 0.0
    o.view("b3f")[...] = 0.0
    rho0 = o.get("u")[:, 0, ng:-ng, ng:-ng, ng:-ng].copy()
    o.initialize()
    o.run(1.0 / CS)   # one period of the right-moving sound wave
    rho1 = o.get("u")[:, 0, ng:-ng, ng:-ng, ng:-ng]
    return np.mean(np.abs(rho1 - rho0)), o


def test_iso_sound_wave_second_order():
    e32, _ = sound_wave(32)
    e64, o = sound_wave(64)
    print(e32, e64, e64 / e32)
    assert e64 / e32 <= 0.3
    assert np.all(o.get("u")[:, 4] == 0.0)


@pytest.mark.parametrize("rs", [0, 1])
def test_iso_conservation_divb_decomposition(rs):
    res = []
    for bnx in ((16, 16, 16), (8, 8, 8)):
        d = make_desc(nx=(16, 16, 16), bnx=bnx, box=((0, 1),) * 3, cfl=0.4, eos=1, iso_cs=CS,
                      rsolver=rs)
        o = Oracle(d)
        load_state(o, turbulent_state(d, seed=4, mach=1.5, iso=True))
        o.initialize()
        t0 = o.totals()
        o.run(0.1)
        t1 = o.totals()
        # totals: mass, three momenta (energy slot unused), three face-field totals
        np.testing.assert_allclose(t1, t0, rtol=1e-13, atol=1e-12)
        assert o.max_divb() < 1e-13   # absolute, B per unit length
        assert np.all(o.get("u")[:, 4] == 0.0)
        from tests.conftest import load_agx
        res.append((load_agx().gather_cells(d, o.get("u")), o.time()))
    assert res[0][1] == res[1][1]
    assert np.array_equal(res[0][0], res[1][0])


@pytest.mark.parametrize("rs", [0, 1])
def test_iso_uniform_state(rs):
    d = make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), eos=1, iso_cs=CS, rsolver=rs)
    o = Oracle(d)
    u = o.view("u")
    u[:, 0] = 1.3
    u[:, 1] = 0.2 * 1.3
    u[:, 2] = -0.1 * 1.3
    u[:, 3] = 0.05 * 1.3
    u[:, 4] = 0.0
    o.view("b1f")[...] = 0.4
    o.view("b2f")[...] = -0.3
    o.view("b3f")[...] = 0.2
    u0 = o.get("u")
    o.initialize()
    o.run(1.0, nlim=3)
    ng = 2
    sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    np.testing.assert_allclose(o.get("u")[sl], u0[sl], rtol=1e-13, atol=1e-15)
