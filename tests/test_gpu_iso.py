"""GPU vs oracle parity for the isothermal EOS (eos = 1): HLLD-iso
(hydro/rsolvers/mhd/hlld_iso.cpp) or the isothermal HLLE branch, isothermal C2P
(eos/isothermal_mhd.cpp:40-73), the update without an energy equation.  The C ABI
carries NHYDRO = 4 variables for u, w and the fluxes, as the reference's arrays do;
the oracle keeps a fifth, zero slot.  Bar: bitwise for the FP64 form (asserted as
|gpu - oracle| <= 1e-12 max|oracle|), 1e-5 relative for the mixed form."""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import turbulent_state

pytestmark = pytest.mark.gpu
CS = 1.1


def close(a, b, name, rtol=1e-12):
    err = np.abs(a - b).max()
    scale = max(np.abs(b).max(), 1e-300)
    print(f"{name}: max|diff| {err:.3e} (scale {scale:.3e}) bitwise={np.array_equal(a, b)}", flush=True)
    assert err <= rtol * scale, name


def iso_pair(agx, rsolver, shape="cube", precision=0, order=2, nghost=2):
    if shape == "cube":
        kw = dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3)
    else:
        kw = dict(nx=(24, 16, 8), bnx=(8, 4, 8), box=((0, 1.5), (0, 1), (0, 0.5)))
    d = make_desc(cfl=0.4, eos=1, iso_cs=CS, rsolver=rsolver, xorder=order, nghost=nghost, **kw)
    u, f1, f2, f3 = turbulent_state(d, seed=31, mach=1.5, iso=True)
    o = Oracle(d)
    for n, a in (("u", u), ("b1f", f1), ("b2f", f2), ("b3f", f3)):
        o.set(n, a)
    dm = type(d).from_buffer_copy(d)
    dm.precision = precision
    m = agx.Mesh(dm)
    m.set("u", u[:, :4])
    for n, a in (("b1f", f1), ("b2f", f2), ("b3f", f3)):
        m.set(n, a)
    return o, m


@pytest.mark.parametrize("rsolver", [0, 1])
def test_iso_fluxes(agx, rsolver):
    o, m = iso_pair(agx, rsolver)
    o.initialize()
    m.initialize()
    assert m.get("u").shape[1] == 4
    close(m.get("w"), o.get("w")[:, :4], "w", rtol=0.0)
    o.L.orc_calculate_fluxes(o.h, 2)
    m.calculate_fluxes(2)
    for n in ("fx1", "fx2", "fx3"):
        close(m.get(n), o.get(n)[:, :4], n)
    for n in ("e3x1", "e2x1", "e1x2", "e3x2", "e1x3", "e2x3", "wx1", "wx2", "wx3"):
        close(m.get(n), o.get(n), n)


@pytest.mark.parametrize("rsolver", [0, 1])
@pytest.mark.parametrize("shape", ["cube", "ragged"])
def test_iso_cycles(agx, rsolver, shape):
    o, m = iso_pair(agx, rsolver, shape)
    o.initialize()
    m.initialize()
    for _ in range(4):
        o.cycle(1e9)
    assert m.cycle(1e9, 4) == 4
    assert m.time()[1] == o.time()[1]
    close(m.get("u"), o.get("u")[:, :4], "u")
    close(m.get("w"), o.get("w")[:, :4], "w")
    for n in ("b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), n)


def test_iso_ppm_cycles(agx):
    o, m = iso_pair(agx, 0, order=3, nghost=3)
    o.initialize()
    m.initialize()
    for _ in range(3):
        o.cycle(1e9)
    m.cycle(1e9, 3)
    close(m.get("u"), o.get("u")[:, :4], "u")
    close(m.get("b1f"), o.get("b1f"), "b1f")


def test_iso_task_level_stage(agx):
    o, m = iso_pair(agx, 0)
    o.initialize()
    m.initialize()
    o.L.orc_calculate_fluxes(o.h, 1)
    m.calculate_fluxes(1)
    o.L.orc_corner_e(o.h)
    m.corner_e()
    o.L.orc_integrate(o.h, 1)
    m.integrate(1)
    ng = 2
    sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    close(m.get("u")[sl], o.get("u")[:, :4][sl], "u after stage 1")


def test_iso_mixed_run(agx):
    o, m = iso_pair(agx, 0, precision=1)
    tlim = 0.04
    o.initialize()
    m.initialize(tlim)
    o.run(tlim)
    m.run(tlim)
    ng = 2
    g = slice(ng, -ng)
    wm, wo = m.get("w")[:, :, g, g, g], o.get("w")[:, :4, g, g, g]
    err = np.abs(wm[:, 0] - wo[:, 0]).max() / np.abs(wo[:, 0]).max()
    print(f"mixed isothermal: rho rel err {err:.3e}", flush=True)
    assert err <= 1e-5
