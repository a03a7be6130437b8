"""GPU (HIP, gfx950) vs CPU oracle parity for the plain FP64 MHD path.

Every comparison runs the same seeded state through the oracle and through the C ABI.
Bar: the plain FP64 path is built without FP contraction and follows the reference's
operand order, so arrays are expected to agree bit for bit; the asserted tolerance is
the path's contract, |gpu - oracle| <= 1e-12 * max|oracle| per array.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def close(a, b, name, rtol=RTOL):
    scale = max(np.abs(b).max(), 1e-300)
    err = np.abs(a - b).max()
    print(f"{name}: max|diff| = {err:.3e} (scale {scale:.3e}) bitwise={np.array_equal(a, b)}", flush=True)
    assert err <= rtol * scale, name


def pair(agx, desc, state=None, linwave=None):
    o = Oracle(desc)
    m = agx.Mesh(desc)
    if linwave is not None:
        o.pgen_linear_wave(*linwave)
        for n in ("u", "b1f", "b2f", "b3f"):
            m.set(n, o.get(n))
    else:
        load_state(o, state)
        load_state(m, state)
    return o, m


CASES = {
    "linwave_1blk": lambda: (make_desc(nx=(32, 16, 16), bnx=(32, 16, 16)), None, (0, 1e-6, 0.0)),
    "turb_8blk": lambda: (make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4),
                          "turb", None),
    "turb_ragged": lambda: (make_desc(nx=(24, 16, 8), bnx=(8, 4, 8), box=((0, 1.5), (0, 1), (0, 0.5))),
                            "turb", None),
}


def build_case(agx, name, rsolver=0):
    desc, st, lw = CASES[name]()
    desc.rsolver = rsolver
    state = turbulent_state(desc, seed=11, mach=1.5) if st else None
    return pair(agx, desc, state, lw)


@pytest.mark.parametrize("case", list(CASES))
def test_initialize(agx, case):
    o, m = build_case(agx, case)
    o.initialize()
    m.initialize()
    for n in ("u", "w", "bcc", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), n, rtol=0.0 if n != "bcc" else RTOL)
    assert m.time()[1] == o.time()[1]


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("order", [1, 2])
@pytest.mark.parametrize("rsolver", [0, 1, 2, 3])
def test_fluxes_and_emf(agx, case, order, rsolver):
    o, m = build_case(agx, case, rsolver)
    o.initialize()
    m.initialize()
    o.L.orc_calculate_fluxes(o.h, order)
    m.calculate_fluxes(order)
    for n in ("fx1", "fx2", "fx3", "e3x1", "e2x1", "e1x2", "e3x2", "e1x3", "e2x3", "wx1", "wx2", "wx3"):
        close(m.get(n), o.get(n), n)
    o.L.orc_corner_e(o.h)
    m.corner_e()
    for n in ("e1", "e2", "e3"):
        close(m.get(n), o.get(n), n)


@pytest.mark.parametrize("case", list(CASES))
def test_stage_updates(agx, case):
    o, m = build_case(agx, case)
    o.initialize()
    m.initialize()
    for stage in (1, 2):
        order = 1 if stage == 1 else 2
        o.L.orc_calculate_fluxes(o.h, order)
        m.calculate_fluxes(order)
        o.L.orc_corner_e(o.h)
        m.corner_e()
        o.L.orc_integrate(o.h, stage)
        m.integrate(stage)
        ng = o.desc.nghost
        # compare interiors (ghost zones of the new register are refreshed next)
        sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
        close(m.get("u")[sl], o.get("u")[sl], f"u stage{stage}")
        close(m.get("u1")[sl], o.get("u1")[sl], f"u1 stage{stage}")
        for n in ("b1f", "b2f", "b3f"):
            a, b = m.get(n), o.get(n)
            close(a[:, ng:-ng, ng:-ng, ng:-ng], b[:, ng:-ng, ng:-ng, ng:-ng], f"{n} stage{stage}")
        o.L.orc_bvals_cc(o.h, 0)
        o.L.orc_bvals_fc(o.h, 0)
        m.bvals(3)
        o.L.orc_cons2prim(o.h)
        m.cons2prim()
        for n in ("u", "w", "bcc", "b1f", "b2f", "b3f"):
            close(m.get(n), o.get(n), f"{n} after stage{stage}")


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("rsolver", [0, 1, 2, 3])
def test_cycles(agx, case, rsolver):
    o, m = build_case(agx, case, rsolver)
    o.initialize()
    m.initialize()
    for _ in range(4):
        o.cycle(1e9)
    assert m.cycle(1e9, 4) == 4
    to, tm = o.time(), m.time()
    assert tm[2] == to[2] == 4
    assert abs(tm[0] - to[0]) <= 1e-14 * to[0] and abs(tm[1] - to[1]) <= 1e-13 * to[1]
    for n in ("u", "w", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), n)


def test_linear_wave_errors_match(agx):
    """Full linear-wave run (64x32x32, t=0.5, the L/R test configuration of
    mhd_linwave.py:88-95): GPU error row equals the oracle's."""
    desc = make_desc(nx=(64, 32, 32), bnx=(32, 16, 16), cfl=0.3)
    o, m = pair(agx, desc, linwave=(0, 1e-6, 0.0))
    o.initialize()
    m.initialize(0.5)
    o.run(0.5)
    m.run(0.5)
    assert m.time()[2] == o.time()[2]
    for n in ("u", "bcc"):
        close(m.get(n), o.get(n), n)
    # error norms from the GPU state, evaluated by the oracle's error routine
    o2 = Oracle(desc)
    o2.set("u", m.get("u"))
    o2.set("bcc", m.get("bcc"))
    eg = o2.linear_wave_errors(0, 1e-6, 0.0)
    eo = o.linear_wave_errors(0, 1e-6, 0.0)
    np.testing.assert_allclose(eg, eo, rtol=1e-9)
    assert eg[17] <= 8.0


def test_divb_preserved(agx):
    desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4)
    st = turbulent_state(desc, seed=5, mach=2.0)
    m = agx.Mesh(desc)
    load_state(m, st)
    m.initialize()
    m.cycle(1e9, 10)
    o = Oracle(desc)
    for n in ("b1f", "b2f", "b3f"):
        o.set(n, m.get(n))
    assert o.max_divb() < 1e-13   # absolute, B per unit length


def test_floors_path(agx):
    """Near-zero pressure: density/pressure floors fire (eos/adiabatic_mhd.cpp:66-84);
    the fused stage must reproduce the reference's floor write-back and ghost C2P."""
    desc = make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=((0, 1),) * 3, cfl=0.3)
    st = turbulent_state(desc, seed=3, p0=1e-19, vamp=0.3, drho=0.5)
    o, m = pair(agx, desc, st)
    o.initialize()
    m.initialize()
    w0 = o.get("w")
    assert (w0[:, 4] <= desc.pfloor * 1.0000001).any()   # floors active
    for _ in range(2):
        o.cycle(1e9)
    m.cycle(1e9, 2)
    for n in ("u", "w", "bcc", "b1f"):
        close(m.get(n), o.get(n), n)


@pytest.mark.parametrize("rsolver", [0, 1, 2])
def test_ppm_fluxes_and_cycles(agx, rsolver):
    """PPM reconstruction (xorder=3, nghost=3): face fluxes, then full VL2 cycles
    (stage 1 donor cell, stage 2 PPM)."""
    desc = make_desc(nx=(24, 24, 24), bnx=(12, 12, 12), box=((0, 1),) * 3, cfl=0.4, xorder=3,
                     nghost=3, rsolver=rsolver)
    st = turbulent_state(desc, seed=17, mach=2.0)
    o, m = pair(agx, desc, st)
    o.initialize()
    m.initialize()
    o.L.orc_calculate_fluxes(o.h, 3)
    m.calculate_fluxes(3)
    for n in ("fx1", "fx2", "fx3", "e3x1", "e2x1", "e1x2", "e3x2", "e1x3", "e2x3", "wx1", "wx2", "wx3"):
        close(m.get(n), o.get(n), n)
    for _ in range(3):
        o.cycle(1e9)
    m.cycle(1e9, 3)
    for n in ("u", "w", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), n)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5])
@pytest.mark.parametrize("case", ["turb_8blk", "turb_ragged"])
def test_cycles_every_tile_shape(agx, case, tile):
    """Every stage form (flux launches + stage kernel in four tile shapes, the one-launch
    stage in two) gives the oracle's cycles bit for bit, and the autotuner leaves the
    state untouched."""
    o, m = build_case(agx, case)
    o.initialize()
    m.initialize()
    u0 = m.get("u")
    choice = m.autotune()
    assert 0 <= choice <= 5
    assert np.array_equal(m.get("u"), u0)
    m.set_tile(tile)
    for _ in range(3):
        o.cycle(1e9)
    assert m.cycle(1e9, 3) == 3
    for n in ("u", "w", "bcc", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), f"{n} tile {tile}", rtol=0.0 if n != "bcc" else RTOL)
