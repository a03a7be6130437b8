"""Mixed-precision form (AGX_PREC_MIXED) against the FP64 oracle.

The mixed form reconstructs and solves the Riemann problem in FP32 from FP32 copies of
w and bcc, and keeps the conserved-variable update, C2P with floors, the corner EMFs and
CT in FP64.  Bars (stated per test): density and pressure within 1e-5 of the FP64 oracle
relative to their maximum over the mesh; the face field's div B at FP64 round-off,
because CT preserves it for any EMF values; the time step within 1e-5.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu
RHO_P_TOL = 1e-5


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def mixed_pair(agx, desc, state):
    o = Oracle(desc)
    load_state(o, state)
    dm = type(desc).from_buffer_copy(desc)
    dm.precision = 1
    m = agx.Mesh(dm)
    load_state(m, state)
    return o, m


def interior(a, ng):
    return a[:, :, ng:-ng, ng:-ng, ng:-ng]


@pytest.mark.parametrize("rsolver", [0, 1])
def test_mixed_fluxes(agx, rsolver):
    desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=rsolver)
    o, m = mixed_pair(agx, desc, turbulent_state(desc, seed=11, mach=1.5))
    o.initialize()
    m.initialize()
    assert abs(m.time()[1] - o.time()[1]) <= 1e-5 * o.time()[1]
    o.L.orc_calculate_fluxes(o.h, 2)
    m.calculate_fluxes(2)
    for n in ("fx1", "fx2", "fx3", "e3x1", "e2x1", "e1x2", "e3x2", "e1x3", "e2x3", "wx1", "wx2", "wx3"):
        e = rel_err(m.get(n), o.get(n))
        print(f"{n}: max rel err {e:.3e}", flush=True)
        assert e <= 1e-5, n


@pytest.mark.parametrize("rsolver", [0, 1])
@pytest.mark.parametrize("shape", ["cube", "ragged"])
def test_mixed_turbulence_run(agx, rsolver, shape):
    if shape == "cube":
        desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=rsolver)
    else:
        desc = make_desc(nx=(24, 16, 8), bnx=(8, 4, 8), box=((0, 1.5), (0, 1), (0, 0.5)), rsolver=rsolver)
    o, m = mixed_pair(agx, desc, turbulent_state(desc, seed=11, mach=1.5))
    tlim = 0.04
    o.initialize()
    m.initialize(tlim)
    o.run(tlim)
    m.run(tlim)
    assert m.time()[0] == pytest.approx(tlim, rel=1e-14) and o.time()[0] == pytest.approx(tlim, rel=1e-14)
    print(f"cycles: mixed {m.time()[2]}, fp64 {o.time()[2]}", flush=True)
    ng = desc.nghost
    wm, wo = interior(m.get("w"), ng), interior(o.get("w"), ng)
    erho = rel_err(wm[:, 0], wo[:, 0])
    ep = rel_err(wm[:, 4], wo[:, 4])
    um, uo = interior(m.get("u"), ng), interior(o.get("u"), ng)
    print(f"rho {erho:.3e}  P {ep:.3e}  E {rel_err(um[:, 4], uo[:, 4]):.3e}", flush=True)
    assert erho <= RHO_P_TOL and ep <= RHO_P_TOL
    for n in ("b1f", "b2f", "b3f"):
        sl = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
        e = rel_err(m.get(n)[sl], o.get(n)[sl])
        print(f"{n}: {e:.3e}", flush=True)
        assert e <= RHO_P_TOL
    o2 = Oracle(desc)
    for n in ("b1f", "b2f", "b3f"):
        o2.set(n, m.get(n))
    assert o2.max_divb() < 1e-12
