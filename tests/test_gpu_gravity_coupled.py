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
