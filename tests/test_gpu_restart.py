"""Restart through the device (agx_restart_write / agx_restart_read): a run stopped
after some cycles, written to a restart file in the reference's layout, read into a new
mesh and initialised (Mesh::Initialize after the restart constructor: gravity solve,
ghost exchange, C2P, NewTimeStep capped by the file's dt) continues exactly as the
uninterrupted run, bit for bit, FP64, with and without multigrid self-gravity.  The file
also round-trips through the oracle: its arrays equal the oracle's state after the same
cycles.  agx_vtk_write writes the reference's VTK files from the device state."""
import os

import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu
FIELDS = ("u", "b1f", "b2f", "b3f")


@pytest.mark.parametrize("sg", [0, 1])
def test_restart_continues_bitwise(agx, sg, tmp_path):
    d = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, self_gravity=sg, four_pi_G=1.0,
                  mg_mode=0, mg_threshold=1e-9 if sg else -1.0)
    st = turbulent_state(d, seed=23, mach=1.5, drho=0.4)
    a = agx.Mesh(d)
    load_state(a, st)
    a.initialize()
    assert a.cycle(1e30, 3) == 3
    path = str(tmp_path / "turb.00001.rst")
    a.restart_write(path)
    o = Oracle(d)
    load_state(o, st)
    o.initialize()
    for _ in range(3):
        o.cycle(1e30)
    # the file holds the state after 3 cycles
    t, dt, n, *arrs = agx.restart_read_host(path, d, [o.get(f).shape for f in FIELDS])
    assert n == 3 and t == a.time()[0] and dt == a.time()[1]
    for f, x in zip(FIELDS, arrs):
        assert np.array_equal(x, a.get(f)), f
        assert np.abs(x - o.get(f)).max() <= 1e-12 * np.abs(o.get(f)).max(), f
    b = agx.Mesh(d)
    b.restart_read(path)
    b.initialize()
    assert b.time() == a.time()
    assert a.cycle(1e30, 3) == 3 and b.cycle(1e30, 3) == 3
    assert b.time() == a.time()
    for f in FIELDS + ("w",) + (("phi",) if sg else ()):
        assert np.array_equal(b.get(f), a.get(f)), f
    a.close()
    b.close()


@pytest.mark.parametrize("sg", [0, 1])
def test_vtk_from_the_device_equals_the_host_writer(agx, sg, tmp_path):
    """agx_vtk_write (the device state after 2 cycles) writes the same bytes as
    agx_vtk_write_host on the downloaded arrays, whose format tests/test_vtk_format.py
    checks against the oracle; and the arrays match the oracle's after the same cycles."""
    d = make_desc(nx=(32, 16, 16), bnx=(16, 16, 16), box=((0, 2), (0, 1), (0, 1)), cfl=0.4, self_gravity=sg,
                  four_pi_G=1.0, mg_mode=0, mg_threshold=1e-9 if sg else -1.0)
    st = turbulent_state(d, seed=5, mach=1.2)
    m = agx.Mesh(d)
    load_state(m, st)
    m.initialize()
    assert m.cycle(1e30, 2) == 2
    (tmp_path / "dev").mkdir()
    (tmp_path / "host").mkdir()
    m.vtk_write(str(tmp_path / "dev" / "run"), "out3", 4, "prim,cons")
    t, _, n = m.time()
    agx.vtk_write_host(str(tmp_path / "host" / "run"), d, t, n, m.get("u"), m.get("w"), m.get("bcc"),
                       m.get("phi") if sg else None, file_id="out3", file_number=4, variable="prim,cons")
    names = sorted(os.listdir(tmp_path / "dev"))
    assert names == sorted(os.listdir(tmp_path / "host")) and len(names) == 2
    for nm in names:
        assert (tmp_path / "dev" / nm).read_bytes() == (tmp_path / "host" / nm).read_bytes(), nm
    o = Oracle(d)
    load_state(o, st)
    o.initialize()
    for _ in range(2):
        o.cycle(1e30)
    for f in ("u", "w"):
        assert np.abs(m.get(f) - o.get(f)).max() <= 1e-12 * np.abs(o.get(f)).max(), f
    m.close()
