"""Pins of the oracle's history sums (orc_history, outputs/history.cpp:135-200).

Closed forms on states whose volume sums are known exactly or to round-off, chosen so
that a wrong column, a dropped factor 0.5, the wrong momentum component in a kinetic
energy, a ghost cell included in a sum or a missing cell volume fails one of them:
- uniform state: every column is (value per unit volume) x (mesh volume);
- one velocity component at a time: only its momentum and kinetic-energy columns move;
- sinusoids over whole periods on the uniform grid: sum sin = 0 and the mean of sin^2 is
  1/2 exactly (up to round-off), so mass, momentum, kinetic, gravitational and magnetic
  energies take the values of the continuous integrals;
- ghost zones filled with large values change nothing;
- the isothermal layout has no total-energy column and moves grav-E one place left.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc

BOX = ((0.0, 2.0), (0.0, 1.0), (0.0, 0.5))
VOL = 2.0 * 1.0 * 0.5


def mesh(eos=0, sg=1):
    # logically cubic MeshBlocks, as the multigrid hierarchy of the phi column needs
    d = make_desc(nx=(16, 8, 8), bnx=(4, 4, 4), box=BOX, self_gravity=sg, eos=eos)
    return d, Oracle(d)


def cell_centres(d):
    x = [(np.arange(n) + 0.5) * (b[1] - b[0]) / n + b[0] for n, b in zip((d.nx1, d.nx2, d.nx3), BOX)]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    return X, Y, Z


def load(o, d, U, bcc, phi, ghost=None):
    from tests.conftest import load_agx
    agx = load_agx()
    for name, G in (("u", U), ("bcc", bcc), ("phi", phi[None])):
        a = agx.scatter_cells(d, G)
        if ghost is not None:   # every ghost cell set to a large value
            ng = d.nghost
            mask = np.ones(a.shape[2:], bool)
            mask[ng:-ng, ng:-ng, ng:-ng] = False
            a[:, :, mask] = ghost
        o.set(name, a)
    return o.history()


def test_uniform_state_columns():
    d, o = mesh()
    shp = (d.nx3, d.nx2, d.nx1)
    rho, v, E, B, phi = 1.7, (0.3, -0.4, 0.25), 2.9, (0.6, -0.2, 0.45), -1.3
    U = np.zeros((5,) + shp)
    U[0] = rho
    for n in range(3):
        U[1 + n] = rho * v[n]
    U[4] = E
    bcc = np.stack([np.full(shp, b) for b in B])
    h = load(o, d, U, bcc, np.full(shp, phi), ghost=1e30)
    want = [rho * VOL] + [rho * x * VOL for x in v] + [0.5 * rho * x * x * VOL for x in v] + [E * VOL] + \
        [0.5 * rho * phi * VOL] + [0.5 * b * b * VOL for b in B]
    assert h.shape == (12,)
    np.testing.assert_allclose(h, want, rtol=1e-13, atol=0)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_one_velocity_component(axis):
    d, o = mesh()
    shp = (d.nx3, d.nx2, d.nx1)
    U = np.zeros((5,) + shp)
    U[0] = 2.0
    U[1 + axis] = 2.0 * 0.5
    h = load(o, d, U, np.zeros((3,) + shp), np.zeros(shp))
    for n in range(3):
        assert h[1 + n] == pytest.approx(2.0 * 0.5 * VOL if n == axis else 0.0, abs=1e-15)
        assert h[4 + n] == pytest.approx(0.5 * 2.0 * 0.25 * VOL if n == axis else 0.0, abs=1e-15)


def test_sinusoids_whole_periods():
    d, o = mesh()
    X, Y, Z = cell_centres(d)
    A, v0, P0, b0 = 0.3, 0.7, 0.2, 0.9
    sx = np.sin(2 * np.pi * X / 2.0)    # one period in x1
    sy = np.sin(2 * np.pi * Y / 1.0)    # one period in x2
    sz = np.sin(2 * np.pi * Z / 0.5)    # one period in x3
    U = np.zeros((5,) + X.shape)
    U[0] = 1.0 + A * sx                 # mass = VOL
    U[1] = v0 * sy                      # rho * vx with rho = 1 + A sx: sum = 0
    U[1] *= 1.0 + A * sx
    U[4] = P0 * (1.0 + sz)
    bcc = np.stack([b0 * sz, b0 * sx, 0.5 * b0 * sy])
    phi = 1.5 * sx                      # 0.5 sum rho phi vol = 0.5 * 1.5 * A * VOL / 2
    h = load(o, d, U, bcc, phi)
    assert h[0] == pytest.approx(VOL, rel=1e-14)
    assert abs(h[1]) < 1e-14 and abs(h[2]) < 1e-15 and abs(h[3]) < 1e-15
    # KE_x = 0.5 sum (rho vx)^2 / rho vol = 0.5 v0^2 sum (1 + A sx) sy^2 vol = 0.25 v0^2 VOL
    assert h[4] == pytest.approx(0.25 * v0 ** 2 * VOL, rel=1e-13)
    assert h[5] == 0.0 and h[6] == 0.0
    assert h[7] == pytest.approx(P0 * VOL, rel=1e-14)
    assert h[8] == pytest.approx(0.5 * 1.5 * A * VOL / 2, rel=1e-13)
    np.testing.assert_allclose(h[9:], [0.25 * b0 ** 2 * VOL, 0.25 * b0 ** 2 * VOL, 0.0625 * b0 ** 2 * VOL],
                               rtol=1e-13)


def test_isothermal_and_no_gravity_layout():
    shp = (8, 8, 16)
    U = np.zeros((5,) + shp)
    U[0] = 3.0
    U[2] = 3.0 * 0.2
    U[4] = 123.0                        # no energy in the isothermal state: never summed
    bcc = np.stack([np.full(shp, 0.1), np.full(shp, 0.2), np.full(shp, 0.3)])
    d, o = mesh(eos=1, sg=1)
    h = load(o, d, U, bcc, np.full(shp, 2.0))
    # mass, 3 mom, 3 KE, grav-E, 3 ME
    assert h.shape == (11,)
    assert h[7] == pytest.approx(0.5 * 3.0 * 2.0 * VOL, rel=1e-13)
    np.testing.assert_allclose(h[8:], [0.5 * 0.01 * VOL, 0.5 * 0.04 * VOL, 0.5 * 0.09 * VOL], rtol=1e-13)
    d, o = mesh(eos=0, sg=0)
    h = load(o, d, U, bcc, np.zeros(shp))
    # mass, 3 mom, 3 KE, tot-E, 3 ME
    assert h.shape == (11,)
    assert h[7] == pytest.approx(123.0 * VOL, rel=1e-13)
    assert h[5] == pytest.approx(0.5 * 3.0 * 0.04 * VOL, rel=1e-13)
