"""Pins the multigrid source mask (MultigridDriver::SetupMultigrid, multigrid_driver.cpp:
430-436, with the collapse problem's spherical mask, pgen/collapse.cpp:66-82) in the
oracle: masking the loaded source is the same solve as loading a density that is
already zero outside the sphere, bit for bit, for periodic and multipole boundaries; the
monopole of the masked solve is the mass inside the sphere; radius <= 0 switches it off.
"""
import math

import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_agx
from tests.test_oracle_mg_multipole import BOX, centres, gaussian

N, B = 32, 16
RC, C = 0.3137, (0.05, -0.03, 0.02)     # no cell centre within 1e-6 of the sphere


def inside(n, c=C, rc=RC):
    x = centres(n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    return (X - c[0]) ** 2 + (Y - c[1]) ** 2 + (Z - c[2]) ** 2 <= rc * rc


def solve(rho, mask, **kw):
    agx = load_agx()
    d = make_desc(nx=(N, N, N), bnx=(B, B, B), box=BOX, self_gravity=1, mg_mode=0, four_pi_G=1.0,
                  mg_threshold=-1.0, mg_niter=4, **kw)
    o = Oracle(d)
    U = np.zeros((5, N, N, N))
    U[0] = rho
    o.set("u", agx.scatter_cells(d, U))
    if mask:
        o.mg_set_source_mask_sphere(*C, RC)
    o.mg_solve(0)
    return o, agx.gather_cells(d, o.get("phi"))[0]


@pytest.mark.parametrize("bc", [dict(mg_bc=0), dict(mg_bc=1, mporder=4, mp_auto_origin=0)])
def test_mask_equals_masked_density(bc):
    rho = 1.0 + gaussian(N, 0.1, c=(0.1, 0.0, -0.05), amp=5.0)
    ins = inside(N)
    x = centres(N)
    dist = np.abs(np.sqrt(((np.stack(np.meshgrid(x, x, x, indexing="ij"))[::-1]
                            - np.array(C)[:, None, None, None]) ** 2).sum(0)) - RC)
    assert dist.min() > 1e-6 and 0 < ins.sum() < N ** 3
    _, a = solve(rho, True, **bc)
    _, b = solve(np.where(ins, rho, 0.0), False, **bc)
    assert np.array_equal(a, b)
    _, c = solve(rho, False, **bc)
    assert not np.array_equal(a, c)


def test_masked_monopole_is_the_mass_inside():
    rho = 1.0 + gaussian(N, 0.1, amp=3.0)
    o, _ = solve(rho, True, mg_bc=1, mporder=4, mp_auto_origin=0)
    _, coef = o.mg_multipole()
    mass = (rho * inside(N)).sum() / N ** 3
    assert coef[0] == pytest.approx(-mass / (4 * math.pi), rel=1e-12)


def test_radius_zero_switches_off():
    rho = 1.0 + gaussian(N, 0.1, amp=3.0)
    agx = load_agx()
    d = make_desc(nx=(N, N, N), bnx=(B, B, B), box=BOX, self_gravity=1, mg_mode=0, four_pi_G=1.0,
                  mg_threshold=-1.0, mg_niter=2)
    U = np.zeros((5, N, N, N))
    U[0] = rho
    o1, o2 = Oracle(d), Oracle(d)
    for o in (o1, o2):
        o.set("u", agx.scatter_cells(d, U))
    o1.mg_set_source_mask_sphere(0.0, 0.0, 0.0, 0.2)
    o1.mg_set_source_mask_sphere(0.0, 0.0, 0.0, 0.0)
    o1.mg_solve(0)
    o2.mg_solve(0)
    assert np.array_equal(o1.get("phi"), o2.get("phi"))
