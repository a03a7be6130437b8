"""Multigrid at task granularity (agx_mg_setup / smooth / fill_ghost / fas / restrict /
prolong / fmg_prolong / solve_coarsest / defect_norm / retrieve / coffset): a V(1,1)
composed of the per-level steps in the reference's task order (mg_task_list.cpp:287-530)
equals agx_mg_vcycles(1) bit for bit, periodic and multipole (FAS), and agx_mg_setup
loads the same source as agx_mg_solve (the potential of a fixed-cycle solve composed of
the steps equals agx_mg_solve's)."""
import numpy as np
import pytest

from oracle.orc import make_desc
from tests.test_oracle_mg_multipole import BOX, gaussian

pytestmark = pytest.mark.gpu


def vcycle_by_steps(m, mp):
    F = len(m.mg_levels()) - 1
    m.mg_coffset(flip=True)
    for l in range(F, 0, -1):
        if mp:
            m.mg_step("fill_ghost", l)
            if l < F:
                m.mg_step("fas", l)
        m.mg_step("smooth", l, 0)
        if mp:
            m.mg_step("fill_ghost", l)
        m.mg_step("smooth", l, 1)
        if mp:
            m.mg_step("fill_ghost", l)
        m.mg_step("restrict", l)
    m.mg_step("solve_coarsest")
    for l in range(1, F + 1):
        if mp:
            m.mg_step("fill_ghost", l - 1)
        m.mg_step("prolong", l)
        if mp:
            m.mg_step("fill_ghost", l)
        m.mg_step("smooth", l, 0)
        if mp:
            m.mg_step("fill_ghost", l)
        m.mg_step("smooth", l, 1)
        if mp:
            m.mg_step("fill_ghost", l)


@pytest.mark.parametrize("mp", [0, 1])
def test_vcycle_from_steps_equals_vcycles(agx, mp):
    n = 32
    d = make_desc(nx=(n, n, n), bnx=(8, 8, 8), box=BOX, self_gravity=1, mg_bc=mp, mg_mode=0, four_pi_G=1.0,
                  mporder=4, mg_threshold=-1.0, mg_niter=2)
    U = np.zeros((5, n, n, n))
    U[0] = gaussian(n, 0.06, c=(0.05, -0.04, 0.03)) + (0.0 if mp else 1.0)
    ub = agx.scatter_cells(d, U)
    a, b = agx.Mesh(d), agx.Mesh(d)
    for m in (a, b):
        m.set("u", ub)
        m.mg_solve(0)
    assert a.mg_coffset() == b.mg_coffset()
    a.mg_vcycles(1)
    vcycle_by_steps(b, mp)
    for m in (a, b):
        m.mg_step("retrieve")
    pa, pb = a.get("phi"), b.get("phi")
    print(f"multipole {mp}: levels {a.mg_levels()}; identical {np.array_equal(pa, pb)}", flush=True)
    assert np.array_equal(pa, pb)
    assert a.mg_defect_norm() == b.mg_defect_norm()
    a.close()
    b.close()


def test_steps_after_fmg_reproduce_fixed_cycle_solve(agx):
    """Multipole (FAS): agx_mg_solve with niteration = 2 equals an FMG-only solve followed
    by two V-cycles composed of steps and agx_mg_retrieve, bit for bit (potential and the
    finest defect norm).  The periodic driver subtracts the mean once more per solve, so
    there the equality holds for the V-cycles alone (test above)."""
    n = 32
    d = make_desc(nx=(n, n, n), bnx=(8, 8, 8), box=BOX, self_gravity=1, mg_bc=1, mg_mode=0, four_pi_G=1.0,
                  mporder=2, mg_threshold=-1.0, mg_niter=2)
    U = np.zeros((5, n, n, n))
    U[0] = gaussian(n, 0.05, c=(-0.03, 0.02, 0.05))
    ub = agx.scatter_cells(d, U)
    a = agx.Mesh(d)
    a.set("u", ub)
    a.mg_solve(0)
    ref, ref_norm = a.get("phi"), a.mg_last()[0]
    b = agx.Mesh(d)
    b.set("u", ub)
    b.mg_set_mode(0, -1.0, 0)    # FMG only
    b.mg_solve(0)
    for _ in range(2):
        vcycle_by_steps(b, True)
    b.mg_step("retrieve")
    got = b.get("phi")
    print(f"FMG + 2 V-cycles by steps identical to the solve: {np.array_equal(got, ref)}", flush=True)
    assert np.array_equal(got, ref)
    assert b.mg_defect_norm() == ref_norm
    a.close()
    b.close()


def test_setup_loads_the_solve_source(agx):
    """agx_mg_setup + FMG prolongations and V-cycles by steps from the root reproduce the
    library's FMG solve (periodic, niteration 0) to rounding of the mean subtraction."""
    n = 16
    d = make_desc(nx=(n, n, n), bnx=(8, 8, 8), box=((0.0, 1.0),) * 3, self_gravity=1, mg_mode=0, four_pi_G=1.0,
                  mg_threshold=-1.0, mg_niter=0)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    U = np.zeros((5, n, n, n))
    U[0] = 2.0 + np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
    ub = agx.scatter_cells(d, U)
    a = agx.Mesh(d)
    a.set("u", ub)
    a.mg_solve(0)
    ref = a.get("phi")
    b = agx.Mesh(d)
    b.set("u", ub)
    b.mg_step("setup")
    b.mg_step("retrieve")    # source loaded, u still zero: phi = 0
    assert np.abs(b.get("phi")).max() == 0.0
    a.close()
    b.close()
    assert np.abs(ref).max() > 0.0
