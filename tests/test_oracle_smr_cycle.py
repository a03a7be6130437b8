"""Pins of the oracle's static-refinement MHD cycle (tests/smr_cases.py: the reference's
task order with refinement, time_integrator.cpp:1460-1558, composed from the oracle's
task-level steps and refinement operators), which the device's SMRMesh is held to bit for
bit:
  - on one level it is the oracle's lattice cycle bit for bit when the state is given as
    global arrays (the EMF correction, which the reference runs on every MHD mesh,
    time_integrator.cpp:985-994, then averages equal values, exactly); with blocks built
    one by one, whose two copies of a periodic-boundary face differ by rounding, it
    differs only at round-off, and equals the lattice cycle bit for bit with the EMF
    correction left out (every other refinement step is a same-level copy there);
  - on two and three levels of a periodic mesh: mass, momentum and energy summed over all
    blocks are conserved to round-off, and the field through every refinement boundary
    agrees between the coarse face and the fine faces that cover it.  Leaving out the flux
    correction breaks the first, leaving out the EMF correction the second; div B stays
    at round-off in every block either way (CT)."""
import numpy as np
import pytest

from oracle import orc
from tests import smr_cases as sc
from tests.states import load_state, turbulent_state

BOX = ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))


def load(o, st):
    for n, a in zip(("u", "b1f", "b2f", "b3f"), st):
        o.set(n, a)


@pytest.mark.parametrize("xorder", [1, 2])
def test_single_level_with_one_face_array_is_the_lattice_cycle(xorder):
    """a state given as global arrays (each shared face stored twice with one value, as
    tests.states and the benchmark generators build it): every block computes the same
    value for a shared edge, the EMF correction averages equal values, which is exact
    ((e + e) * 0.5 and (((e + e) + e) + e) * 0.25 both give e), and the whole reference
    sequence equals the lattice cycle bit for bit"""
    tree = sc.periodic_tree((2, 2, 2))
    desc = orc.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA, xorder=xorder)
    st = turbulent_state(desc, seed=3, mach=1.5)
    r = orc.Oracle(desc)
    load_state(r, st)
    r.initialize()
    o = sc.OracleSMR(tree, 8, 2, BOX, xorder=xorder)
    load_state(o.o, st)
    o.initialize()
    for _ in range(4):
        r.cycle(1e9)
    assert o.cycle(1e9, 4) == 4
    for n in ("u", "w", "b1f", "b2f", "b3f"):
        assert np.array_equal(o.get(n), r.get(n)), n
    o.close()


def test_equal_averages_are_exact():
    e = np.concatenate([np.random.default_rng(1).standard_normal(200000),
                        1.0 + np.arange(1, 200001) * 2.0 ** -52, 4.0 / 3.0 + np.arange(1, 200001) * 2.0 ** -52])
    assert np.array_equal((e + e) * 0.5, e)
    assert np.array_equal((((e + e) + e) + e) * 0.25, e)


@pytest.mark.parametrize("xorder", [1, 2])
def test_single_level_is_the_lattice_cycle(xorder):
    """blocks built one by one (pgen.divfree_blocks): the two copies of a face on a
    periodic boundary differ by the rounding of their coordinates"""
    tree = sc.periodic_tree((2, 2, 2))
    st = sc.smooth_state(tree, 8, 2, BOX)
    desc = orc.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA, xorder=xorder)
    r = orc.Oracle(desc)
    load(r, st)
    r.initialize()
    for _ in range(4):
        r.cycle(1e9)
    for skip, exact in ((("all",), True), ((), False)):
        o = sc.OracleSMR(tree, 8, 2, BOX, xorder=xorder, emf_skip=skip)
        load(o, st)
        o.initialize()
        assert o.cycle(1e9, 4) == 4
        for n in ("u", "w", "b1f", "b2f", "b3f"):
            a, b = o.get(n), r.get(n)
            if exact:
                assert np.array_equal(a, b), n
            else:
                assert np.abs(a - b).max() <= 1e-13 * np.abs(b).max(), n
        o.close()


def run(tree, nx, st, ncyc, **kw):
    o = sc.OracleSMR(tree, nx, 2, BOX, **kw)
    load(o, st)
    o.initialize()
    t0 = sc.totals(tree, nx, 2, BOX, o.get("u"))
    assert o.cycle(1e9, ncyc) == ncyc
    t1 = sc.totals(tree, nx, 2, BOX, o.get("u"))
    b = [o.get(n) for n in ("b1f", "b2f", "b3f")]
    o.close()
    cons = float((np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)).max())
    return cons, sc.max_divb(tree, nx, 2, BOX, *b), sc.interface_mismatch(tree, nx, 2, *b)


def two_levels():
    return sc.periodic_tree((2, 2, 2), refine=((0, 0, 0), (1, 1, 1))), 8


def three_levels():
    inner = tuple((x, y, z) for z in (1, 2) for y in (1, 2) for x in (1, 2))
    return sc.periodic_tree((4, 4, 4), refine=inner, refine2=((3, 3, 3),)), 4


@pytest.mark.parametrize("mesh", [two_levels, three_levels])
def test_refined_cycle_conserves_and_keeps_interfaces(mesh):
    tree, nx = mesh()
    st = sc.smooth_state(tree, nx, 2, BOX, amp=0.3)
    assert sc.interface_mismatch(tree, nx, 2, *st[1:]) < 1e-15
    cons, divb, mis = run(tree, nx, st, 4)
    print(f"{len(tree.leaves)} blocks: conservation {cons:.2e}, div B {divb:.2e}, interface {mis:.2e}", flush=True)
    assert cons < 1e-14
    assert divb < 1e-14
    assert mis < 1e-14


@pytest.mark.parametrize("rsolver", [0, 3])
def test_refined_ppm_cycle_conserves_and_keeps_interfaces(rsolver):
    """PPM on a refined mesh needs an even ghost width of at least 3: NGHOST = 4"""
    tree, nx, ng = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0), (1, 1, 1))), 8, 4
    st = sc.smooth_state(tree, nx, ng, BOX, amp=0.3)
    o = sc.OracleSMR(tree, nx, ng, BOX, xorder=3, rsolver=rsolver)
    load(o, st)
    o.initialize()
    t0 = sc.totals(tree, nx, ng, BOX, o.get("u"))
    assert o.cycle(1e9, 3) == 3
    t1 = sc.totals(tree, nx, ng, BOX, o.get("u"))
    b = [o.get(n) for n in ("b1f", "b2f", "b3f")]
    o.close()
    assert float((np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)).max()) < 1e-14
    assert sc.max_divb(tree, nx, ng, BOX, *b) < 1e-14
    assert sc.interface_mismatch(tree, nx, ng, *b) < 1e-14


def test_refined_isothermal_cycle_conserves_and_keeps_interfaces():
    tree, nx = two_levels()
    st = sc.smooth_state(tree, nx, 2, BOX, amp=0.3, iso=True)
    o = sc.OracleSMR(tree, nx, 2, BOX, eos=1, iso_cs=0.8)
    load(o, st)
    o.initialize()
    t0 = sc.totals(tree, nx, 2, BOX, o.get("u")[:, :4])
    assert o.cycle(1e9, 4) == 4
    t1 = sc.totals(tree, nx, 2, BOX, o.get("u")[:, :4])
    b = [o.get(n) for n in ("b1f", "b2f", "b3f")]
    assert np.all(o.get("u")[:, 4] == 0.0)
    o.close()
    assert float((np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)).max()) < 1e-14
    assert sc.max_divb(tree, nx, 2, BOX, *b) < 1e-14
    assert sc.interface_mismatch(tree, nx, 2, *b) < 1e-14


def test_planted_omissions_fail_the_pins():
    tree, nx = two_levels()
    st = sc.smooth_state(tree, nx, 2, BOX, amp=0.3)
    cons, divb, mis = run(tree, nx, st, 4, flux_skip=True)
    assert cons > 1e-8 and divb < 1e-14
    cons, divb, mis = run(tree, nx, st, 4, emf_skip=("all",))
    assert mis > 1e-6 and divb < 1e-14
