"""Pins the oracle's EMF correction across MeshBlocks (oracle/agx_oracle_refine.c:
orc_ref_emf_same_level / _to_coarser / _set / _clear / _average, orc_ref_count_fine_edges)
and the host neighbour search and work lists (athena-gamma_amd/smr.py).

The reference corrects the edge EMFs after each stage so that the face fluxes of
neighbouring blocks stay consistent (flux_correction_fc.cpp:623-1660): blocks on one
level end with the same EMF on every shared edge, and a coarse block's EMF on an edge
that finer blocks cover is their length-weighted mean.  Checks, on a periodic mesh of
2^3 root blocks with one of them refined (15 blocks, faces, edges and edge-only contacts
between levels):
  - a linear EMF field (already consistent everywhere) comes out unchanged, bit for
    bit, on the two-level mesh and on a uniform mesh;
  - random EMFs come out consistent: one value per shared edge on each level, and
    coarse = mean of the covering fine edges, over every coarse edge at the interface
    (768);
  - each planted mistake (no clear, no average, quadrants swapped) breaks one of these;
  - CountFineEdges on known neighbourhoods, and the oracle, the host module and the
    library's host function agree on every block of the mesh.
"""
import numpy as np
import pytest

from oracle import orc
from tests import emf_cases as ec
from tests.conftest import load_smr

NX, NG = 8, 2


@pytest.fixture(scope="module")
def mesh():
    t = ec.tree()
    return t, ec.coords(t, NX, NG)


def test_linear_field_unchanged_two_levels(mesh):
    t, tab = mesh
    E = ec.linear_emfs(t, NX, NG, tab)
    out, _ = ec.run_oracle(orc, t, E, NX, NG, tab)
    for a in range(3):
        assert np.array_equal(out[a], E[a]), f"e{a + 1}"


def test_linear_field_unchanged_uniform():
    t = ec.tree(refine=())
    tab = ec.coords(t, NX, NG)
    E = ec.linear_emfs(t, NX, NG, tab)
    plan = load_smr().emf_plan(t)
    assert not plan.to_coarser and not plan.clear and not plan.set_finer
    out, _ = ec.run_oracle(orc, t, E, NX, NG, tab, plan=plan)
    for a in range(3):
        assert np.array_equal(out[a], E[a]), f"e{a + 1}"


def test_random_emfs_become_consistent(mesh):
    t, tab = mesh
    E = ec.random_emfs(len(t.leaves), NX, NG, seed=5)
    spread0, err0, n0 = ec.check_consistency(t, E, NX, NG)
    assert spread0 > 0.1 and err0 > 0.1            # inconsistent before
    out, _ = ec.run_oracle(orc, t, E, NX, NG, tab)
    spread, err, ncmp = ec.check_consistency(t, out, NX, NG)
    print(f"same-level spread {spread:.2e}, coarse vs restricted fine {err:.2e} over {ncmp} edges", flush=True)
    # every coarse edge on the refined block's surface: 6 faces x 2 x 9 x 8, less the
    # 12 x 8 edges two faces share
    assert ncmp == n0 == 6 * 144 - 12 * 8
    assert spread < 4e-15 and err < 4e-15


@pytest.mark.parametrize("mistake", ["clear", "average", "swap_quadrants"])
def test_planted_mistakes_are_caught(mesh, mistake):
    t, tab = mesh
    E = ec.linear_emfs(t, NX, NG, tab)
    plan = load_smr().emf_plan(t)
    skip = ()
    if mistake == "swap_quadrants":
        for d in plan.set_finer:
            if d["fid"] < 6:
                d["fi1"], d["fi2"] = d["fi2"], d["fi1"]
    else:
        skip = (mistake,)
    out, _ = ec.run_oracle(orc, t, E, NX, NG, tab, skip=skip, plan=plan)
    assert any(not np.array_equal(out[a], E[a]) for a in range(3))


def lv(**pos):
    """27 relative levels: 0 everywhere (same level), overrides by name 'oxoyoz'."""
    v = [0] * 27
    for name, level in pos.items():
        o = [{"m": -1, "z": 0, "p": 1}[c] for c in name]
        v[((o[2] + 1) * 3 + o[1] + 1) * 3 + o[0] + 1] = level
    return v


def test_count_fine_edges_known():
    # all neighbours on the same level: every edge has 4 blocks, none finer
    f, n = orc.ref_count_fine_edges(lv())
    assert f == [True] * 12 and n == [4] * 12
    # a finer x1- face: the four edges on that face see one finer block position
    f, n = orc.ref_count_fine_edges(lv(mzz=1))
    on = [0, 2, 4, 6]          # x1x2 edges at ox1=-1 (eid 0, 2) and x1x3 edges (4, 6)
    assert [e for e in range(12) if not f[e]] == on and all(n[e] == 1 for e in on)
    assert all(n[e] == 4 for e in range(12) if e not in on)
    # a finer x1x2 edge neighbour (-1, -1, 0) alone: edge 0 only
    f, n = orc.ref_count_fine_edges(lv(mmz=1))
    assert [e for e in range(12) if not f[e]] == [0] and n[0] == 1
    # finer x1- face and x2- face around edge 0: two finer positions there
    f, n = orc.ref_count_fine_edges(lv(mzz=1, zmz=1))
    assert n[0] == 2 and not f[0]
    # coarser or missing neighbours (-1) are not counted
    f, n = orc.ref_count_fine_edges(lv(mzz=-1, mmz=-1))
    assert f[0] and n[0] == 2


def test_count_fine_edges_three_ways(mesh, agx):
    t, _ = mesh
    smr = load_smr()
    for g in range(len(t.leaves)):
        nbl = t.nblevel(g)
        a = orc.ref_count_fine_edges(nbl)
        assert smr.count_fine_edges(nbl) == a
        assert agx.ref_count_fine_edges(nbl) == a


def test_neighbor_lists_reciprocal(mesh):
    """every face/edge neighbour relation is seen from both sides, with the finer side's
    quadrant covering each coarse face exactly once"""
    t, _ = mesh
    seen = {}
    for g in range(len(t.leaves)):
        for nb in t.neighbors(g):
            seen.setdefault((g, nb.gid, nb.ox), []).append(nb)
    for (g, h, ox), nbs in seen.items():
        back = [n for n in t.neighbors(h) if n.gid == g and n.ox == tuple(-o for o in ox)]
        rel = nbs[0].rel
        if rel == 0 or rel == 1:
            assert len(back) == 1 and back[0].rel == -rel, (g, h, ox)
    # coarse block 0 (root (1,0,0)) has its x1- and x1+ faces on the refined root block
    quads = sorted((n.fid, n.fi1, n.fi2) for n in t.neighbors(0) if n.rel == 1 and n.kind == "face")
    assert quads == sorted((f, a, b) for f in (0, 1) for a in (0, 1) for b in (0, 1))


def test_random_emfs_consistent_with_physical_boundaries():
    """4^3 root blocks, outflow boundaries, two refined root blocks side by side (fine
    blocks of two refined regions meet): the same invariants; faces on the domain
    boundary are not averaged (AverageFluxBoundary skips them)"""
    from tests import ghost_cases as gc
    t = gc.tree(((1, 1, 1), (2, 1, 1)))
    tab = load_smr().block_coords(t, NX, NG)
    E = ec.random_emfs(len(t.leaves), NX, NG, seed=21)
    out, _ = ec.run_oracle(orc, t, E, NX, NG, tab)
    spread, err, ncmp = ec.check_consistency(t, out, NX, NG)
    print(f"spread {spread:.2e}, coarse vs fine {err:.2e} over {ncmp} edges", flush=True)
    assert ncmp > 1000 and spread < 4e-15 and err < 4e-15
