"""MHD on a static-refinement mesh on the device (SMRMesh: agx_mesh_set_levels, the
task-level steps and the refinement operators on the mesh's own arrays, in the reference's
task order with flux and EMF correction, time_integrator.cpp:1460-1558) against:
  - the oracle's composition of the same cycle (tests/smr_cases.py, pinned by
    test_oracle_smr_cycle.py) on one, two and three levels: identical bit for bit;
  - the single-level lattice cycle (agx_cycle) on the same blocks: bit for bit for a
    state given as global arrays (the same-level EMF average then averages equal values,
    exactly), within round-off for blocks built one by one (the two copies of a
    periodic-boundary face differ by rounding, and only the reference's sequence averages
    the edges they lead to);
  - what the corrections guarantee on a periodic mesh: mass, momentum and energy summed
    over all blocks conserved to round-off (flux correction), the field through every
    refinement boundary equal on both sides (EMF correction), div B at round-off."""
import numpy as np
import pytest

from tests import smr_cases as sc

pytestmark = pytest.mark.gpu
BOX = ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))
FIELDS = ("u", "w", "bcc", "b1f", "b2f", "b3f")


def gpu_run(agx, tree, nx, state, ncyc, xorder=2, rsolver=0, ng=2):
    m = agx.SMRMesh(tree, nx, ng, BOX, cfl=0.3, xorder=xorder, rsolver=rsolver)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), state):
            m.set(name, a)
        m.initialize()
        assert m.cycle(1e9, ncyc) == ncyc
        out = {f: m.get(f) for f in FIELDS}
        out["time"] = m.time()
        return out
    finally:
        m.close()


def oracle_run(tree, nx, state, ncyc, xorder=2, rsolver=0, ng=2):
    o = sc.OracleSMR(tree, nx, ng, BOX, cfl=0.3, xorder=xorder, rsolver=rsolver)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), state):
            o.set(name, a)
        o.initialize()
        assert o.cycle(1e9, ncyc) == ncyc
        out = {f: o.get(f) for f in FIELDS}
        out["time"] = o.time()
        return out
    finally:
        o.close()


def same(a, b, what):
    ok = np.array_equal(a, b)
    if not ok:
        d = np.abs(a - b)
        print(f"{what}: max |diff| {np.nanmax(d):.3e} at {np.unravel_index(np.nanargmax(d), d.shape)}", flush=True)
    assert ok, what


def test_single_level_global_arrays_equal_lattice(agx):
    """a state given as global arrays: the reference's sequence with the EMF correction
    (SMRMesh) and the lattice cycle (agx_cycle, fused stages) agree bit for bit"""
    from tests.states import load_state, turbulent_state
    tree = sc.periodic_tree((2, 2, 2))
    desc = agx.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA)
    st = turbulent_state(desc, seed=3, mach=1.5)
    s = agx.SMRMesh(tree, 8, 2, BOX, cfl=0.3)
    m = agx.Mesh(desc)
    try:
        load_state(s.m, st)
        load_state(m, st)
        s.initialize()
        m.initialize()
        assert s.cycle(1e9, 4) == 4
        assert m.cycle(1e9, 4) == 4
        assert s.time() == m.time()
        for f in FIELDS:
            same(s.get(f), m.get(f), f"global arrays: {f}")
    finally:
        s.close()
        m.close()


@pytest.mark.parametrize("tile", [1, 5])
def test_lattice_with_unequal_face_copies_equals_oracle(agx, tile):
    """blocks built one by one (the copies of a periodic-boundary face differ by rounding):
    the lattice cycle converts its ghosts from each block's own faces, as the oracle's
    lattice cycle does (no primitive push), bit for bit, split and one-launch forms"""
    from oracle import orc
    tree = sc.periodic_tree((2, 2, 2))
    u, b1, b2, b3 = sc.smooth_state(tree, 8, 2, BOX)
    assert not np.array_equal(b1[0][:, :, 2], b1[1][:, :, 10])   # the x = 0 and x = 1 copies differ
    desc = agx.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA)
    odesc = orc.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA)
    m = agx.Mesh(desc)
    o = orc.Oracle(odesc)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), (u, b1, b2, b3)):
            m.set(name, a)
            o.set(name, a)
        m.set_tile(tile)
        m.initialize()
        o.initialize()
        assert m.cycle(1e9, 3) == 3
        for _ in range(3):
            o.cycle(1e9)
        assert m.time() == o.time()
        for f in FIELDS:
            same(m.get(f), o.get(f), f"unequal copies, tile {tile}: {f}")
    finally:
        m.close()
        o.close()


def test_single_level_equals_oracle_and_lattice(agx):
    tree = sc.periodic_tree((2, 2, 2))
    nx = 8
    state = sc.smooth_state(tree, nx, 2, BOX)
    got = gpu_run(agx, tree, nx, state, 3)
    ref = oracle_run(tree, nx, state, 3)
    assert got["time"] == ref["time"]
    for f in FIELDS:
        same(got[f], ref[f], f"one level: {f}")
    desc = agx.make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, nghost=2, gamma=sc.GAMMA)
    m = agx.Mesh(desc)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), state):
            m.set(name, a)
        m.initialize()
        assert m.cycle(1e9, 3) == 3
        for f in FIELDS:
            lat = m.get(f)
            assert np.abs(got[f] - lat).max() <= 1e-13 * np.abs(lat).max(), f
    finally:
        m.close()


@pytest.mark.parametrize("levels", [2, 3])
def test_refined_mesh_equals_oracle(agx, levels):
    if levels == 2:
        tree, nx = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0),)), 8
    else:
        inner = tuple((x, y, z) for z in (1, 2) for y in (1, 2) for x in (1, 2))
        tree, nx = sc.periodic_tree((4, 4, 4), refine=inner, refine2=((3, 3, 3),)), 4
    print(f"{levels} levels: {len(tree.leaves)} blocks of {nx}^3", flush=True)
    state = sc.smooth_state(tree, nx, 2, BOX)
    got = gpu_run(agx, tree, nx, state, 3)
    ref = oracle_run(tree, nx, state, 3)
    assert got["time"] == ref["time"]
    for f in FIELDS:
        same(got[f], ref[f], f"{levels} levels: {f}")


@pytest.mark.parametrize("rsolver", [0, 2, 3])
def test_refined_ppm_equals_oracle(agx, rsolver):
    """PPM (NGHOST = 4, the even width mesh refinement needs) with HLLD, Roe and LHLLD"""
    tree, nx, ng = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0),)), 8, 4
    state = sc.smooth_state(tree, nx, ng, BOX, amp=0.3)
    got = gpu_run(agx, tree, nx, state, 3, xorder=3, rsolver=rsolver, ng=ng)
    ref = oracle_run(tree, nx, state, 3, xorder=3, rsolver=rsolver, ng=ng)
    assert got["time"] == ref["time"]
    for f in FIELDS:
        same(got[f], ref[f], f"PPM rsolver {rsolver}: {f}")


def test_refined_isothermal_equals_oracle(agx):
    tree, nx = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0),)), 8
    u, b1, b2, b3 = sc.smooth_state(tree, nx, 2, BOX, amp=0.3, iso=True)
    m = agx.SMRMesh(tree, nx, 2, BOX, cfl=0.3, eos=1, iso_cs=0.8)
    o = sc.OracleSMR(tree, nx, 2, BOX, cfl=0.3, eos=1, iso_cs=0.8)
    try:
        m.set("u", u[:, :4])
        o.set("u", u)
        for name, a in zip(("b1f", "b2f", "b3f"), (b1, b2, b3)):
            m.set(name, a)
            o.set(name, a)
        m.initialize()
        o.initialize()
        assert m.cycle(1e9, 3) == 3
        assert o.cycle(1e9, 3) == 3
        assert m.time() == o.time()
        same(m.get("u"), o.get("u")[:, :4], "isothermal: u")
        same(m.get("w")[:, :4], o.get("w")[:, :4], "isothermal: w")
        for f in ("bcc", "b1f", "b2f", "b3f"):
            same(m.get(f), o.get(f), f"isothermal: {f}")
    finally:
        m.close()
        o.close()


def test_odd_ghost_width_refused(agx):
    tree = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0),))
    with pytest.raises(agx.AgxError, match="odd ghost width"):
        agx.SMRMesh(tree, 8, 3, BOX, xorder=3)


@pytest.mark.parametrize("xorder,rsolver", [(2, 0), (2, 1), (1, 0)])
def test_refined_mesh_conserves_and_keeps_interfaces(agx, xorder, rsolver):
    tree, nx, ng = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0), (1, 1, 1))), 8, 2
    state = sc.smooth_state(tree, nx, ng, BOX, amp=0.3)
    m = agx.SMRMesh(tree, nx, ng, BOX, cfl=0.3, xorder=xorder, rsolver=rsolver)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), state):
            m.set(name, a)
        m.initialize()
        t0 = sc.totals(tree, nx, ng, BOX, m.get("u"))
        assert m.cycle(1e9, 10) == 10
        t1 = sc.totals(tree, nx, ng, BOX, m.get("u"))
        b = [m.get(n) for n in ("b1f", "b2f", "b3f")]
    finally:
        m.close()
    rel = float((np.abs(t1 - t0) / np.maximum(np.abs(t0), 1.0)).max())
    divb, mis = sc.max_divb(tree, nx, ng, BOX, *b), sc.interface_mismatch(tree, nx, ng, *b)
    print(f"xorder {xorder} rsolver {rsolver}: conservation {rel:.2e}, div B {divb:.2e}, interface {mis:.2e}",
          flush=True)
    assert rel < 1e-14
    assert divb < 1e-14
    assert mis < 1e-14


def test_lattice_entry_points_refuse_a_refined_mesh(agx):
    tree = sc.periodic_tree((2, 2, 2), refine=((0, 0, 0),))
    m = agx.SMRMesh(tree, 8, 2, BOX)
    try:
        for call in (lambda: m.m.cycle(1e9, 1), lambda: m.m.initialize(), lambda: m.m.bvals(3),
                     lambda: m.m.autotune()):
            with pytest.raises(agx.AgxError, match="refinement levels"):
                call()
    finally:
        m.close()


def test_uneven_blocks_on_three_levels_equal_oracle(agx):
    """MeshBlocks of 16^3 and 32^3 cells on three levels (smr.uneven_leaves: 7 root blocks
    of 32^3, 7 level-1 blocks of 32^3 and 7 of 16^3, 8 level-2 blocks of 16^3).  The state
    is given per uneven block, split into the 127 leaves of 16^3 that one device mesh holds,
    run 3 cycles and merged back per block: equal bit for bit to the oracle's run of the
    same mesh as 16^3 blocks."""
    from tests.conftest import load_smr
    from tests.test_smr_uneven import three_level_blocks
    smr = load_smr()
    nx, ng = 16, 2
    blocks = three_level_blocks()
    leaves, owner = smr.uneven_leaves(blocks)
    tree = smr.SMRTree((4, 4, 4), leaves)
    names = ("u", "b1f", "b2f", "b3f")
    leaf_state = sc.smooth_state(tree, nx, ng, BOX, amp=0.3)
    per_block = [smr.merge_uneven(blocks, owner, n, a, nx, ng) for n, a in zip(names, leaf_state)]
    assert {(a.shape[-1] - 2 * ng) for a in per_block[0]} == {16, 32}
    state = [smr.split_uneven(blocks, owner, n, arrs, nx, ng) for n, arrs in zip(names, per_block)]
    for n, a, b in zip(names, state, leaf_state):
        inner = (slice(None),) * (a.ndim - 3) + smr._interior(n, nx, ng)
        assert np.array_equal(a[inner], b[inner]), n
    got = gpu_run(agx, tree, nx, state, 3)
    ref = oracle_run(tree, nx, state, 3)
    assert got["time"] == ref["time"]
    for f in FIELDS:
        same(got[f], ref[f], f"uneven blocks: {f}")
        mg = smr.merge_uneven(blocks, owner, f, got[f], nx, ng)
        mo = smr.merge_uneven(blocks, owner, f, ref[f], nx, ng)
        for bi, (a, b) in enumerate(zip(mg, mo)):
            same(a, b, f"uneven block {bi}: {f}")
