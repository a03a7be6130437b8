"""hipGraph replay of the fused cycle (agx_set_graph) against the oracle and the eager path.

A captured cycle reads dt and time from device memory and advances them on the device,
and the face registers rotate with period two cycles, so one graph per register
assignment is replayed.  Each test replays the graphs several times with host work
(downloads, eager cycles) between replays and checks every replay: bit for bit against
the oracle in FP64, bit for bit against the eager GPU path in the mixed form.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu
FIELDS = ("u", "w", "b1f", "b2f", "b3f")


def turb_desc(**kw):
    return make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, **kw)


def same(a, b, name):
    d = np.abs(a - b).max()
    print(f"{name}: max|diff| {d:.3e}", flush=True)
    assert np.array_equal(a, b), name


@pytest.mark.parametrize("order", [2, 3])
def test_graph_cycles_match_oracle(agx, order):
    desc = turb_desc(xorder=order, nghost=3 if order == 3 else 2)
    st = turbulent_state(desc, seed=5, mach=1.5)
    o = Oracle(desc)
    m = agx.Mesh(desc)
    load_state(o, st)
    load_state(m, st)
    o.initialize()
    m.initialize()
    m.set_graph(True)
    total = 0
    # graph, graph again (both register parities), eager, graph: every replay checked
    for n, graph in ((3, True), (4, True), (1, False), (2, True)):
        m.set_graph(graph)
        assert m.cycle(1e9, n) == n
        for _ in range(n):
            o.cycle(1e9)
        total += n
        assert m.time() == o.time(), (m.time(), o.time())
        for f in FIELDS:
            same(m.get(f), o.get(f), f"{f} after {total} cycles (graph={graph})")


def test_graph_stops_at_tlim(agx):
    """The last step is clipped to tlim on the device, as Mesh::NewTimeStep does."""
    desc = turb_desc()
    st = turbulent_state(desc, seed=6, mach=1.0)
    o = Oracle(desc)
    m = agx.Mesh(desc)
    load_state(o, st)
    load_state(m, st)
    tlim = 0.013
    o.initialize()
    m.initialize(tlim)
    m.set_graph(True)
    nm = m.cycle(tlim, 1000)
    no = o.run(tlim)
    print(f"cycles: graph {nm}, oracle {no}; time {m.time()}", flush=True)
    assert m.time() == o.time()
    assert m.time()[0] == tlim
    for f in FIELDS:
        same(m.get(f), o.get(f), f)
    assert m.cycle(tlim, 5) == 0


def test_graph_mixed_equals_eager(agx):
    desc = turb_desc(precision=1)
    st = turbulent_state(desc, seed=7, mach=1.5)
    runs = []
    for graph in (False, True):
        m = agx.Mesh(desc)
        load_state(m, st)
        m.initialize()
        m.set_graph(graph)
        out = []
        for n in (2, 3, 2):
            assert m.cycle(1e9, n) == n
            out.append((m.time(), [m.get(f) for f in FIELDS]))
        runs.append(out)
        m.close()
    for (te, ae), (tg, ag) in zip(*runs):
        assert te == tg
        for f, a, b in zip(FIELDS, ae, ag):
            same(b, a, f"mixed {f}")


def test_graph_rebuilt_after_tile_change(agx):
    desc = turb_desc()
    st = turbulent_state(desc, seed=8, mach=1.5)
    o = Oracle(desc)
    m = agx.Mesh(desc)
    load_state(o, st)
    load_state(m, st)
    o.initialize()
    m.initialize()
    m.set_graph(True)
    for tile in (1, 5, 3, 4, 0):
        m.set_tile(tile)
        assert m.cycle(1e9, 2) == 2
        for _ in range(2):
            o.cycle(1e9)
        for f in FIELDS:
            same(m.get(f), o.get(f), f"{f} tile {tile}")


GRAV_FORMS = [(0, 0, 1), (1, 0, 1), (0, 1, 1), (1, 1, 1), (0, 1, 0), (1, 1, 0)]


def grav_desc(precision, mg_bc, auto_origin, **kw):
    """32^3 turbulence with multigrid self-gravity: periodic (mg_bc 0) or multipole
    boundaries about the centre of mass (auto_origin 1) or about a fixed origin"""
    mp = {} if auto_origin else dict(mp_auto_origin=0, mp_origin=(0.45, 0.5, 0.55))
    return turb_desc(self_gravity=1, four_pi_G=1.0, mg_mode=0, precision=precision, mg_bc=mg_bc, **mp, **kw)


@pytest.mark.parametrize("precision,mg_bc,auto_origin", GRAV_FORMS)
def test_graph_with_fixed_count_gravity_equals_eager(agx, precision, mg_bc, auto_origin):
    """Self-gravity with a fixed V-cycle count (niteration), periodic or multipole
    boundaries: both multigrid solves of a cycle are recorded with the MHD stages.  Graph
    replays, interleaved with eager cycles and downloads, equal the eager path bit for
    bit, potential included."""
    desc = grav_desc(precision, mg_bc, auto_origin, mg_threshold=-1.0, mg_niter=3)
    st = turbulent_state(desc, seed=9, mach=1.5, drho=0.5)
    runs = []
    for graph in (False, True):
        m = agx.Mesh(desc)
        load_state(m, st)
        m.initialize()
        snaps = []
        for n, g in ((2, graph), (3, graph), (1, False), (2, graph)):
            m.set_graph(g)
            assert m.cycle(1e9, n) == n
            snap = {f: m.get(f) for f in FIELDS + ("phi",)}
            snap["time"] = m.time()
            snap["iters"] = m.mg_last()[1]
            snaps.append(snap)
        if graph:
            norm, iters = m.mg_last()
            assert iters == 3 and np.isnan(norm)
        runs.append(snaps)
        m.close()
    for k, (a, b) in enumerate(zip(*runs)):
        assert a["time"] == b["time"], k
        assert a["iters"] == b["iters"] == 3
        for f in FIELDS + ("phi",):
            same(b[f], a[f], f"segment {k}: {f} (precision {precision}, mg_bc {mg_bc}, auto origin {auto_origin})")


@pytest.mark.parametrize("precision,mg_bc,auto_origin", GRAV_FORMS)
def test_graph_with_threshold_gravity_equals_eager(agx, precision, mg_bc, auto_origin):
    """Self-gravity solved to a threshold (SolveIterative): each solve's V-cycle loop is
    recorded with the count the same stage's last solve needed, gated on the device by the
    stopping rule, and continued eagerly when it needs more.  The thresholds change between
    segments so that every case runs: no recorded pass (first cycle), exactly as many as
    needed, more than needed (gated off) and fewer (continued).  Fields, potential, time,
    the last defect norm and V-cycle count equal the eager path bit for bit."""
    desc = grav_desc(precision, mg_bc, auto_origin, mg_threshold=1e-8, mg_niter=-1)
    st = turbulent_state(desc, seed=9, mach=1.5, drho=0.5)
    plan = ((2, True, 1e-8), (3, True, 1e-8), (1, False, 1e-8), (2, True, 1e-10), (2, True, 1e-6),
            (1, True, 1e-9))
    runs = []
    for graph in (False, True):
        m = agx.Mesh(desc)
        load_state(m, st)
        m.initialize()
        snaps = []
        for n, g, eps in plan:
            m.mg_set_mode(0, eps, -1)
            m.set_graph(g and graph)
            assert m.cycle(1e9, n) == n
            snap = {f: m.get(f) for f in FIELDS + ("phi",)}
            snap["time"] = m.time()
            snap["last"] = m.mg_last()
            snaps.append(snap)
            print(f"graph={graph} eps={eps:g}: last norm {snap['last'][0]:.6e}, {snap['last'][1]} V-cycles",
                  flush=True)
        runs.append(snaps)
        m.close()
    for k, (a, b) in enumerate(zip(*runs)):
        assert a["time"] == b["time"], k
        assert a["last"][1] == b["last"][1], (k, a["last"], b["last"])
        assert a["last"][0] == b["last"][0], (k, a["last"], b["last"])
        for f in FIELDS + ("phi",):
            same(b[f], a[f], f"segment {k}: {f} (precision {precision}, mg_bc {mg_bc}, auto origin {auto_origin})")
    assert len({s["last"][1] for s in runs[0]}) > 1   # the V-cycle count changed along the plan
