"""GPU (HIP, gfx950) EMF correction across MeshBlocks against the CPU oracle, through
the C ABI (agx_ref_emf_same_level / _to_coarser / _set / _clear / _average).

Meshes: a periodic cube of 2^3 root blocks, one or two of them refined once (faces,
quadrants, edge-only contacts between levels), and the uniform mesh.  The work lists
come from the host module (athena-gamma_amd/smr.py) in the reference's neighbour order;
the oracle runs them item by item, the library per class and quadrant.  Bar: FP64 in the
reference's summation order, so every packed buffer and every edge array, including the
elements no item writes, is identical bit for bit.  The corrected fields also satisfy
the cross-block consistency the oracle is pinned on.
"""
import numpy as np
import pytest

from oracle import orc
from tests import emf_cases as ec
from tests.conftest import load_smr

pytestmark = pytest.mark.gpu
NG = 2


def run_gpu(agx, t, E, nx, ng, tab, plan):
    keep = []

    def dev(a):
        d = agx.DeviceArray(np.ascontiguousarray(a))
        keep.append(d)
        return d

    nb = len(t.leaves)
    g = agx.RefGeom(nx, nx, nx, ng, ng, 0)
    fc, arrs = agx.device_coords(tab)
    keep += arrs
    e = [dev(a) for a in E]
    ss = max(1, max((agx.ref_emf_count_same_level(g, d["fid"]) for d in plan.same_level), default=1))
    sc = max(1, max((agx.ref_emf_count(g, d["fid"]) for d in plan.to_coarser), default=1))
    bs = dev(np.zeros((max(1, len(plan.same_level)), ss)))
    bc = dev(np.zeros((max(1, len(plan.to_coarser)), sc)))
    IT = agx.RefItem
    L = lambda kind, items, nbf: agx.RefList(g, kind, [IT(**d) for d in items], nbf, nb)   # noqa: E731
    l_same = L("emf_same_level", plan.same_level, nb)
    l_coarse = L("emf_to_coarser", plan.to_coarser, nb)
    l_set_same = L("emf_set_same", plan.set_same, max(1, len(plan.same_level)))
    l_clear = L("emf_clear", plan.clear, nb)
    l_set_fine = L("emf_set_finer", plan.set_finer, max(1, len(plan.to_coarser)))
    l_avg = L("emf_average", plan.average, nb)
    p = [x.ptr for x in e]
    l_same.emf_same_level(*p, bs.ptr, ss)
    l_coarse.emf_to_coarser(*p, fc, bc.ptr, sc)
    l_set_same.emf_set(*p, bs.ptr, ss)
    l_clear.emf_clear(*p)
    l_set_fine.emf_set(*p, bc.ptr, sc)
    l_avg.emf_average(*p)
    agx.RefList.sync()
    out = [x.get() for x in e]
    bufs = (bs.get(), bc.get())
    for lst in (l_same, l_coarse, l_set_same, l_clear, l_set_fine, l_avg):
        lst.close()
    return out, bufs


@pytest.mark.parametrize("refine,nx,field", [
    (((0, 0, 0),), 8, "random"),
    (((0, 0, 0),), 8, "linear"),
    (((0, 0, 0), (1, 1, 1)), 8, "random"),   # two refined root blocks touching at a corner
    (((1, 0, 0),), 4, "random"),
    ((), 8, "random"),
])
def test_emf_correction_matches_oracle(agx, refine, nx, field):
    t = ec.tree(refine=refine)
    tab = ec.coords(t, nx, NG)
    plan = load_smr().emf_plan(t)
    nb = len(t.leaves)
    E = ec.random_emfs(nb, nx, NG, seed=17) if field == "random" else ec.linear_emfs(t, nx, NG, tab)
    print(f"mesh refine={refine} nx={nx} {field}: {nb} blocks, items same {len(plan.same_level)} "
          f"coarser {len(plan.to_coarser)} clear {len(plan.clear)} finer {len(plan.set_finer)} "
          f"average {len(plan.average)}", flush=True)
    ref, (obs, _, obc, _) = ec.run_oracle(orc, t, E, nx, NG, tab, plan=plan)
    got, (gbs, gbc) = run_gpu(agx, t, E, nx, NG, tab, plan)
    assert np.array_equal(gbs, obs), "same-level buffers"
    assert np.array_equal(gbc, obc), "to-coarser buffers"
    for a in range(3):
        d = np.abs(got[a] - ref[a]).max()
        print(f"  e{a + 1}: max|gpu - oracle| = {d:.3e}", flush=True)
        assert np.array_equal(got[a], ref[a]), f"e{a + 1}"
    spread, err, _ = ec.check_consistency(t, got, nx, NG)
    assert spread < 4e-15 and err < 4e-15
    if field == "linear":
        for a in range(3):
            assert np.array_equal(got[a], E[a])
