"""GPU (HIP, gfx950) cell-centred ghost exchange across refinement levels against the
CPU oracle, through the C ABI (agx_ref_copy_cc, agx_ref_restrict_cc, agx_ref_prolong_cc).

The host plan (athena-gamma_amd/smr.py: cc_plan) runs on the GPU as six batched
launches: same-level copies, restriction into the senders' coarse buffers, copies to
coarser blocks, copies into finer blocks' coarse buffers, same-level ghost restriction,
prolongation.  The oracle runs the prolongation step per block and coarser neighbour in
the reference's order (ProlongateBoundaries).  Bar: FP64 in the reference's operand
order, so u and the coarse buffers, every element, are identical bit for bit; a linear
field is reproduced exactly (NGHOST 2).
"""
import numpy as np
import pytest

from oracle import orc
from tests import ghost_cases as gc
from tests.conftest import load_smr

pytestmark = pytest.mark.gpu
NX = 8


def run_gpu(agx, t, u, cu, ng, ftab, ctab, nvar):
    keep = []
    plan = load_smr().cc_plan(t, NX, ng)
    nb = len(t.leaves)
    g = gc.geom(agx, NX, ng)
    fc, a1 = agx.device_coords(ftab)
    cc, a2 = agx.device_coords(ctab)
    keep += a1 + a2
    du, dcu = agx.DeviceArray(u), agx.DeviceArray(cu)
    L = lambda kind, items: agx.RefList(g, kind, [agx.RefItem(**d) for d in items], nb, nb)   # noqa: E731
    lists = [L("copy_cc_bb", plan.copy_same), L("restrict_cc", plan.restrict_send),
             L("copy_cc_cb", plan.copy_from_finer), L("copy_cc_bc", plan.copy_to_finer),
             L("restrict_cc", plan.restrict_same), L("prolong_cc", plan.prolong)]
    lists[0].copy_cc(nvar, du.ptr, du.ptr)
    lists[1].restrict_cc(nvar, du.ptr, fc, dcu.ptr)
    lists[2].copy_cc(nvar, dcu.ptr, du.ptr)
    lists[3].copy_cc(nvar, du.ptr, dcu.ptr)
    lists[4].restrict_cc(nvar, du.ptr, fc, dcu.ptr)
    lists[5].prolong_cc(nvar, dcu.ptr, cc, du.ptr, fc)
    agx.RefList.sync()
    out = du.get(), dcu.get()
    for lst in lists:
        lst.close()
    return out


@pytest.mark.parametrize("refine,ng", [(((1, 1, 1),), 2), (((1, 1, 1), (2, 1, 1)), 2), (((1, 1, 1),), 4)])
def test_ghost_exchange_matches_oracle(agx, refine, ng):
    t = gc.tree(refine)
    ftab, ctab = gc.coords(t, NX, ng), gc.coords(t, NX, ng, coarse=True)
    nvar = 5
    rng = np.random.default_rng(23)
    u = rng.standard_normal((len(t.leaves), nvar) + (NX + 2 * ng,) * 3)
    cu = rng.standard_normal((len(t.leaves), nvar) + (NX // 2 + 2 * ng,) * 3)
    print(f"refine={refine} ng={ng}: {len(t.leaves)} blocks, nvar {nvar}", flush=True)
    ru, rcu = gc.run_oracle(orc, t, u, cu, NX, ng, ftab, ctab, nvar, order="reference")
    gu, gcu = run_gpu(agx, t, u, cu, ng, ftab, ctab, nvar)
    print(f"  u: max|gpu - oracle| {np.abs(gu - ru).max():.3e}, coarse {np.abs(gcu - rcu).max():.3e}", flush=True)
    assert np.array_equal(gu, ru) and np.array_equal(gcu, rcu)


def test_ghost_exchange_linear_field_exact(agx):
    t = gc.tree(((1, 1, 1), (2, 1, 1)))
    ng = 2
    ftab, ctab = gc.coords(t, NX, ng), gc.coords(t, NX, ng, coarse=True)
    exact = gc.linear(t, ftab, 2)
    u = exact.copy()
    gc.poison_exchanged_ghosts(t, u, NX, ng)
    cu = np.full((len(t.leaves), 2) + (NX // 2 + 2 * ng,) * 3, np.nan)
    gu, _ = run_gpu(agx, t, u, cu, ng, ftab, ctab, 2)
    assert not np.isnan(gu).any()
    assert np.array_equal(gu, exact)
