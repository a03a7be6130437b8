"""GPU (HIP, gfx950) SMR ghost zones for MHD against the CPU oracle, through the C ABI:
exchange of u (cells) and b (faces) across levels, then ProlongateBoundaries per slot
of coarser neighbours (restriction of same-level ghosts, agx_ref_cons_to_prim on the
coarse buffer, prolongation of primitives and faces, agx_ref_prim_to_cons).  The oracle
runs the reference's per-block neighbour order.  States have 2% of cells below the
density or pressure floor.  Bar: FP64 in the reference's operand order, every array
identical bit for bit.
"""
import numpy as np
import pytest

from oracle import orc
from tests import ghost_cases as gc
from tests.conftest import load_smr

pytestmark = pytest.mark.gpu
NX, NG = 8, 2


def run_gpu(agx, t, A, ftab, ctab):
    smr = load_smr()
    cc, fc = smr.cc_plan(t, NX, NG), smr.fc_plan(t, NX, NG)
    slots, _ = smr.mhd_prolong_plan(t, NX, NG, cc, fc)
    nb = len(t.leaves)
    g = gc.geom(agx, NX, NG)
    keep = []
    fcd, a1 = agx.device_coords(ftab)
    ccd, a2 = agx.device_coords(ctab)
    keep += a1 + a2
    D = {k: ([agx.DeviceArray(x) for x in v] if isinstance(v, list) else agx.DeviceArray(v)) for k, v in A.items()}
    nbuf = len(fc.pack_b) + len(fc.pack_c)
    stride = max([1] + [(d["ei"] - d["si"] + 1) * (d["ej"] - d["sj"] + 1) * (d["ek"] - d["sk"] + 1)
                        for d in fc.pack_b + fc.pack_c])
    buf = agx.DeviceArray(np.zeros((max(1, nbuf), stride)))
    lists = []

    def L(kind, items, nbf=nb):
        x = agx.RefList(g, kind, [agx.RefItem(**d) for d in items], nbf, nb)
        lists.append(x)
        return x
    u, cu = D["u"].ptr, D["cu"].ptr
    b, cb = [x.ptr for x in D["b"]], [x.ptr for x in D["cb"]]
    L("copy_cc_bb", cc.copy_same).copy_cc(5, u, u)
    L("restrict_cc", cc.restrict_send).restrict_cc(5, u, fcd, cu)
    L("copy_cc_cb", cc.copy_from_finer).copy_cc(5, cu, u)
    L("copy_cc_bc", cc.copy_to_finer).copy_cc(5, u, cu)
    for c in range(3):
        L(f"restrict_fc{c + 1}", fc.restrict_send[c]).restrict_fc(b[c], fcd, cb[c])
    L("pack_fc_b", fc.pack_b).pack_fc(*b, buf.ptr, stride)
    L("pack_fc_c", fc.pack_c).pack_fc(*cb, buf.ptr + len(fc.pack_b) * stride * 8, stride)
    L("unpack_fc_b", fc.unpack_b, max(1, nbuf)).unpack_fc(*b, buf.ptr, stride)
    L("unpack_fc_c", fc.unpack_c, max(1, nbuf)).unpack_fc(*cb, buf.ptr, stride)
    w, bcc, cw, cbcc = D["w"].ptr, D["bcc"].ptr, D["cw"].ptr, D["cbcc"].ptr
    for st in slots:
        L("restrict_cc", st["restrict_cc"]).restrict_cc(5, u, fcd, cu)
        for c in range(3):
            L(f"restrict_fc{c + 1}", st["restrict_fc"][c]).restrict_fc(b[c], fcd, cb[c])
        L("cons_to_prim", st["c2p"]).cons_to_prim(cu, *cb, cw, cbcc, gc.GAMMA, gc.DFLOOR, gc.PFLOOR)
        L("prolong_cc", st["prolong_cc"]).prolong_cc(5, cw, ccd, w, fcd)
        for c in range(3):
            L(f"prolong_fc{c + 1}", st["prolong_fc"][c]).prolong_fc_shared(cb[c], ccd, b[c], fcd)
        L("prolong_fc_int", st["prolong_int"]).prolong_fc_internal(*b, fcd)
        L("prim_to_cons", st["p2c"]).prim_to_cons(w, *b, bcc, u, gc.GAMMA)
    agx.RefList.sync()
    out = {k: ([x.get() for x in v] if isinstance(v, list) else v.get()) for k, v in D.items()}
    for x in lists:
        x.close()
    return out


@pytest.mark.parametrize("refine", [((1, 1, 1),), ((1, 1, 1), (2, 1, 1))])
def test_mhd_ghosts_match_oracle(agx, refine):
    t = gc.tree(refine)
    ftab, ctab = gc.coords(t, NX, NG), gc.coords(t, NX, NG, coarse=True)
    A = gc.mhd_state(t, NX, NG, seed=6)
    print(f"refine={refine}: {len(t.leaves)} blocks", flush=True)
    ref = gc.run_oracle_mhd(orc, t, A, NX, NG, ftab, ctab, order="reference")
    got = run_gpu(agx, t, A, ftab, ctab)
    for k in ref:
        x, y = (ref[k], got[k]) if isinstance(ref[k], list) else ([ref[k]], [got[k]])
        for xx, yy in zip(x, y):
            print(f"  {k}: max|gpu - oracle| {np.abs(xx - yy).max():.3e}", flush=True)
            assert np.array_equal(xx, yy), k


@pytest.mark.parametrize("fused,levels", [(True, 2), (False, 2), (True, 3)])
def test_smr_boundary_driver(agx, fused, levels):
    """the package's SMRBoundary (every list built once from the tree) against the
    oracle: boundary_step on an MHD state with floors (fused: four launches per slot of
    the prolongation, or one per list), correct_emf on random EMFs"""
    from tests import emf_cases as ec
    smr = load_smr()
    t = gc.tree(((1, 1, 1), (2, 1, 1))) if levels == 2 else gc.tree(*gc.THREE)
    ftab, ctab = gc.coords(t, NX, NG), gc.coords(t, NX, NG, coarse=True)
    A = gc.mhd_state(t, NX, NG, seed=12)
    ref = gc.run_oracle_mhd(orc, t, A, NX, NG, ftab, ctab, order="reference")
    E = ec.random_emfs(len(t.leaves), NX, NG, seed=13)
    eref, _ = ec.run_oracle(orc, t, E, NX, NG, smr.block_coords(t, NX, NG))
    sb = agx.SMRBoundary(t, NX, NG, gc.GAMMA, gc.DFLOOR, gc.PFLOOR)
    D = {k: ([agx.DeviceArray(x) for x in v] if isinstance(v, list) else agx.DeviceArray(v)) for k, v in A.items()}
    de = [agx.DeviceArray(x) for x in E]
    P = {k: ([x.ptr for x in v] if isinstance(v, list) else v.ptr) for k, v in D.items()}
    sb.boundary_step(P["u"], P["b"], P["w"], P["bcc"], P["cu"], P["cb"], P["cw"], P["cbcc"], fused=fused)
    sb.correct_emf(*[x.ptr for x in de])
    sb.sync()
    for k in ref:
        x, y = (ref[k], D[k]) if isinstance(ref[k], list) else ([ref[k]], [D[k]])
        for xx, yy in zip(x, y):
            assert np.array_equal(xx, yy.get()), k
    for a in range(3):
        assert np.array_equal(eref[a], de[a].get()), f"e{a + 1}"
    sb.close()


def test_smr_boundary_graph_replay(agx):
    """the boundary step and EMF correction captured in a hipGraph: three replays, each
    from the same initial arrays, with eager work and synchronisations in between, each
    identical bit for bit to the eager run"""
    from tests import emf_cases as ec
    t = gc.tree(((1, 1, 1), (2, 1, 1)))
    A = gc.mhd_state(t, NX, NG, seed=14)
    E = ec.random_emfs(len(t.leaves), NX, NG, seed=15)
    sb = agx.SMRBoundary(t, NX, NG, gc.GAMMA, gc.DFLOOR, gc.PFLOOR)
    D = {k: ([agx.DeviceArray(x) for x in v] if isinstance(v, list) else agx.DeviceArray(v)) for k, v in A.items()}
    de = [agx.DeviceArray(x) for x in E]
    P = {k: ([x.ptr for x in v] if isinstance(v, list) else v.ptr) for k, v in D.items()}
    args = (P["u"], P["b"], P["w"], P["bcc"], P["cu"], P["cb"], P["cw"], P["cbcc"])
    emf = [x.ptr for x in de]

    def reset():
        for k, v in A.items():
            if isinstance(v, list):
                for d, h in zip(D[k], v):
                    d.set(h)
            else:
                D[k].set(v)
        for d, h in zip(de, E):
            d.set(h)

    def snapshot():
        out = {k: ([x.get() for x in v] if isinstance(v, list) else [v.get()]) for k, v in D.items()}
        out["emf"] = [x.get() for x in de]
        return out
    sb.boundary_step(*args)
    sb.correct_emf(*emf)
    sb.sync()
    eager = snapshot()
    reset()
    sb.capture(*args, emf=emf)
    other = [agx.DeviceArray(x) for x in ec.random_emfs(len(t.leaves), NX, NG, seed=16)]
    for rep in range(3):
        reset()
        sb.replay()
        sb.sync()
        got = snapshot()
        for k in eager:
            for xx, yy in zip(eager[k], got[k]):
                assert np.array_equal(xx, yy), (rep, k)
        print(f"replay {rep}: identical to the eager run", flush=True)
        sb.correct_emf(*[x.ptr for x in other])     # eager work between replays
        sb.sync()
    sb.close()
