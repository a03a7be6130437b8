"""GPU (HIP, gfx950) face-field ghost exchange across refinement levels against the CPU
oracle, through the C ABI (agx_ref_pack_fc / agx_ref_unpack_fc, agx_ref_restrict_fc,
agx_ref_prolong_fc_shared, agx_ref_prolong_fc_internal).

The host plan (athena-gamma_amd/smr.py: fc_plan) runs on the GPU as whole-mesh launches:
restriction into the senders' coarse faces, packs, unpacks by launch group (neighbour
class and quadrant), same-level face restriction, shared and interior prolongation.  The
oracle runs the reference's neighbour order.  Bar: FP64 in the reference's operand
order, so every block face and coarse face is identical bit for bit; a linear
divergence-free field is reproduced exactly.
"""
import numpy as np
import pytest

from oracle import orc
from tests import ghost_cases as gc
from tests.conftest import load_smr

pytestmark = pytest.mark.gpu
NX, NG = 8, 2


def run_gpu(agx, t, f, cf, ftab, ctab):
    keep = []
    plan = load_smr().fc_plan(t, NX, NG)
    nb = len(t.leaves)
    g = gc.geom(agx, NX, NG)
    fc, a1 = agx.device_coords(ftab)
    cc, a2 = agx.device_coords(ctab)
    keep += a1 + a2
    df = [agx.DeviceArray(a) for a in f]
    dcf = [agx.DeviceArray(a) for a in cf]
    nbuf = len(plan.pack_b) + len(plan.pack_c)
    stride = max([1] + [(d["ei"] - d["si"] + 1) * (d["ej"] - d["sj"] + 1) * (d["ek"] - d["sk"] + 1)
                        for d in plan.pack_b + plan.pack_c])
    buf = agx.DeviceArray(np.zeros((max(1, nbuf), stride)))
    keep += df + dcf + [buf]
    L = lambda kind, items, nbf=nb: agx.RefList(g, kind, [agx.RefItem(**d) for d in items], nbf, nb)   # noqa: E731
    rsend = [L(f"restrict_fc{c + 1}", plan.restrict_send[c]) for c in range(3)]
    pb, pc = L("pack_fc_b", plan.pack_b), L("pack_fc_c", plan.pack_c)
    ub, uc = L("unpack_fc_b", plan.unpack_b, max(1, nbuf)), L("unpack_fc_c", plan.unpack_c, max(1, nbuf))
    rsame = [L(f"restrict_fc{c + 1}", plan.restrict_same[c]) for c in range(3)]
    sh = [L(f"prolong_fc{c + 1}", plan.prolong_shared[c]) for c in range(3)]
    pi = L("prolong_fc_int", plan.prolong_int)
    F = [x.ptr for x in df]
    CF = [x.ptr for x in dcf]
    for c in range(3):
        rsend[c].restrict_fc(F[c], fc, CF[c])
    pb.pack_fc(*F, buf.ptr, stride)
    pc.pack_fc(*CF, buf.ptr + len(plan.pack_b) * stride * 8, stride)
    ub.unpack_fc(*F, buf.ptr, stride)
    uc.unpack_fc(*CF, buf.ptr, stride)
    for c in range(3):
        rsame[c].restrict_fc(F[c], fc, CF[c])
    for c in range(3):
        sh[c].prolong_fc_shared(CF[c], cc, F[c], fc)
    pi.prolong_fc_internal(*F, fc)
    agx.RefList.sync()
    out = [x.get() for x in df], [x.get() for x in dcf]
    for x in rsend + [pb, pc, ub, uc] + rsame + sh + [pi]:
        x.close()
    return out


@pytest.mark.parametrize("refine", [((1, 1, 1),), ((1, 1, 1), (2, 1, 1)), ()])
def test_face_exchange_matches_oracle(agx, refine):
    t = gc.tree(refine)
    ftab, ctab = gc.coords(t, NX, NG), gc.coords(t, NX, NG, coarse=True)
    rng = np.random.default_rng(29)
    f = [rng.standard_normal(a.shape) for a in gc.linear_faces(t, NX, NG, ftab)]
    cf = [rng.standard_normal(a.shape) for a in gc.linear_faces(t, NX, NG, ctab, coarse=True)]
    print(f"refine={refine}: {len(t.leaves)} blocks", flush=True)
    rf, rcf = gc.run_oracle_fc(orc, t, f, cf, NX, NG, ftab, ctab, order="reference")
    gf, gcf = run_gpu(agx, t, f, cf, ftab, ctab)
    for c in range(3):
        print(f"  x{c + 1}f: max|gpu - oracle| {np.abs(gf[c] - rf[c]).max():.3e}, coarse "
              f"{np.abs(gcf[c] - rcf[c]).max():.3e}", flush=True)
        assert np.array_equal(gf[c], rf[c]) and np.array_equal(gcf[c], rcf[c]), c


def test_face_exchange_linear_field_exact(agx):
    t = gc.tree(((1, 1, 1), (2, 1, 1)))
    ftab, ctab = gc.coords(t, NX, NG), gc.coords(t, NX, NG, coarse=True)
    exact = gc.linear_faces(t, NX, NG, ftab)
    f = [a.copy() for a in exact]
    gc.poison_face_ghosts(t, f, NX, NG)
    cf = [np.full(a.shape, np.nan) for a in gc.linear_faces(t, NX, NG, ctab, coarse=True)]
    gf, _ = run_gpu(agx, t, f, cf, ftab, ctab)
    for c in range(3):
        assert not np.isnan(gf[c]).any() and np.array_equal(gf[c], exact[c]), c
