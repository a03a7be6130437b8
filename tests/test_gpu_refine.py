"""GPU (HIP, gfx950) refinement operators against the CPU oracle, through the C ABI.

Every operator runs on random data over stretched per-block coordinates.  Its item
lists mix whole-block ranges, ghost-zone ranges and partial ranges on several blocks.
The same lists go to the oracle (oracle/agx_oracle_refine.c).  Bar: FP64 in the
reference's operand order on both sides, so every output array, including the elements
no item writes, is identical bit for bit.
"""
import numpy as np
import pytest

from oracle import orc
from tests.refine_cases import Geo, geom_struct, item

pytestmark = pytest.mark.gpu


class Dev:
    """Device copies of host arrays through the library's agx_device_* entry points."""

    def __init__(self):
        self.keep = []
        from tests.conftest import load_agx
        self.agx = load_agx()

    def put(self, a):
        t = self.agx.DeviceArray(a)
        self.keep.append(t)
        return t

    def coords(self, agx, tab):
        c, arrs = agx.device_coords(tab)
        self.keep += arrs
        return c


def same(gpu_t, ref, name):
    a = gpu_t.get()
    d = np.abs(a - ref).max() if a.size else 0.0
    print(f"{name}: max|gpu - oracle| = {d:.3e}, bitwise {np.array_equal(a, ref)}", flush=True)
    assert np.array_equal(a, ref), name


def lists(agx, g, kind, items, nbf, nbc):
    gg = geom_struct(agx.RefGeom, g)
    return agx.RefList(gg, kind, [item(agx.RefItem, *it[:8], **it[8]) if len(it) > 8 else item(agx.RefItem, *it)
                                  for it in items], nbf, nbc)


def oitems(items):
    return [item(orc.RefItem, *it[:8], **it[8]) if len(it) > 8 else item(orc.RefItem, *it) for it in items]


@pytest.mark.parametrize("ng", [2, 3])
def test_restrict_and_prolong_cc(agx, ng):
    g = Geo(nx=(8, 8, 8), ng=ng, nb=3, seed=11, stretch=0.25)
    nvar = 5
    rs = np.random.default_rng(12)
    fine = rs.normal(size=g.cells(nvar))
    coarse = rs.normal(size=g.ccells(nvar))
    cis, cie, cng = g.cis(0), g.cie(0), g.cng
    # restriction: whole interiors, and a range reaching into the coarse ghosts
    ritems = [(0, 0, cis, cie, cis, cie, cis, cie), (1, 1, cng - 1, cie, cis, cis + 1, cie, cie),
              (2, 2, cis + 1, cis + 2, cng - 1, cie + 1, cis, cie)]
    # prolongation: interiors and ghost-zone slabs (the coarse stencil needs +-1)
    pitems = [(0, 0, cis, cie, cis, cie, cis, cie), (1, 1, cis - 1, cis - 1, cis, cie, cis, cie),
              (1, 1, cie + 1, cie + 1, cis, cie, cis, cie), (2, 2, cis, cie, cis - 1, cie + 1, cis, cis + 1)]
    d = Dev()
    fc, cc = d.coords(agx, g.fine), d.coords(agx, g.coarse)
    fine_t, coarse_t = d.put(fine), d.put(coarse.copy())
    lr = lists(agx, g, "restrict_cc", ritems, g.nb, g.nb)
    lr.restrict_cc(nvar, fine_t.ptr, fc, coarse_t.ptr)
    agx.RefList.sync()
    ref = coarse.copy()
    orc.ref_restrict_cc(geom_struct(orc.RefGeom, g), nvar, fine, g.fine, ref, oitems(ritems))
    same(coarse_t, ref, f"restrict_cc ng={ng}")
    fine2_t = d.put(fine.copy())
    lp = lists(agx, g, "prolong_cc", pitems, g.nb, g.nb)
    lp.prolong_cc(nvar, coarse_t.ptr, cc, fine2_t.ptr, fc)
    agx.RefList.sync()
    ref2 = fine.copy()
    orc.ref_prolong_cc(geom_struct(orc.RefGeom, g), nvar, ref, g.coarse, ref2, g.fine, oitems(pitems))
    same(fine2_t, ref2, f"prolong_cc ng={ng}")


@pytest.mark.parametrize("d", [0, 1, 2])
def test_face_restrict_and_shared_prolong(agx, d):
    g = Geo(nx=(8, 8, 8), ng=2, nb=2, seed=13 + d, stretch=0.3)
    rs = np.random.default_rng(14 + d)
    fine = rs.normal(size=g.faces(d))
    coarse = rs.normal(size=g.faces(d, coarse=True))
    cis, cie = g.cis(0), g.cie(0)
    hi = [cie, cie, cie]
    hi[d] += 1
    ritems = [(0, 0, cis, hi[0], cis, hi[1], cis, hi[2]), (1, 1, cis, cis + 1, cis + 1, hi[1], cis, hi[2])]
    lo = [cis - 1, cis, cis]   # a ghost slab on the x1 side, faces along d
    pitems = [(0, 0, cis, hi[0], cis, hi[1], cis, hi[2]), (1, 1, lo[0], lo[0], lo[1], hi[1], lo[2], hi[2])]
    dv = Dev()
    fc, cc = dv.coords(agx, g.fine), dv.coords(agx, g.coarse)
    fine_t, coarse_t = dv.put(fine), dv.put(coarse.copy())
    G = geom_struct(orc.RefGeom, g)
    lists(agx, g, f"restrict_fc{d + 1}", ritems, g.nb, g.nb).restrict_fc(fine_t.ptr, fc, coarse_t.ptr)
    agx.RefList.sync()
    ref = coarse.copy()
    orc.ref_restrict_fc(G, d, fine, g.fine, ref, oitems(ritems))
    same(coarse_t, ref, f"restrict_fc dir {d}")
    fine2_t = dv.put(fine.copy())
    lists(agx, g, f"prolong_fc{d + 1}", pitems, g.nb, g.nb).prolong_fc_shared(coarse_t.ptr, cc,
                                                                             fine2_t.ptr, fc)
    agx.RefList.sync()
    ref2 = fine.copy()
    orc.ref_prolong_fc_shared(G, d, ref, g.coarse, ref2, g.fine, oitems(pitems))
    same(fine2_t, ref2, f"prolong_fc_shared dir {d}")


def test_internal_field(agx):
    g = Geo(nx=(8, 8, 8), ng=2, nb=2, seed=17, stretch=0.3)
    rs = np.random.default_rng(18)
    b = [rs.normal(size=g.faces(d)) for d in range(3)]
    cis, cie = g.cis(0), g.cie(0)
    items = [(0, 0, cis, cie, cis, cie, cis, cie), (1, 1, cis - 1, cis - 1, cis, cie, cis, cie + 1)]
    dv = Dev()
    fc = dv.coords(agx, g.fine)
    bt = [dv.put(x.copy()) for x in b]
    lists(agx, g, "prolong_fc_int", items, g.nb, g.nb).prolong_fc_internal(bt[0].ptr, bt[1].ptr,
                                                                           bt[2].ptr, fc)
    agx.RefList.sync()
    ref = [x.copy() for x in b]
    orc.ref_prolong_fc_internal(geom_struct(orc.RefGeom, g), ref[0], ref[1], ref[2], g.fine, oitems(items))
    for dd in range(3):
        same(bt[dd], ref[dd], f"internal x{dd + 1}f")


def test_flux_from_finer_and_emf_to_coarser(agx):
    g = Geo(nx=(8, 8, 8), ng=2, nb=4, seed=19, stretch=0.2)
    nvar = 5
    rs = np.random.default_rng(20)
    ff = [rs.normal(size=g.fluxes(d, nvar)) for d in range(3)]
    cf = [rs.normal(size=g.fluxes(d, nvar)) for d in range(3)]
    # finer blocks 0..3 onto faces / quadrants of coarser blocks 0..3 (pool of 4 as well)
    items = [(0, 1, 0, 0, 0, 0, 0, 0, dict(fid=1, fi1=1, fi2=0)), (1, 1, 0, 0, 0, 0, 0, 0, dict(fid=1, fi1=0, fi2=1)),
             (2, 0, 0, 0, 0, 0, 0, 0, dict(fid=2, fi1=1, fi2=1)), (3, 2, 0, 0, 0, 0, 0, 0, dict(fid=5, fi1=0, fi2=0)),
             (0, 3, 0, 0, 0, 0, 0, 0, dict(fid=0, fi1=1, fi2=1)), (1, 3, 0, 0, 0, 0, 0, 0, dict(fid=4, fi1=1, fi2=0))]
    dv = Dev()
    fc = dv.coords(agx, g.fine)
    fft = [dv.put(x) for x in ff]
    cft = [dv.put(x.copy()) for x in cf]
    lists(agx, g, "flux_from_finer", items, g.nb, g.nb).flux_from_finer(
        nvar, *[t.ptr for t in fft], fc, *[t.ptr for t in cft])
    agx.RefList.sync()
    ref = [x.copy() for x in cf]
    orc.ref_flux_from_finer(geom_struct(orc.RefGeom, g), nvar, ff[0], ff[1], ff[2], g.fine, ref[0], ref[1], ref[2],
                            oitems(items))
    for d in range(3):
        same(cft[d], ref[d], f"flux_from_finer x{d + 1}flux")
    e = [rs.normal(size=g.edges(d)) for d in range(3)]
    eitems = [(b, 0, 0, 0, 0, 0, 0, 0, dict(fid=f)) for b in range(g.nb) for f in range(6)]
    gg = geom_struct(agx.RefGeom, g)
    stride = max(agx.ref_emf_count(gg, f) for f in range(6)) + 3
    buf = np.full(stride * len(eitems), -7.0)
    et = [dv.put(x) for x in e]
    bt = dv.put(buf.copy())
    lists(agx, g, "emf_to_coarser", eitems, g.nb, g.nb).emf_to_coarser(*[t.ptr for t in et], fc,
                                                                        bt.ptr, stride)
    agx.RefList.sync()
    orc.ref_emf_to_coarser(geom_struct(orc.RefGeom, g), e[0], e[1], e[2], g.fine, buf, stride, oitems(eitems))
    same(bt, buf, "emf_to_coarser")
