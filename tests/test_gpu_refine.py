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


@pytest.mark.parametrize("case", range(6))
def test_ghost_prolongation_pipeline(agx, case):
    """ProlongateGhostCells for one coarser neighbour, with ranges from
    agx_ref_prolong_ranges: cell prolongation of w, shared faces over the face ranges,
    interior faces, then bcc and u on the prolonged cells; bitwise against the oracle."""
    import itertools
    rng = np.random.default_rng(30 + case)
    oxs = [ox for ox in itertools.product((-1, 0, 1), repeat=3) if ox != (0, 0, 0)]
    ox = oxs[rng.integers(len(oxs))]
    par = tuple(int(v) for v in rng.integers(0, 2, 3))
    lv = [-1] * 27 if case == 0 else list(int(v) for v in rng.integers(-1, 2, 27))
    lv[13] = 0
    lv[((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1] = -1
    g = Geo(nx=(8, 8, 8), ng=2, nb=1, seed=31 + case, stretch=0.2)
    cw = rng.normal(size=g.ccells(5))
    cw[:, 0] = np.abs(cw[:, 0]) + 0.5
    cb = [rng.normal(size=g.faces(d, coarse=True)) for d in range(3)]
    fw = rng.normal(size=g.cells(5))
    fb = [rng.normal(size=g.faces(d)) for d in range(3)]
    bcc0, u0 = rng.normal(size=g.cells(3)), rng.normal(size=g.cells(5))
    gamma = 5.0 / 3.0
    ga = geom_struct(agx.RefGeom, g)
    r = agx.ref_prolong_ranges(ga, ox, par, lv)
    print(f"ox {ox} parity {par}: ranges {r}", flush=True)
    cc = (0, 0, r[0], r[1], r[2], r[3], r[4], r[5])
    fr = [(0, 0, r[6], r[7], r[2], r[3], r[4], r[5]), (0, 0, r[0], r[1], r[8], r[9], r[4], r[5]),
          (0, 0, r[0], r[1], r[2], r[3], r[10], r[11])]
    dv = Dev()
    fcd, ccd = dv.coords(agx, g.fine), dv.coords(agx, g.coarse)
    cw_t, fw_t = dv.put(cw), dv.put(fw)
    cb_t, fb_t = [dv.put(x) for x in cb], [dv.put(x) for x in fb]
    bcc_t, u_t = dv.put(bcc0), dv.put(u0)
    lists(agx, g, "prolong_cc", [cc], 1, 1).prolong_cc(5, cw_t.ptr, ccd, fw_t.ptr, fcd)
    for d in range(3):
        lists(agx, g, f"prolong_fc{d + 1}", [fr[d]], 1, 1).prolong_fc_shared(cb_t[d].ptr, ccd, fb_t[d].ptr, fcd)
    lists(agx, g, "prolong_fc_int", [cc], 1, 1).prolong_fc_internal(fb_t[0].ptr, fb_t[1].ptr, fb_t[2].ptr, fcd)
    lists(agx, g, "prim_to_cons", [cc], 1, 1).prim_to_cons(fw_t.ptr, fb_t[0].ptr, fb_t[1].ptr, fb_t[2].ptr,
                                                           bcc_t.ptr, u_t.ptr, gamma)
    agx.RefList.sync()
    G = geom_struct(orc.RefGeom, g)
    assert orc.ref_prolong_ranges(G, ox, par, lv) == r
    rw, rb, rbcc, ru = fw.copy(), [x.copy() for x in fb], bcc0.copy(), u0.copy()
    orc.ref_prolong_cc(G, 5, cw, g.coarse, rw, g.fine, oitems([cc]))
    for d in range(3):
        orc.ref_prolong_fc_shared(G, d, cb[d], g.coarse, rb[d], g.fine, oitems([fr[d]]))
    orc.ref_prolong_fc_internal(G, rb[0], rb[1], rb[2], g.fine, oitems([cc]))
    orc.ref_prim_to_cons(G, rw, rb[0], rb[1], rb[2], rbcc, ru, gamma, 0, oitems([cc]))
    same(fw_t, rw, "w")
    for d in range(3):
        same(fb_t[d], rb[d], f"x{d + 1}f")
    same(bcc_t, rbcc, "bcc")
    same(u_t, ru, "u")
