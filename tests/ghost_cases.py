"""Meshes and drivers for the cell-centred ghost-exchange tests (host numpy).

A 4 x 4 x 4 root grid of nx^3-cell MeshBlocks on the unit cube with outflow boundaries;
root blocks listed in `refine` are refined once.  Blocks hold u (nb, nvar, n, n, n),
n = nx + 2 ng, and coarse buffers (nb, nvar, nc, nc, nc), nc = nx/2 + 2 ng.  Ghost cells
on the domain boundary keep what they hold (the physical boundary functions are not
part of the exchange).
"""
import itertools

import numpy as np

from tests.conftest import load_smr

NRB = (4, 4, 4)
COEF = [(1.0, 3.0, -2.0, 5.0), (-2.0, 1.0, 4.0, -3.0)]


def tree(refine, refine2=()):
    """root blocks in `refine` refined once; level-1 locations in `refine2` once more"""
    smr = load_smr()
    leaves = [(0, x, y, z) for z, y, x in itertools.product(*(range(n) for n in NRB[::-1]))
              if (x, y, z) not in refine]
    for (x, y, z) in refine:
        leaves += [(1, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)
                   if (2 * x + c1, 2 * y + c2, 2 * z + c3) not in refine2]
    for (x, y, z) in refine2:
        leaves += [(2, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)]
    return smr.SMRTree(NRB, leaves, periodic=(False, False, False))


# three levels: the central 2^3 root blocks refined, then one inner level-1 block again
THREE = (tuple((x, y, z) for z in (1, 2) for y in (1, 2) for x in (1, 2)), ((3, 3, 3),))


def coords(t, nx, ng, coarse=False):
    """per-block x?v, dx?f tables of the blocks, or of their coarse buffers"""
    n = (nx // 2 + 2 * ng) if coarse else nx + 2 * ng
    tab = {k: [] for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")}
    for (lev, *lx) in t.leaves:
        for d in range(3):
            h = 1.0 / ((NRB[d] * nx) << lev)
            if coarse:
                f = (lx[d] * nx + 2 * (np.arange(n + 1) - ng)) * h
            else:
                f = (lx[d] * nx - ng + np.arange(n + 1)) * h
            tab[f"x{d + 1}v"].append(0.5 * (f[1:] + f[:-1]))
            tab[f"dx{d + 1}f"].append(f[1:] - f[:-1])
    return {k: np.ascontiguousarray(np.array(v)) for k, v in tab.items()}


def linear(t, tab, nvar):
    nb, n = tab["x1v"].shape
    u = np.empty((nb, nvar, n, n, n))
    for v in range(nvar):
        c0, c1, c2, c3 = COEF[v]
        u[:, v] = (c0 + c1 * tab["x1v"][:, None, None, :] + c2 * tab["x2v"][:, None, :, None]
                   + c3 * tab["x3v"][:, :, None, None])
    return u


def poison_exchanged_ghosts(t, u, nx, ng):
    """NaN into every ghost region that has a neighbour (what the exchange must fill)"""
    lo, hi = ng, ng + nx - 1
    for g in range(len(t.leaves)):
        for ox in itertools.product((-1, 0, 1), repeat=3):
            if ox == (0, 0, 0) or t.find(g, ox) is None:
                continue
            sl = []
            for d in (2, 1, 0):
                o = ox[d]
                sl.append(slice(lo, hi + 1) if o == 0 else (slice(hi + 1, hi + 1 + ng) if o > 0 else slice(0, lo)))
            u[(g, slice(None), *sl)] = np.nan


def geom(mod, nx, ng):
    return mod.RefGeom(nx, nx, nx, ng, ng, 0)


def run_oracle(orc, t, u, cu, nx, ng, ftab, ctab, nvar, plan=None, order="reference"):
    """One exchange with prolongation on copies of (u, cu).  order 'reference': the
    prolongation step per block and coarser neighbour, restriction then prolongation, as
    ProlongateBoundaries; 'batched': every restriction, then every prolongation, as the
    library's lists run."""
    plan = plan or load_smr().cc_plan(t, nx, ng)
    g = geom(orc, nx, ng)
    u, cu = np.ascontiguousarray(u.copy()), np.ascontiguousarray(cu.copy())
    I = lambda ds: [orc.RefItem(**d) for d in ds]   # noqa: E731
    orc.ref_copy_cc(g, nvar, u, 0, u, 0, I(plan.copy_same))
    orc.ref_restrict_cc(g, nvar, u, ftab, cu, I(plan.restrict_send))
    orc.ref_copy_cc(g, nvar, cu, 1, u, 0, I(plan.copy_from_finer))
    orc.ref_copy_cc(g, nvar, u, 0, cu, 1, I(plan.copy_to_finer))
    uncovered = np.zeros(u.shape, bool)
    if order == "reference":
        for rs, p in plan.per_neighbor:
            orc.ref_restrict_cc(g, nvar, u, ftab, cu, I(rs))
            mark_uncovered(cu, p, ng, uncovered)
            orc.ref_prolong_cc(g, nvar, cu, ctab, u, ftab, I([p]))
    else:
        orc.ref_restrict_cc(g, nvar, u, ftab, cu, I(plan.restrict_same))
        orc.ref_prolong_cc(g, nvar, cu, ctab, u, ftab, I(plan.prolong))
    run_oracle.uncovered = uncovered
    return u, cu


def mark_uncovered(cu, p, ng, mask):
    """fine cells of prolongation item p whose stencil (the coarse cell and its six face
    neighbours, ProlongateCellCenteredValues) reads a coarse cell that is still NaN"""
    b = p["cb"]
    for ck in range(p["sk"], p["ek"] + 1):
        for cj in range(p["sj"], p["ej"] + 1):
            for ci in range(p["si"], p["ei"] + 1):
                st = [(ck, cj, ci)] + [(ck + a, cj + bb, ci + c) for a, bb, c in
                                       ((1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1))]
                if any(np.isnan(cu[(b, slice(None)) + s]).any() for s in st):
                    fk, fj, fi = [(c - ng) * 2 + ng for c in (ck, cj, ci)]
                    mask[b, :, fk:fk + 2, fj:fj + 2, fi:fi + 2] = True


# ---------------------------------------------------------------- face fields
BCOEF = [(1.0, 4.0, 2.0, -3.0), (-1.0, 1.0, -2.0, 5.0), (2.0, 3.0, -1.0, -2.0)]   # div B = 4 - 2 - 2 = 0


def face_coords(t, nx, ng, coarse=False):
    """x?f face positions per block (block faces or coarse-buffer faces)"""
    n = (nx // 2 + 2 * ng) if coarse else nx + 2 * ng
    out = []
    for d in range(3):
        rows = []
        for (lev, *lx) in t.leaves:
            hh = 1.0 / ((NRB[d] * nx) << lev)
            rows.append(((lx[d] * nx + 2 * (np.arange(n + 1) - ng)) if coarse else (lx[d] * nx - ng + np.arange(n + 1))) * hh)
        out.append(np.array(rows))
    return out


def linear_faces(t, nx, ng, tab, coarse=False):
    """b_c on its faces: a linear, divergence-free field (dyadic, so restriction,
    prolongation and the interior-face formula are exact)"""
    xf = face_coords(t, nx, ng, coarse)
    nb, n = tab["x1v"].shape
    out = []
    for c in range(3):
        pos = [xf[d] if d == c else tab[f"x{d + 1}v"] for d in range(3)]
        c0, c1, c2, c3 = BCOEF[c]
        out.append(np.ascontiguousarray(c0 + c1 * pos[0][:, None, None, :] + c2 * pos[1][:, None, :, None]
                                        + c3 * pos[2][:, :, None, None]))
    return out


def poison_face_ghosts(t, f, nx, ng):
    lo, hi = ng, ng + nx - 1
    for g in range(len(t.leaves)):
        for ox in itertools.product((-1, 0, 1), repeat=3):
            if ox == (0, 0, 0) or t.find(g, ox) is None:
                continue
            for c in range(3):
                sl = []
                for d in (2, 1, 0):
                    e = d == c
                    o = ox[d]
                    sl.append(slice(lo, hi + 1 + e) if o == 0 else
                              (slice(hi + 1 + e, hi + ng + 1 + e) if o > 0 else slice(0, lo)))
                f[c][(g, *sl)] = np.nan


def run_oracle_fc(orc, t, f, cf, nx, ng, ftab, ctab, plan=None, order="reference", coverage=None):
    """One face-field exchange with prolongation on copies of the block faces f[3] and
    coarse faces cf[3].  order 'reference': packs, unpacks in list (neighbour) order,
    then per block and coarser neighbour restriction, shared and interior prolongation;
    'batched': unpacks by launch group and the prolongation step list by list, as the
    library runs them.  coverage: a list that receives, per shared prolongation, whether
    a coarse face its minmod stencil reads was still NaN."""
    plan = plan or load_smr().fc_plan(t, nx, ng)
    g = geom(orc, nx, ng)
    f = [np.ascontiguousarray(a.copy()) for a in f]
    cf = [np.ascontiguousarray(a.copy()) for a in cf]
    I = lambda ds: [orc.RefItem(**d) for d in ds]   # noqa: E731
    for c in range(3):
        orc.ref_restrict_fc(g, c, f[c], ftab, cf[c], I(plan.restrict_send[c]))
    stride = max([1] + [(d["ei"] - d["si"] + 1) * (d["ej"] - d["sj"] + 1) * (d["ek"] - d["sk"] + 1)
                        for d in plan.pack_b + plan.pack_c])
    buf = np.zeros((max(1, len(plan.pack_b) + len(plan.pack_c)), stride))
    orc.ref_pack_fc(g, f, 0, buf, stride, I(plan.pack_b))
    bc = np.ascontiguousarray(buf[len(plan.pack_b):]) if plan.pack_c else np.zeros((1, stride))
    orc.ref_pack_fc(g, cf, 1, bc, stride, I(plan.pack_c))
    if plan.pack_c:
        buf[len(plan.pack_b):] = bc
    ub, uc = plan.unpack_b, plan.unpack_c
    if order == "batched":
        ub = sorted(ub, key=lambda d: d["aux"] >> 2)
        uc = sorted(uc, key=lambda d: d["aux"] >> 2)
    orc.ref_unpack_fc(g, f, 0, buf, stride, I(ub))
    orc.ref_unpack_fc(g, cf, 1, buf, stride, I(uc))
    if order == "reference":
        for rs, sh, it in plan.per_neighbor:
            for c in range(3):
                orc.ref_restrict_fc(g, c, f[c], ftab, cf[c], I(rs[c]))
            for c in range(3):
                if coverage is not None:
                    coverage.append(_fc_stencil_has_nan(cf[c], sh[c], c))
                orc.ref_prolong_fc_shared(g, c, cf[c], ctab, f[c], ftab, I([sh[c]]))
            orc.ref_prolong_fc_internal(g, f[0], f[1], f[2], ftab, I([it]))
    else:
        for c in range(3):
            orc.ref_restrict_fc(g, c, f[c], ftab, cf[c], I(plan.restrict_same[c]))
        for c in range(3):
            orc.ref_prolong_fc_shared(g, c, cf[c], ctab, f[c], ftab, I(plan.prolong_shared[c]))
        orc.ref_prolong_fc_internal(g, f[0], f[1], f[2], ftab, I(plan.prolong_int))
    return f, cf


def _fc_stencil_has_nan(cfc, p, c):
    """ProlongateSharedField reads each coarse face and its four neighbours along the two
    tangential directions (mesh_refinement.cpp:541-762), no diagonal ones"""
    a = cfc[p["cb"]]
    lo = [p["si"], p["sj"], p["sk"]]
    hi = [p["ei"], p["ej"], p["ek"]]
    boxes = [(0, 0, 0)]
    for d in range(3):
        if d != c:
            for s in (-1, 1):
                o = [0, 0, 0]
                o[d] = s
                boxes.append(tuple(o))
    for o in boxes:
        sub = a[lo[2] + o[2]:hi[2] + 1 + o[2], lo[1] + o[1]:hi[1] + 1 + o[1], lo[0] + o[0]:hi[0] + 1 + o[0]]
        if np.isnan(sub).any():
            return True
    return False


# ---------------------------------------------------------------- MHD pipeline
GAMMA, DFLOOR, PFLOOR = 5.0 / 3.0, 1e-4, 1e-3


def mhd_state(t, nx, ng, seed, floors=True):
    """random primitive state and faces; u from them; a few cells below the floors"""
    rng = np.random.default_rng(seed)
    nb = len(t.leaves)
    n, nc = nx + 2 * ng, nx // 2 + 2 * ng
    b = [rng.normal(0.0, 0.5, s) for s in ((nb, n, n, n + 1), (nb, n, n + 1, n), (nb, n + 1, n, n))]
    bcc = np.stack([0.5 * (b[0][..., :-1] + b[0][..., 1:]), 0.5 * (b[1][:, :, :-1] + b[1][:, :, 1:]),
                    0.5 * (b[2][:, :-1] + b[2][:, 1:])], axis=1)
    rho = rng.uniform(0.5, 1.5, (nb, n, n, n))
    v = rng.normal(0.0, 0.3, (nb, 3, n, n, n))
    p = rng.uniform(0.3, 1.0, (nb, n, n, n))
    u = np.empty((nb, 5, n, n, n))
    u[:, 0] = rho
    u[:, 1:4] = rho[:, None] * v
    u[:, 4] = p / (GAMMA - 1.0) + 0.5 * rho * (v ** 2).sum(1) + 0.5 * (bcc ** 2).sum(1)
    if floors:
        sel = rng.random((nb, n, n, n)) < 0.02
        u[:, 0][sel] = 1e-6
        sel = rng.random((nb, n, n, n)) < 0.02
        u[:, 4][sel] -= p[sel] / (GAMMA - 1.0) * 1.5
    arrs = dict(u=u, b=b, w=rng.standard_normal((nb, 5, n, n, n)), bcc=rng.standard_normal((nb, 3, n, n, n)),
                cu=rng.standard_normal((nb, 5, nc, nc, nc)), cw=rng.standard_normal((nb, 5, nc, nc, nc)),
                cbcc=rng.standard_normal((nb, 3, nc, nc, nc)),
                cb=[rng.standard_normal(s) for s in ((nb, nc, nc, nc + 1), (nb, nc, nc + 1, nc), (nb, nc + 1, nc, nc))])
    return {k: (np.ascontiguousarray(v) if not isinstance(v, list) else [np.ascontiguousarray(x) for x in v])
            for k, v in arrs.items()}


def run_oracle_mhd(orc, t, A, nx, ng, ftab, ctab, order="reference", gamma=GAMMA, dfloor=DFLOOR, pfloor=PFLOOR,
                   iso=False):
    """ghost exchange of u (cells) and b (faces) across levels, then ProlongateBoundaries
    for MHD, on copies of the arrays in A (u, w, cu, cw with NHYDRO = 4 variables for the
    isothermal EOS)"""
    smr = load_smr()
    cc, fc = smr.cc_plan(t, nx, ng), smr.fc_plan(t, nx, ng)
    slots, seq = smr.mhd_prolong_plan(t, nx, ng, cc, fc)
    A = {k: (v.copy() if not isinstance(v, list) else [x.copy() for x in v]) for k, v in A.items()}
    g = geom(orc, nx, ng)
    I = lambda ds: [orc.RefItem(**d) for d in ds]   # noqa: E731
    nh = 4 if iso else 5
    u, cu = A["u"], A["cu"]
    orc.ref_copy_cc(g, nh, u, 0, u, 0, I(cc.copy_same))
    orc.ref_restrict_cc(g, nh, u, ftab, cu, I(cc.restrict_send))
    orc.ref_copy_cc(g, nh, cu, 1, u, 0, I(cc.copy_from_finer))
    orc.ref_copy_cc(g, nh, u, 0, cu, 1, I(cc.copy_to_finer))
    A["b"], A["cb"] = run_oracle_fc(orc, t, A["b"], A["cb"], nx, ng, ftab, ctab, plan=_exchange_only(fc),
                                    order=order)
    for st in (seq if order == "reference" else slots):
        orc.ref_restrict_cc(g, nh, u, ftab, cu, I(st["restrict_cc"]))
        for c in range(3):
            orc.ref_restrict_fc(g, c, A["b"][c], ftab, A["cb"][c], I(st["restrict_fc"][c]))
        orc.ref_cons_to_prim(g, cu, A["cb"], A["cw"], A["cbcc"], gamma, dfloor, pfloor, int(iso), I(st["c2p"]))
        orc.ref_prolong_cc(g, nh, A["cw"], ctab, A["w"], ftab, I(st["prolong_cc"]))
        for c in range(3):
            orc.ref_prolong_fc_shared(g, c, A["cb"][c], ctab, A["b"][c], ftab, I(st["prolong_fc"][c]))
        orc.ref_prolong_fc_internal(g, A["b"][0], A["b"][1], A["b"][2], ftab, I(st["prolong_int"]))
        orc.ref_prim_to_cons(g, A["w"], A["b"][0], A["b"][1], A["b"][2], A["bcc"], u, gamma, int(iso), I(st["p2c"]))
    return A


def _exchange_only(fc):
    """the face plan without its prolongation step (which the MHD pipeline runs itself)"""
    import dataclasses
    return dataclasses.replace(fc, restrict_same=[[], [], []], prolong_shared=[[], [], []], prolong_int=[],
                               per_neighbor=[])
