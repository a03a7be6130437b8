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


def tree(refine):
    smr = load_smr()
    leaves = [(0, x, y, z) for z, y, x in itertools.product(*(range(n) for n in NRB[::-1]))
              if (x, y, z) not in refine]
    for (x, y, z) in refine:
        leaves += [(1, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)]
    return smr.SMRTree(NRB, leaves, periodic=(False, False, False))


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
