"""Meshes, EMF arrays and cross-block checks for the EMF-correction tests (host numpy).

A mesh is a periodic unit cube of 2 x 2 x 2 root MeshBlocks of nx^3 cells, some of them
refined once into 8 children (athena-gamma_amd/smr.py builds the tree, the neighbours and
the work lists).  Every block keeps its own edge arrays in the reference's layouts:
x1e (nb, n+1, n+1, n), x2e (nb, n+1, n, n+1), x3e (nb, n, n+1, n+1), n = nx + 2 ng.
"""
import itertools

import numpy as np

from tests.conftest import load_smr

NRB = (2, 2, 2)


def tree(refine=((0, 0, 0),)):
    smr = load_smr()
    leaves = [(0, x, y, z) for z, y, x in itertools.product(range(NRB[2]), range(NRB[1]), range(NRB[0]))
              if (x, y, z) not in refine]
    for (x, y, z) in refine:
        leaves += [(1, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)]
    return smr.SMRTree(NRB, leaves)


def coords(t, nx, ng):
    """per-block x?v, dx?f, x?f tables (uniform spacing per level)"""
    n = nx + 2 * ng
    tab = {k: [] for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f", "x1f", "x2f", "x3f")}
    for (lev, *lx) in t.leaves:
        for d in range(3):
            h = 1.0 / ((NRB[d] * nx) << lev)
            f = (lx[d] * nx - ng + np.arange(n + 1)) * h
            tab[f"x{d + 1}f"].append(f)
            tab[f"x{d + 1}v"].append(0.5 * (f[1:] + f[:-1]))
            tab[f"dx{d + 1}f"].append(np.full(n, h))
    return {k: np.ascontiguousarray(np.array(v)) for k, v in tab.items()}


def shapes(nb, nx, ng):
    n = nx + 2 * ng
    return [(nb, n + 1, n + 1, n), (nb, n + 1, n, n + 1), (nb, n, n + 1, n + 1)]


def random_emfs(nb, nx, ng, seed):
    rng = np.random.default_rng(seed)
    return [rng.standard_normal(s) for s in shapes(nb, nx, ng)]


def linear_emfs(t, nx, ng, tab):
    """e_a = c_a . (x, y, z) + c_a0 at each edge's midpoint, transverse face coordinates
    taken mod 1 (periodic).  Dyadic positions and small integer coefficients: every
    restriction and average of these values is exact."""
    nb = len(t.leaves)
    out = []
    coef = [(3, 1, -2, 5), (-1, 4, 2, -3), (2, -5, 1, 4)]
    for a, s in enumerate(shapes(nb, nx, ng)):
        e = np.empty(s)
        for b in range(nb):
            pos = []
            for d in range(3):
                if d == a:
                    pos.append(tab[f"x{d + 1}v"][b])
                else:
                    pos.append(np.mod(tab[f"x{d + 1}f"][b], 1.0))
            x, y, z = pos
            c0, c1, c2, c3 = coef[a]
            e[b] = c0 + c1 * x[None, None, :] + c2 * y[None, :, None] + c3 * z[:, None, None]
        out.append(e)
    return out


def edge_table(t, E, nx, ng):
    """{(level, arr, gx, gy, gz): [values]} over every edge of every block's active
    region: start points in the level's global index space, wrapped in periodic
    directions."""
    tab = {}
    lo, hi = ng, ng + nx
    for b, (lev, *lx) in enumerate(t.leaves):
        N = [(t.nrb[d] * nx) << lev for d in range(3)]
        for a in range(3):
            rng = [range(lo, hi + 1 if d != a else hi) for d in range(3)]   # i, j, k ranges
            wrap = [N[d] if t.periodic[d] else 1 << 30 for d in range(3)]   # positions, periodic wrap only
            for k in rng[2]:
                gz = (lx[2] * nx + k - ng) % wrap[2]
                for j in rng[1]:
                    gy = (lx[1] * nx + j - ng) % wrap[1]
                    row = E[a][b, k, j]
                    for i in rng[0]:
                        gx = (lx[0] * nx + i - ng) % wrap[0]
                        tab.setdefault((lev, a, gx, gy, gz), []).append(float(row[i]))
    return tab


def check_consistency(t, E, nx, ng, tol=4e-15):
    """(same-level spread, coarse-vs-restricted-fine error, coarse edges compared):
    every edge shared by blocks of one level carries one value in all of them, and every
    edge of one level that edges of the next finer level cover equals their length-
    weighted mean (equal lengths)."""
    tab = edge_table(t, E, nx, ng)
    spread = max((max(v) - min(v) for v in tab.values() if len(v) > 1), default=0.0)
    err, ncmp = 0.0, 0
    for (lev, a, gx, gy, gz), v in tab.items():
        p = [2 * gx, 2 * gy, 2 * gz]
        q = list(p)
        q[a] += 1
        fa, fb = tab.get((lev + 1, a, *p)), tab.get((lev + 1, a, *q))
        if fa is None or fb is None:
            continue
        ncmp += 1
        err = max(err, abs(v[0] - (fa[0] + fb[0]) * 0.5))
    return spread, err, ncmp


def items_of(cls, dicts):
    return [cls(**d) for d in dicts]


def run_oracle(orc, t, E, nx, ng, tab, skip=(), plan=None):
    """One EMF correction on copies of E with the oracle, in the reference's order;
    `skip` leaves out steps ('clear', 'average') to plant a mistake, or all of them
    ('all')."""
    smr = load_smr()
    plan = plan or smr.emf_plan(t)
    g = orc.RefGeom(nx, nx, nx, ng, ng, 0)
    e1, e2, e3 = [np.ascontiguousarray(a.copy()) for a in E]
    if "all" in skip:
        return [e1, e2, e3], None
    ss = max(1, max((orc_count_same(nx, d["fid"]) for d in plan.same_level), default=1))
    sc = max(1, max((orc_count_coarser(nx, d["fid"]) for d in plan.to_coarser), default=1))
    buf_s = np.zeros((max(1, len(plan.same_level)), ss))
    buf_c = np.zeros((max(1, len(plan.to_coarser)), sc))
    IT = orc.RefItem
    orc.ref_emf_same_level(g, e1, e2, e3, buf_s, ss, items_of(IT, plan.same_level))
    orc.ref_emf_to_coarser(g, e1, e2, e3, tab, buf_c, sc, items_of(IT, plan.to_coarser))
    orc.ref_emf_set(g, e1, e2, e3, buf_s, ss, 0, items_of(IT, plan.set_same))
    if "clear" not in skip:
        orc.ref_emf_clear(g, e1, e2, e3, items_of(IT, plan.clear))
    orc.ref_emf_set(g, e1, e2, e3, buf_c, sc, 1, items_of(IT, plan.set_finer))
    if "average" not in skip:
        orc.ref_emf_average(g, e1, e2, e3, items_of(IT, plan.average))
    return [e1, e2, e3], (buf_s, ss, buf_c, sc)


def orc_count_same(nx, fid):
    """values of a same-level buffer (cubic blocks): a face packs two edge components
    of (nx+1) x nx, an edge nx values"""
    return 2 * (nx + 1) * nx if fid < 6 else nx


def orc_count_coarser(nx, fid):
    h = nx // 2
    return 2 * (h + 1) * h if fid < 6 else h
