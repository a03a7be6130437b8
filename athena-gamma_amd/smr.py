"""Static-refinement block tree, neighbour search and EMF-correction work lists (host).

A mesh of 3D MeshBlocks on several refinement levels is given as its leaves, each a
LogicalLocation (level, lx1, lx2, lx3) with level 0 = the root grid of nrb blocks.  This
module finds every block's face, edge and corner neighbours the way the reference does:
  - MeshBlockTree::FindNeighbor (mesh/meshblock_tree.cpp:362-450): the same-level
    location at the offset, wrapped on periodic boundaries, is a leaf (same level),
    lies inside a coarser leaf, or is refined (its children are the neighbours);
  - BoundaryBase::SearchAndSetNeighbors (bvals/bvals_base.cpp:299-700): the neighbour
    order (x1, x2, x3 faces, then x1x2, x1x3, x2x3 edges), the finer children of a
    face (f2 outer, f1 inner) or an edge, the face id fid and edge id eid
    (NeighborBlock::SetNeighbor, bvals_base.cpp:43-66), the rule that a coarser edge
    neighbour is kept only by the block at that corner of its parent, and nblevel;
  - FaceCenteredBoundaryVariable::CountFineEdges (bvals/fc/bvals_fc.cpp:1128-1191).
It then builds the work lists of
  - the EMF correction of one cycle in the order of SendFluxCorrection /
    ReceiveFluxCorrection (bvals/fc/flux_correction_fc.cpp:623-688, 1610-1660): pack
    every same-level and to-coarser buffer from the blocks' own EMFs, then per block add
    the same-level buffers, clear the parts bordering finer blocks, add the finer
    buffers and average (faces and edges);
  - the ghost exchange of cell-centred variables across levels
    (CellCenteredBoundaryVariable, bvals/cc/bvals_cc.cpp:201-479, and the cell part of
    BoundaryValues::ProlongateBoundaries, bvals/bvals_refine.cpp:93-276): same-level
    copies, restriction and copy to coarser blocks, copy of coarser data into the coarse
    buffer, restriction of same-level ghosts and prolongation (faces, edges, corners).

Plain host logic: no device work.  Polar and shearing boundaries are not handled.
"""
from dataclasses import dataclass

import numpy as np

# AGX_REF_* kinds of the EMF correction (include/athena_gamma_amd.h)
K_EMF_TO_COARSER, K_EMF_SAME_LEVEL, K_EMF_SET_SAME, K_EMF_SET_FINER, K_EMF_CLEAR, K_EMF_AVERAGE = \
    10, 12, 13, 14, 15, 16


@dataclass
class Neighbor:
    gid: int
    rel: int          # neighbour level minus this block's level: -1, 0, +1
    kind: str         # "face", "edge" or "corner"
    ox: tuple         # (ox1, ox2, ox3)
    fid: int = -1     # BoundaryFace for faces
    eid: int = -1     # edge id for edges
    fi1: int = 0
    fi2: int = 0

    @property
    def code(self):
        """fid of a work item: the face 0..5 or 6 + the edge id."""
        return self.fid if self.kind == "face" else 6 + self.eid


def face_id(ox):
    for d in range(3):
        if ox[d]:
            return 2 * d + (ox[d] > 0)
    raise ValueError(ox)


def edge_id(ox):
    o1, o2, o3 = ox
    if o3 == 0:
        return ((o1 + 1) >> 1) | ((o2 + 1) & 2)
    if o2 == 0:
        return 4 + (((o1 + 1) >> 1) | ((o3 + 1) & 2))
    return 8 + (((o2 + 1) >> 1) | ((o3 + 1) & 2))


class SMRTree:
    def __init__(self, nrb, leaves, periodic=(True, True, True)):
        self.nrb = tuple(int(x) for x in nrb)
        self.leaves = [tuple(int(v) for v in loc) for loc in leaves]
        self.periodic = tuple(bool(p) for p in periodic)
        self.gid = {loc: g for g, loc in enumerate(self.leaves)}
        if len(self.gid) != len(self.leaves):
            raise ValueError("duplicate leaves")
        self.maxlev = max(loc[0] for loc in self.leaves)
        vol = sum(8 ** (self.maxlev - loc[0]) for loc in self.leaves)
        if vol != self.nrb[0] * self.nrb[1] * self.nrb[2] * 8 ** self.maxlev:
            raise ValueError("leaves do not tile the root grid")
        for loc in self.leaves:
            if self._covering(loc) != loc:
                raise ValueError(f"leaf {loc} overlaps another leaf")
        self._nb = [self._search(g) for g in range(len(self.leaves))]
        for g in range(len(self.leaves)):
            for lev in self.nblevel(g):
                if lev > 1:
                    raise ValueError("neighbouring blocks differ by more than one level")

    # ---- tree queries
    def _covering(self, loc):
        """the leaf containing the level-`loc[0]` location, at its level or coarser"""
        lev, x, y, z = loc
        while lev >= 0:
            if (lev, x, y, z) in self.gid:
                return (lev, x, y, z)
            lev, x, y, z = lev - 1, x >> 1, y >> 1, z >> 1
        return None

    def _wrap(self, loc, ox):
        lev = loc[0]
        pos = []
        for d in range(3):
            n = self.nrb[d] << lev
            p = loc[1 + d] + ox[d]
            if p < 0 or p >= n:
                if not self.periodic[d]:
                    return None
                p %= n
            pos.append(p)
        return (lev, *pos)

    def find(self, g, ox):
        """FindNeighbor: ('same'|'coarser', gid) or ('finer', location at this level) or None"""
        loc = self.leaves[g]
        t = self._wrap(loc, ox)
        if t is None:
            return None
        if t in self.gid:
            return ("same", self.gid[t])
        c = self._covering(t)
        if c is not None:
            return ("coarser", self.gid[c])
        return ("finer", t)

    def _child(self, t, c1, c2, c3):
        lev, x, y, z = t
        return self.gid[(lev + 1, 2 * x + c1, 2 * y + c2, 2 * z + c3)]

    # ---- SearchAndSetNeighbors (faces and edges)
    def _search(self, g):
        loc = self.leaves[g]
        myox = [((loc[1 + d] & 1) * 2 - 1) for d in range(3)]
        out = []
        for d in range(3):                       # faces
            for n in (-1, 1):
                ox = [0, 0, 0]
                ox[d] = n
                r = self.find(g, ox)
                if r is None:
                    continue
                fid = face_id(ox)
                if r[0] == "finer":
                    ff = 1 - (n + 1) // 2
                    for f2 in range(2):
                        for f1 in range(2):   # GetLeaf: f1, f2 along the face's first and second tangent
                            if d == 0:
                                idx = (ff, f1, f2)
                            elif d == 1:
                                idx = (f1, ff, f2)
                            else:
                                idx = (f1, f2, ff)
                            out.append(Neighbor(self._child(r[1], *idx), 1, "face", tuple(ox), fid=fid, fi1=f1, fi2=f2))
                else:
                    out.append(Neighbor(r[1], 0 if r[0] == "same" else -1, "face", tuple(ox), fid=fid))
        for da, db in ((0, 1), (0, 2), (1, 2)):   # edges: x1x2, x1x3, x2x3
            for m in (-1, 1):
                for n in (-1, 1):
                    ox = [0, 0, 0]
                    ox[da], ox[db] = n, m
                    r = self.find(g, ox)
                    if r is None:
                        continue
                    eid = edge_id(ox)
                    if r[0] == "finer":
                        ffa, ffb = 1 - (n + 1) // 2, 1 - (m + 1) // 2
                        free = 3 - da - db
                        for f1 in range(2):
                            idx = [0, 0, 0]
                            idx[da], idx[db], idx[free] = ffa, ffb, f1
                            out.append(Neighbor(self._child(r[1], *idx), 1, "edge", tuple(ox), eid=eid, fi1=f1))
                    else:
                        rel = 0 if r[0] == "same" else -1
                        if rel == 0 or (myox[da] == n and myox[db] == m):
                            out.append(Neighbor(r[1], rel, "edge", tuple(ox), eid=eid))
        for l in (-1, 1):                         # corners
            for m in (-1, 1):
                for n in (-1, 1):
                    ox = (n, m, l)
                    r = self.find(g, ox)
                    if r is None:
                        continue
                    if r[0] == "finer":
                        out.append(Neighbor(self._child(r[1], 1 - (n + 1) // 2, 1 - (m + 1) // 2, 1 - (l + 1) // 2),
                                            1, "corner", ox))
                    elif r[0] == "same" or myox == [n, m, l]:
                        out.append(Neighbor(r[1], 0 if r[0] == "same" else -1, "corner", ox))
        return out

    def neighbors(self, g, corners=True):
        return self._nb[g] if corners else [n for n in self._nb[g] if n.kind != "corner"]

    def nblevel(self, g):
        """27 levels relative to block g at ((ox3+1)*3 + ox2+1)*3 + ox1+1; -1 where the
        neighbour is coarser or missing (as the reference's -1 for a missing one)."""
        lv = [-1] * 27
        lv[13] = 0
        for o3 in (-1, 0, 1):
            for o2 in (-1, 0, 1):
                for o1 in (-1, 0, 1):
                    if (o1, o2, o3) == (0, 0, 0):
                        continue
                    r = self.find(g, (o1, o2, o3))
                    if r is None:
                        continue
                    lv[((o3 + 1) * 3 + o2 + 1) * 3 + o1 + 1] = {"same": 0, "coarser": -1, "finer": 1}[r[0]]
        return lv

    def boundary_is_block(self, g, fid):
        """the face has a block or periodic boundary (AverageFluxBoundary's test)"""
        ox = [0, 0, 0]
        ox[fid >> 1] = 1 if fid & 1 else -1
        return self._wrap(self.leaves[g], ox) is not None


def count_fine_edges(nblevel):
    """CountFineEdges on 27 relative levels: (edge_flag[12], nedge_fine[12])."""
    def lv(o3, o2, o1):
        return nblevel[((o3 + 1) * 3 + o2 + 1) * 3 + o1 + 1]
    flags, counts = [], []
    for plane in range(3):
        for os_ in (-1, 1):
            for of in (-1, 1):
                nf, fl = 0, 0
                for a in range(max(os_ - 1, -1), min(os_ + 1, 1) + 1):
                    for c in range(max(of - 1, -1), min(of + 1, 1) + 1):
                        lev = lv(0, a, c) if plane == 0 else (lv(a, 0, c) if plane == 1 else lv(a, c, 0))
                        if lev > fl:
                            fl, nf = fl + 1, 0
                        if lev == fl:
                            nf += 1
                flags.append(fl == 0)
                counts.append(nf)
    return flags, counts


def _item(fb=0, cb=0, fid=0, fi1=0, fi2=0):
    return dict(fb=fb, cb=cb, si=0, ei=0, sj=0, ej=0, sk=0, ek=0, fid=fid, fi1=fi1, fi2=fi2, aux=0)


@dataclass
class EMFPlan:
    """Work lists (dicts of agx_ref_item fields) of one EMF correction, in the reference's
    order.  to_coarser and same_level pack slots 0.. of their own buffers; set_same reads
    slots of the same-level buffer, set_finer slots of the to-coarser buffer."""
    to_coarser: list
    same_level: list
    set_same: list
    clear: list
    set_finer: list
    average: list


def emf_plan(tree):
    nblk = len(tree.leaves)
    flags = [count_fine_edges(tree.nblevel(g)) for g in range(nblk)]
    to_coarser, same_level = [], []
    slot_same, slot_fine = {}, {}
    for g in range(nblk):                        # SendFluxCorrection, every block
        for nb in tree.neighbors(g, corners=False):
            if nb.rel == 0 and (nb.kind == "face" or flags[g][0][nb.eid]):
                slot_same[(nb.gid, tuple(-o for o in nb.ox), g)] = len(same_level)
                same_level.append(_item(fb=g, fid=nb.code))
            elif nb.rel == -1:
                slot_fine[(nb.gid, tuple(-o for o in nb.ox), g)] = len(to_coarser)
                to_coarser.append(_item(fb=g, fid=nb.code))
    set_same, clear, set_finer, average = [], [], [], []
    for g in range(nblk):                        # ReceiveFluxCorrection, every block
        eflag, nfine = flags[g]
        lv = tree.nblevel(g)
        for nb in tree.neighbors(g, corners=False):
            if nb.rel == 0 and (nb.kind == "face" or eflag[nb.eid]):
                set_same.append(_item(fb=slot_same[(g, nb.ox, nb.gid)], cb=g, fid=nb.code))
        for fid in range(6):                     # ClearCoarseFluxBoundary
            ox = [0, 0, 0]
            ox[fid >> 1] = 1 if fid & 1 else -1
            if lv[((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1] > 0:
                clear.append(_item(cb=g, fid=fid))
        for e in range(12):
            if not eflag[e]:
                clear.append(_item(cb=g, fid=6 + e))
        for nb in tree.neighbors(g, corners=False):
            if nb.rel == 1:
                set_finer.append(_item(fb=slot_fine[(g, nb.ox, nb.gid)], cb=g, fid=nb.code, fi1=nb.fi1, fi2=nb.fi2))
        for fid in range(6):                     # AverageFluxBoundary
            if not tree.boundary_is_block(g, fid):
                continue
            ox = [0, 0, 0]
            ox[fid >> 1] = 1 if fid & 1 else -1
            rel = lv[((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1]
            if rel == 0:
                average.append(_item(cb=g, fid=fid, fi1=0))
            elif rel > 0:
                average.append(_item(cb=g, fid=fid, fi1=1))
        for e in range(12):
            if nfine[e] != 1:
                average.append(_item(cb=g, fid=6 + e, fi1=nfine[e]))
    return EMFPlan(to_coarser, same_level, set_same, clear, set_finer, average)


# ---------------------------------------------------------------- cell-centred ghosts
@dataclass
class CCPlan:
    """Work lists (dicts of agx_ref_item fields) of one cell-centred ghost exchange with
    prolongation, to run in this order: copy_same (AGX_REF_COPY_CC_BB, block -> block),
    restrict_send (RESTRICT_CC, block -> its own coarse buffer), copy_from_finer
    (COPY_CC_CB, coarse buffer -> coarser block's ghosts), copy_to_finer (COPY_CC_BC,
    block -> finer block's coarse buffer), restrict_same (RESTRICT_CC) and prolong
    (PROLONG_CC).  per_neighbor lists, for each block in neighbour order, its coarser
    neighbours' (restrict_same items, prolong item): the reference's own order."""
    copy_same: list
    restrict_send: list
    copy_from_finer: list
    copy_to_finer: list
    restrict_same: list
    prolong: list
    per_neighbor: list


def _box(fb, cb, si, ei, sj, ej, sk, ek, dst=(0, 0, 0)):
    return dict(fb=fb, cb=cb, si=si, ei=ei, sj=sj, ej=ej, sk=sk, ek=ek, fid=dst[0], fi1=dst[1], fi2=dst[2], aux=0)


def _fi_axes(ox):
    """which of fi1 / fi2 splits each direction with ox == 0 (bvals_cc.cpp:262-283)"""
    ax = [None, None, None]
    free = [d for d in range(3) if ox[d] == 0]
    if len(free) == 2:
        ax[free[0]], ax[free[1]] = "fi1", "fi2"
    elif len(free) == 1:
        ax[free[0]] = "fi1"
    return ax


def cc_plan(tree, nx, ng):
    """Ghost exchange of cubic blocks of nx^3 cells and ng ghosts; coarse buffers of
    nx/2 + 2 ng cells, interior from ng (mesh/meshblock.cpp:59-85)."""
    nblk = len(tree.leaves)
    lo, hi = ng, ng + nx - 1                    # is, ie (same in j, k)
    clo, chi = ng, ng + nx // 2 - 1             # cis, cie
    cng = (ng + 1) // 2 + 1                     # cnghost
    h = nx // 2
    copy_same, restrict_send, copy_from_finer, copy_to_finer = [], [], [], []
    restrict_same, prolong, per_neighbor = [], [], []

    def back(src, dst, ox):
        """the neighbour entry of block src that points at block dst from the other side"""
        o = tuple(-x for x in ox)
        for nb in tree.neighbors(src):
            if nb.gid == dst and nb.ox == o:
                return nb
        raise KeyError((src, dst, ox))

    for g in range(nblk):
        lx = tree.leaves[g][1:]
        for nb in tree.neighbors(g):
            ox, s = nb.ox, nb.gid
            if nb.rel == 0:
                # LoadBoundaryBufferSameLevel of s towards g, SetBoundarySameLevel of g
                src, dst = [], []
                for d in range(3):
                    o = -ox[d]
                    src.append((hi - ng + 1 if o > 0 else lo, lo + ng - 1 if o < 0 else hi))
                    dst.append(lo if ox[d] == 0 else (hi + 1 if ox[d] > 0 else lo - ng))
                copy_same.append(_box(s, g, *src[0], *src[1], *src[2], dst=dst))
            elif nb.rel == 1:
                # s restricts its slab into its coarse buffer and packs it
                # (LoadBoundaryBufferToCoarser); g unpacks (SetBoundaryFromFiner)
                cn = ng - 1
                src, dst = [], []
                ax = _fi_axes(ox)
                for d in range(3):
                    o = -ox[d]
                    src.append((chi - cn if o > 0 else clo, clo + cn if o < 0 else chi))
                    if ox[d] == 0:
                        dst.append(lo + (h if getattr(nb, ax[d]) == 1 else 0))
                    else:
                        dst.append(hi + 1 if ox[d] > 0 else lo - ng)
                restrict_send.append(_box(s, s, *src[0], *src[1], *src[2]))
                copy_from_finer.append(_box(s, g, *src[0], *src[1], *src[2], dst=dst))
            else:
                # s packs for its finer neighbour g (LoadBoundaryBufferToFiner), g unpacks
                # into its coarse buffer (SetBoundaryFromCoarser)
                sb = back(s, g, ox)
                cn = cng - 1
                ax = _fi_axes(sb.ox)
                src, dst = [], []
                for d in range(3):
                    o = sb.ox[d]
                    a, b = (hi - cn if o > 0 else lo), (lo + cn if o < 0 else hi)
                    if o == 0:
                        if getattr(sb, ax[d]) == 1:
                            a += h - cng
                        else:
                            b -= h - cng
                    src.append((a, b))
                    if ox[d] == 0:
                        dst.append(clo if lx[d] & 1 == 0 else clo - cng)
                    else:
                        dst.append(chi + 1 if ox[d] > 0 else clo - cng)
                copy_to_finer.append(_box(s, g, *src[0], *src[1], *src[2], dst=dst))
        # ProlongateBoundaries: per coarser neighbour, restrict the same-level ghost
        # positions around it, then prolongate its ghost range
        lv = tree.nblevel(g)
        for nb in tree.neighbors(g):
            if nb.rel >= 0:
                continue
            ox = nb.ox
            rs = []
            rng = [range(max(ox[d] - 1, -1), min(ox[d] + 1, 1) + 1) for d in range(3)]
            for nk in rng[2]:
                for nj in rng[1]:
                    for ni in rng[0]:
                        if (ni, nj, nk) == (0, 0, 0) or lv[((nk + 1) * 3 + nj + 1) * 3 + ni + 1] != 0:
                            continue
                        r = []
                        for d, nn in enumerate((ni, nj, nk)):
                            if nn == 0:
                                a, b = clo, chi
                                if ox[d] == 1:
                                    a = chi
                                elif ox[d] == -1:
                                    b = clo
                            elif nn == 1:
                                a = b = chi + 1
                            else:
                                a = b = clo - 1
                            r.append((a, b))
                        rs.append(_box(g, g, *r[0], *r[1], *r[2]))
            pr = []
            cn = cng - 1
            for d in range(3):
                if ox[d] == 0:
                    a, b = clo, chi
                    if lx[d] & 1 == 0:
                        b += cn
                    else:
                        a -= cn
                elif ox[d] > 0:
                    a, b = chi + 1, chi + cn
                else:
                    a, b = clo - cn, clo - 1
                pr.append((a, b))
            p = _box(g, g, *pr[0], *pr[1], *pr[2])
            restrict_same += rs
            prolong.append(p)
            per_neighbor.append((rs, p))
    return CCPlan(copy_same, restrict_send, copy_from_finer, copy_to_finer, restrict_same, prolong, per_neighbor)


def block_coords(tree, nx, ng, coarse=False, box=(1.0, 1.0, 1.0)):
    """Per-block x?v / dx?f tables (dict of (nb, n) float64 arrays) of cubic blocks of
    nx^3 cells on a uniform Cartesian root grid covering [0, box): the blocks' own cells
    (n = nx + 2 ng) or their coarse buffers (n = nx/2 + 2 ng, coarse cell ng starting
    at the block's first face)."""
    n = (nx // 2 + 2 * ng) if coarse else nx + 2 * ng
    tab = {k: [] for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")}
    for (lev, *lx) in tree.leaves:
        for d in range(3):
            h = box[d] / ((tree.nrb[d] * nx) << lev)
            if coarse:
                f = (lx[d] * nx + 2 * (np.arange(n + 1) - ng)) * h
            else:
                f = (lx[d] * nx - ng + np.arange(n + 1)) * h
            tab[f"x{d + 1}v"].append(0.5 * (f[1:] + f[:-1]))
            tab[f"dx{d + 1}f"].append(f[1:] - f[:-1])
    return {k: np.ascontiguousarray(np.array(v)) for k, v in tab.items()}


def _meshgen_x(index, nrange):
    """ComputeMeshGeneratorX with the symmetric form of the uniform generator
    (mesh/mesh.hpp:389-405)"""
    noffset = index - nrange // 2
    noffset_ceil = index - (nrange + 1) // 2
    return float(noffset + noffset_ceil) / (2.0 * nrange)


def _uniform_gen(x, xmin, xmax):
    """UniformMeshGeneratorX? (mesh/mesh.hpp:465-488)"""
    return 0.5 * (xmin + xmax) + (x * xmax - x * xmin)


def mesh_coords(tree, nx, ng, box, coarse=False):
    """Per-block x?v / dx?f tables (dict of (nb, n) float64 arrays) as the reference's
    Coordinates builds them for each MeshBlock of a uniform Cartesian mesh over
    box = ((x1min, x1max), (x2min, x2max), (x3min, x3max)): block extents from
    Mesh::SetBlockSizeAndBoundaries (mesh/mesh.cpp:1668-1730), faces from the uniform
    generator at each face's offset on the block's level, the block's own first and last
    faces set to its extents, dx?f = the block width over its cells, x?v the face
    midpoints (coordinates/coordinates.cpp:86-145, cartesian.cpp:26-36).  coarse: the
    coarse buffer (nx/2 cells, every other fine face, coarse_flag of the same
    constructor)."""
    tab = {k: [] for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")}
    nc = nx // 2 if coarse else nx
    for (lev, *lx) in tree.leaves:
        for d in range(3):
            xmin, xmax = float(box[d][0]), float(box[d][1])
            nblk = tree.nrb[d] << lev
            nroot = (tree.nrb[d] * nx) << lev
            bmin = xmin if lx[d] == 0 else _uniform_gen(_meshgen_x(lx[d], nblk), xmin, xmax)
            bmax = xmax if lx[d] == nblk - 1 else _uniform_gen(_meshgen_x(lx[d] + 1, nblk), xmin, xmax)
            dx = (bmax - bmin) / nc
            n = nc + 2 * ng
            xf = []
            for i in range(n + 1):
                off = ((i - ng) * 2 if coarse else (i - ng)) + lx[d] * nx
                xf.append(_uniform_gen(_meshgen_x(off, nroot), xmin, xmax))
            xf[ng] = bmin
            xf[ng + nc] = bmax
            tab[f"x{d + 1}v"].append([0.5 * (xf[i + 1] + xf[i]) for i in range(n)])
            tab[f"dx{d + 1}f"].append([dx] * n)
    return {k: np.ascontiguousarray(np.array(v, dtype=np.float64)) for k, v in tab.items()}


def face_positions(tree, nx, ng, box):
    """per block the three arrays of its n + 1 face positions (n = nx + 2 ng) along x1,
    x2, x3: the uniform generator at each face's offset on the block's level, the block's
    own first and last faces set to its extents (as mesh_coords)"""
    out = []
    for (lev, *lx) in tree.leaves:
        per = []
        for d in range(3):
            xmin, xmax = float(box[d][0]), float(box[d][1])
            nblk, nroot = tree.nrb[d] << lev, (tree.nrb[d] * nx) << lev
            bmin = xmin if lx[d] == 0 else _uniform_gen(_meshgen_x(lx[d], nblk), xmin, xmax)
            bmax = xmax if lx[d] == nblk - 1 else _uniform_gen(_meshgen_x(lx[d] + 1, nblk), xmin, xmax)
            xf = [_uniform_gen(_meshgen_x(i - ng + lx[d] * nx, nroot), xmin, xmax) for i in range(nx + 2 * ng + 1)]
            xf[ng], xf[ng + nx] = bmin, bmax
            per.append(np.array(xf))
        out.append(per)
    return out


def flux_plan(tree):
    """Work list of the hydro flux correction (flux_correction_cc.cpp:77-169,240-283): one
    AGX_REF_FLUX_FROM_FINER item per finer face neighbour of a block, fb = the finer
    block, cb = the block, fid = its face, fi1/fi2 = the quadrant, in the order of
    ReceiveFluxCorrection over blocks and their neighbours.  The four items of a face
    write disjoint quadrants."""
    out = []
    for g in range(len(tree.leaves)):
        for nb in tree.neighbors(g, corners=False):
            if nb.kind == "face" and nb.rel == 1:
                out.append(_item(fb=nb.gid, cb=g, fid=nb.fid, fi1=nb.fi1, fi2=nb.fi2))
    return out


# ---------------------------------------------------------------- face-field ghosts
@dataclass
class FCPlan:
    """Work lists of one face-field (x1f, x2f, x3f) ghost exchange with prolongation, to
    run in this order: restrict_send[c] (AGX_REF_RESTRICT_FC1..3, block faces -> own
    coarse faces), pack_b (PACK_FC_B: same-level and to-finer packs, buffer slots 0..),
    pack_c (PACK_FC_C: to-coarser packs, slots from len(pack_b)), unpack_b (UNPACK_FC_B:
    same-level and from-finer, into block faces), unpack_c (UNPACK_FC_C: from coarser,
    into coarse faces), restrict_same[c] (RESTRICT_FC), prolong_shared[c] (PROLONG_FC1..3)
    and prolong_int (PROLONG_FC_INT).  aux = component (1..3) + 4 x launch group (the
    neighbour class and quadrant, in the reference's order).  per_neighbor: per block and
    coarser neighbour, (restrict[c], shared[c], internal item) in the reference's order."""
    restrict_send: list
    pack_b: list
    pack_c: list
    unpack_b: list
    unpack_c: list
    restrict_same: list
    prolong_shared: list
    prolong_int: list
    per_neighbor: list


def _group(nb):
    cls = {"face": 0, "edge": 1, "corner": 2}[nb.kind]
    return cls * 4 + ((nb.fi1 + 2 * nb.fi2) if nb.rel == 1 else 0)


def _rng_item(fb, cb, r, aux=0):
    (si, ei), (sj, ej), (sk, ek) = r
    return dict(fb=fb, cb=cb, si=si, ei=ei, sj=sj, ej=ej, sk=sk, ek=ek, fid=0, fi1=0, fi2=0, aux=aux)


def fc_plan(tree, nx, ng):
    """Face-field ghost exchange of cubic blocks (bvals/fc/bvals_fc.cpp:344-977, the face
    part of RestrictGhostCellsOnSameLevel and ProlongateGhostCells, bvals_refine.cpp:
    289-330, 483-521)."""
    nblk = len(tree.leaves)
    lo, hi = ng, ng + nx - 1
    clo, chi = ng, ng + nx // 2 - 1
    cng = (ng + 1) // 2 + 1
    h = nx // 2
    ml = tree.maxlev > 0                       # Mesh::multilevel
    fi_ax = _fi_axes

    def fi_of(nb, ax, d):
        return getattr(nb, ax[d]) if ax[d] else 0

    def sl_pack(o, nonface, c):
        r = []
        for d in range(3):
            if d == c:
                a, b = (lo, hi + 1) if o[d] == 0 else ((hi - ng + 1, hi) if o[d] > 0 else (lo + 1, lo + ng))
                if ml and nonface:
                    if o[d] > 0:
                        b += 1
                    elif o[d] < 0:
                        a -= 1
            else:
                a, b = (lo, hi) if o[d] == 0 else ((hi - ng + 1, hi) if o[d] > 0 else (lo, lo + ng - 1))
            r.append((a, b))
        return r

    def sl_unpack(ox, nonface, c):
        r = []
        for d in range(3):
            if d == c:
                a, b = (lo, hi + 1) if ox[d] == 0 else ((hi + 2, hi + ng + 1) if ox[d] > 0 else (lo - ng, lo - 1))
                if ml and nonface:
                    if ox[d] > 0:
                        a -= 1
                    elif ox[d] < 0:
                        b += 1
            else:
                a, b = (lo, hi) if ox[d] == 0 else ((hi + 1, hi + ng) if ox[d] > 0 else (lo - ng, lo - 1))
            r.append((a, b))
        return r

    def tc_pack(o, nonface, c):                # LoadBoundaryBufferToCoarser, coarse faces
        r = []
        for d in range(3):
            if d == c:
                a, b = (clo, chi + 1) if o[d] == 0 else ((chi - ng + 1, chi) if o[d] > 0 else (clo + 1, clo + ng))
                if nonface:
                    if o[d] > 0:
                        b += 1
                    elif o[d] < 0:
                        a -= 1
            else:
                a, b = (clo, chi) if o[d] == 0 else ((chi - ng + 1, chi) if o[d] > 0 else (clo, clo + ng - 1))
            r.append((a, b))
        return r

    def ff_unpack(nb, c):                      # SetBoundaryFromFiner
        ox, nonface, ax = nb.ox, nb.kind != "face", fi_ax(nb.ox)
        r = []
        for d in range(3):
            e = d == c
            if ox[d] == 0:
                a, b = lo, hi + e
                if fi_of(nb, ax, d) == 1:
                    a += h
                else:
                    b -= h
            elif ox[d] > 0:
                a, b = hi + 1 + e, hi + ng + e
            else:
                a, b = lo - ng, lo - 1
            if e and nonface:
                if ox[d] > 0:
                    a -= 1
                elif ox[d] < 0:
                    b += 1
            r.append((a, b))
        return r

    def tf_pack(sb, c):                        # LoadBoundaryBufferToFiner, sender's view sb
        o, ax = sb.ox, fi_ax(sb.ox)
        r = []
        for d in range(3):
            e = d == c
            if o[d] == 0:
                a, b = lo, hi + e
                if fi_of(sb, ax, d) == 1:
                    a += h - cng
                else:
                    b -= h - cng
            elif o[d] > 0:
                a, b = (hi + 1 - cng, hi + 1) if e else (hi - (cng - 1), hi)
            else:
                a, b = (lo, lo + cng) if e else (lo, lo + cng - 1)
            r.append((a, b))
        return r

    def fc_unpack(ox, lx, c):                  # SetBoundaryFromCoarser, coarse faces
        r = []
        for d in range(3):
            e = d == c
            if ox[d] == 0:
                a, b = clo, chi + e
                if lx[d] & 1 == 0:
                    b += cng
                else:
                    a -= cng
            elif ox[d] > 0:
                a, b = chi + 1, chi + cng + e
            else:
                a, b = clo - cng, clo - 1 + e
            r.append((a, b))
        return r

    def back(src, dst, ox):
        o = tuple(-x for x in ox)
        for nb in tree.neighbors(src):
            if nb.gid == dst and nb.ox == o:
                return nb
        raise KeyError((src, dst, ox))

    restrict_send = [[], [], []]
    pack_b, pack_c, unpack_b, unpack_c = [], [], [], []
    pend_c = []                                # unpacks reading pack_c slots (offset later)
    for g in range(nblk):
        lx = tree.leaves[g][1:]
        for nb in tree.neighbors(g):
            s, ox, grp = nb.gid, nb.ox, _group(nb)
            o = tuple(-x for x in ox)
            for c in range(3):
                aux = (c + 1) + 4 * grp
                if nb.rel == 0:
                    slot = len(pack_b)
                    pack_b.append(_rng_item(s, s, sl_pack(o, nb.kind != "face", c), c + 1))
                    unpack_b.append(_rng_item(slot, g, sl_unpack(ox, nb.kind != "face", c), aux))
                elif nb.rel == 1:
                    r = tc_pack(o, nb.kind != "face", c)
                    restrict_send[c].append(_rng_item(s, s, r))
                    pend_c.append(len(unpack_b))
                    unpack_b.append(_rng_item(len(pack_c), g, ff_unpack(nb, c), aux))
                    pack_c.append(_rng_item(s, s, r, c + 1))
                else:
                    sb = back(s, g, ox)
                    slot = len(pack_b)
                    pack_b.append(_rng_item(s, s, tf_pack(sb, c), c + 1))
                    unpack_c.append(_rng_item(slot, g, fc_unpack(ox, lx, c), aux))
    for m in pend_c:
        unpack_b[m]["fb"] += len(pack_b)

    restrict_same, prolong_shared, prolong_int, per_neighbor = [[], [], []], [[], [], []], [], []
    cn = cng - 1
    for g in range(nblk):
        lx = tree.leaves[g][1:]
        lv = tree.nblevel(g)
        L = lambda o: lv[((o[2] + 1) * 3 + o[1] + 1) * 3 + o[0] + 1]   # noqa: E731
        for nb in tree.neighbors(g):
            if nb.rel >= 0:
                continue
            ox = nb.ox
            rs = [[], [], []]
            rng = [range(max(ox[d] - 1, -1), min(ox[d] + 1, 1) + 1) for d in range(3)]
            for nk in rng[2]:
                for nj in rng[1]:
                    for ni in rng[0]:
                        pos = (ni, nj, nk)
                        if pos == (0, 0, 0) or L(pos) != 0:
                            continue
                        r = []
                        for d, nn in enumerate(pos):
                            if nn == 0:
                                a, b = clo, chi
                                if ox[d] == 1:
                                    a = chi
                                elif ox[d] == -1:
                                    b = clo
                            elif nn == 1:
                                a = b = chi + 1
                            else:
                                a = b = clo - 1
                            r.append((a, b))
                        for c in range(3):
                            a, b = r[c][0], r[c][1] + 1
                            pm = list(pos)
                            pm[c] -= 1
                            pp = list(pos)
                            pp[c] += 1
                            if a == clo and L(pm) < 0:
                                a += 1
                            if b == chi + 1 and L(pp) < 0:
                                b -= 1
                            rr = list(r)
                            rr[c] = (a, b)
                            rs[c].append(_rng_item(g, g, rr))
            cell = []
            for d in range(3):
                if ox[d] == 0:
                    a, b = clo, chi
                    if lx[d] & 1 == 0:
                        b += cn
                    else:
                        a -= cn
                elif ox[d] > 0:
                    a, b = chi + 1, chi + cn
                else:
                    a, b = clo - cn, clo - 1
                cell.append((a, b))
            sh = []
            for c in range(3):
                a, b = cell[c][0], cell[c][1] + 1
                om = list(ox)
                om[c] -= 1
                op = list(ox)
                op[c] += 1
                if ox[c] >= 0 and L(om) >= 0:
                    a += 1
                if ox[c] <= 0 and L(op) >= 0:
                    b -= 1
                rr = list(cell)
                rr[c] = (a, b)
                sh.append(_rng_item(g, g, rr))
            it = _rng_item(g, g, cell)
            for c in range(3):
                restrict_same[c] += rs[c]
                prolong_shared[c].append(sh[c])
            prolong_int.append(it)
            per_neighbor.append((rs, sh, it))
    return FCPlan(restrict_send, pack_b, pack_c, unpack_b, unpack_c, restrict_same, prolong_shared, prolong_int,
                  per_neighbor)


# ---------------------------------------------------------------- MHD ghost prolongation
def mhd_prolong_plan(tree, nx, ng, cc=None, fc=None):
    """ProlongateBoundaries for MHD on every block (bvals/bvals_refine.cpp:93-229): per
    coarser neighbour, restriction of the same-level ghosts (u, faces), ConservedToPrimitive
    on the coarse buffer over the prolongation range widened by one cell towards every
    existing neighbour (ApplyPhysicalBoundariesOnCoarseLevel, :336-384), prolongation of
    the primitives and faces, then bcc and u on the fine cells (ProlongateGhostCells).

    Returns (slots, order).  order: one dict of single-step lists per block and coarser
    neighbour, in the reference's order.  slots[k]: the same lists for the k-th coarser
    neighbour of every block, batched; running the slots in turn gives the reference's
    per-block sequence (a floor in ConservedToPrimitive makes the sequence matter)."""
    cc = cc or cc_plan(tree, nx, ng)
    fc = fc or fc_plan(tree, nx, ng)
    order, slots = [], []
    m = 0
    for g in range(len(tree.leaves)):
        k = 0
        for nb in tree.neighbors(g):
            if nb.rel >= 0:
                continue
            rs_c, p = cc.per_neighbor[m]
            rs_f, sh, it = fc.per_neighbor[m]
            m += 1
            r = [(p["si"], p["ei"]), (p["sj"], p["ej"]), (p["sk"], p["ek"])]
            for d in range(3):
                if nb.ox[d] == 0:
                    e = [0, 0, 0]
                    e[d] = -1
                    fm = 1 if tree.find(g, tuple(e)) is not None else 0
                    e[d] = 1
                    fp = 1 if tree.find(g, tuple(e)) is not None else 0
                    if not (fm and fp):
                        # ApplyPhysicalBoundariesOnCoarseLevel (:401-445) would run the
                        # domain's boundary function on the coarse buffer here
                        raise ValueError(f"block {g} has a coarser neighbour and a physical x{d + 1} "
                                         "boundary: boundary functions on the coarse buffer are "
                                         "not built (periodic meshes, or refinement away from "
                                         "the domain boundary)")
                else:
                    fm = fp = 1
                r[d] = (r[d][0] - fm, r[d][1] + fp)
            step = dict(restrict_cc=list(rs_c), restrict_fc=[list(x) for x in rs_f],
                        c2p=[_rng_item(g, g, r)], prolong_cc=[p], prolong_fc=[[x] for x in sh],
                        prolong_int=[it], p2c=[dict(it)])
            order.append(step)
            if len(slots) <= k:
                slots.append(dict(restrict_cc=[], restrict_fc=[[], [], []], c2p=[], prolong_cc=[],
                                  prolong_fc=[[], [], []], prolong_int=[], p2c=[]))
            sl = slots[k]
            for key in ("restrict_cc", "c2p", "prolong_cc", "prolong_int", "p2c"):
                sl[key] += step[key]
            for c in range(3):
                sl["restrict_fc"][c] += step["restrict_fc"][c]
                sl["prolong_fc"][c] += step["prolong_fc"][c]
            k += 1
    return slots, order


# ---- MeshBlocks of uneven sizes
#
# The reference's MeshBlocks all have the same cell count.  A mesh whose blocks have
# 2^p times that many cells a side on any level is run as the tree of equal blocks it
# divides into: block (level, x, y, z, p) has its lower corner at the level-`level`
# position (x, y, z) counted in blocks of the base size nx (each a multiple of 2^p) and
# becomes the 8^p leaves (level, x + a, y + b, z + c), 0 <= a, b, c < 2^p.  A block's
# update does not depend on where the mesh is cut into blocks, so one device mesh of the
# leaves runs every block of every size and level in the same launches.

def uneven_leaves(blocks):
    """the leaves of uneven `blocks` and, per leaf, (block index, a, b, c)"""
    leaves, owner = [], []
    for bi, (lev, x, y, z, p) in enumerate(blocks):
        s = 1 << p
        if x % s or y % s or z % s:
            raise ValueError(f"block {bi}: corner ({x}, {y}, {z}) is not a multiple of {s}")
        for c in range(s):
            for b in range(s):
                for a in range(s):
                    leaves.append((lev, x + a, y + b, z + c))
                    owner.append((bi, a, b, c))
    return leaves, owner


def _window(name, a, b, c, nx, ng):
    """index window (k, j, i) of leaf (a, b, c) in its block's array, ghosts included"""
    n = nx + 2 * ng
    e = {"b1f": (0, 0, 1), "b2f": (0, 1, 0), "b3f": (1, 0, 0)}.get(name, (0, 0, 0))
    return (slice(c * nx, c * nx + n + e[0]), slice(b * nx, b * nx + n + e[1]), slice(a * nx, a * nx + n + e[2]))


def _interior(name, nx, ng):
    """a leaf's own cells (or faces, both bounding faces included) within its window"""
    e = {"b1f": (0, 0, 1), "b2f": (0, 1, 0), "b3f": (1, 0, 0)}.get(name, (0, 0, 0))
    return tuple(slice(ng, ng + nx + e[d]) for d in range(3))


def split_uneven(blocks, owner, name, arrays, nx, ng):
    """leaf arrays of variable `name` ("u", "w", "bcc": (nvar, n, n, n); "b1f", "b2f",
    "b3f": the face layouts) from one array per uneven block, ghosts included (a leaf's
    ghosts inside its block are its neighbours' cells)"""
    out = []
    for bi, a, b, c in owner:
        w = _window(name, a, b, c, nx, ng)
        src = arrays[bi]
        out.append(np.array(src[(Ellipsis,) + w]))
    return np.stack(out)


def merge_uneven(blocks, owner, name, leaves_arr, nx, ng):
    """one array per uneven block from the leaf arrays: cells and faces from the leaf that
    owns them, the block's ghost zones from its boundary leaves"""
    out = []
    for lev, x, y, z, p in blocks:
        n = (nx << p) + 2 * ng
        e = {"b1f": (0, 0, 1), "b2f": (0, 1, 0), "b3f": (1, 0, 0)}.get(name, (0, 0, 0))
        lead = leaves_arr.shape[1:-3]
        out.append(np.zeros(lead + (n + e[0], n + e[1], n + e[2])))
    for g, (bi, a, b, c) in enumerate(owner):   # whole windows: the block's ghost zones
        out[bi][(Ellipsis,) + _window(name, a, b, c, nx, ng)] = leaves_arr[g]
    inner = _interior(name, nx, ng)
    for g, (bi, a, b, c) in enumerate(owner):   # then every leaf's own cells and faces
        w = _window(name, a, b, c, nx, ng)
        dst = tuple(slice(w[d].start + inner[d].start, w[d].start + inner[d].stop) for d in range(3))
        out[bi][(Ellipsis,) + dst] = leaves_arr[g][(Ellipsis,) + inner]
    return out
