"""Static-refinement MHD cycles for the tests: periodic block trees, the package's smooth
divergence-free state on every block of every level (pgen.divfree_blocks), diagnostics
(conservation, div B, agreement of the field through refinement boundaries) and the
oracle's composition of the cycle (the reference's task order with refinement,
time_integrator.cpp:1460-1558) from the oracle's task-level steps and refinement
operators, on its dense arrays."""
import itertools

import numpy as np

from oracle import orc
from tests import emf_cases as ec
from tests import ghost_cases as gc
from tests.conftest import load_pgen, load_smr

GAMMA = 5.0 / 3.0


def periodic_tree(nrb, refine=(), refine2=()):
    """root blocks in `refine` refined once, level-1 locations in `refine2` once more"""
    smr = load_smr()
    leaves = [(0, x, y, z) for z, y, x in itertools.product(*(range(n) for n in nrb[::-1]))
              if (x, y, z) not in refine]
    for (x, y, z) in refine:
        leaves += [(1, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)
                   if (2 * x + c1, 2 * y + c2, 2 * z + c3) not in refine2]
    for (x, y, z) in refine2:
        leaves += [(2, 2 * x + c1, 2 * y + c2, 2 * z + c3) for c3, c2, c1 in itertools.product(range(2), repeat=3)]
    return smr.SMRTree(nrb, leaves, periodic=(True, True, True))


def face_coords(tree, nx, ng, box):
    return load_smr().face_positions(tree, nx, ng, box)


def smooth_state(tree, nx, ng, box, amp=0.2, b0=(0.6, 0.0, 0.0), seed=0, iso=False):
    """pgen.divfree_blocks on the tree's blocks (iso: the energy slot zero)"""
    u, b1, b2, b3 = load_pgen().divfree_blocks(face_coords(tree, nx, ng, box), box, GAMMA, amp, b0, seed)
    if iso:
        u[:, 4] = 0.0
    return u, b1, b2, b3


def max_divb(tree, nx, ng, box, b1, b2, b3):
    """max over interior cells of |div B| dx / max|B| (each block with its own widths)"""
    xf = face_coords(tree, nx, ng, box)
    s = slice(ng, ng + nx)
    bmax = max(np.abs(b1).max(), np.abs(b2).max(), np.abs(b3).max())
    worst = 0.0
    for g in range(len(tree.leaves)):
        d1, d2, d3 = (np.diff(x)[s] for x in xf[g])
        div = ((b1[g][s, s, ng + 1:ng + nx + 1] - b1[g][s, s, s]) / d1[None, None, :] +
               (b2[g][s, ng + 1:ng + nx + 1, s] - b2[g][s, s, s]) / d2[None, :, None] +
               (b3[g][ng + 1:ng + nx + 1, s, s] - b3[g][s, s, s]) / d3[:, None, None])
        worst = max(worst, float(np.abs(div).max() * d1.min() / bmax))
    return worst


def interface_mismatch(tree, nx, ng, b1, b2, b3):
    """max over coarse/fine interfaces of |coarse face value - mean of the 2x2 fine faces
    that cover it| / max|B|: the two copies of the field through a refinement boundary
    agree while the EMF correction makes both sides' CT use the same edge values"""
    s = [b1, b2, b3]
    bmax = max(np.abs(b1).max(), np.abs(b2).max(), np.abs(b3).max())
    h = nx // 2
    worst = 0.0
    for g in range(len(tree.leaves)):
        for nb in tree.neighbors(g, corners=False):
            if nb.kind != "face" or nb.rel != 1:
                continue
            d, side = nb.fid >> 1, nb.fid & 1
            f = s[d]
            xc = ng + nx if side else ng            # coarse block's face index along d
            xf = ng if side else ng + nx            # the finer block's face on the interface
            # tangential index ranges: the quadrant (fi1, fi2) of the coarse face
            t = [a for a in range(3) if a != d]     # fast, slow tangential axes (x-order)
            lo = [ng + (nb.fi1 * h), ng + (nb.fi2 * h)]
            for q in range(h):
                for p in range(h):
                    ic = [0, 0, 0]
                    ic[d], ic[t[0]], ic[t[1]] = xc, lo[0] + p, lo[1] + q
                    cv = f[g][ic[2], ic[1], ic[0]]
                    acc = 0.0
                    for bq in range(2):
                        for bp in range(2):
                            jf = [0, 0, 0]
                            jf[d], jf[t[0]], jf[t[1]] = xf, ng + 2 * p + bp, ng + 2 * q + bq
                            acc += f[nb.gid][jf[2], jf[1], jf[0]]
                    worst = max(worst, abs(cv - 0.25 * acc) / bmax)
    return worst


def totals(tree, nx, ng, box, u):
    """volume sums over every block's interior of the conserved variables"""
    xf = face_coords(tree, nx, ng, box)
    s = slice(ng, ng + nx)
    nv = u.shape[1]
    out = np.zeros(nv)
    for g in range(len(tree.leaves)):
        vol = float(np.diff(xf[g][0])[ng]) * float(np.diff(xf[g][1])[ng]) * float(np.diff(xf[g][2])[ng])
        out += u[g][:, s, s, s].reshape(nv, -1).sum(1) * vol
    return out


class OracleSMR:
    """The oracle's static-refinement cycle: the same task order as the library's
    SMRMesh (fluxes, flux correction, corner EMFs, EMF correction, update and CT, ghost
    exchange with prolongation, C2P), each step the oracle's restatement on its dense
    arrays; the prolongation in the reference's per-block neighbour order."""

    def __init__(self, tree, nx, ng=2, box=((0.0, 1.0),) * 3, cfl=0.3, xorder=2, rsolver=0, emf_skip=(),
                 flux_skip=False, eos=0, iso_cs=1.0):
        smr = load_smr()
        self.tree, self.nx, self.ng, self.xorder = tree, nx, ng, xorder
        self.emf_skip = emf_skip   # steps of the EMF correction left out (emf_cases.run_oracle)
        self.flux_skip = flux_skip   # leave out the hydro flux correction
        nb = len(tree.leaves)
        desc = orc.make_desc(nx=(nb * nx, nx, nx), bnx=(nx, nx, nx), box=((0.0, float(nb)), (0.0, 1.0), (0.0, 1.0)),
                             nghost=ng, xorder=xorder, rsolver=rsolver, gamma=GAMMA, cfl=cfl, eos=eos, iso_cs=iso_cs)
        self.iso, self.nh = eos == 1, 4 if eos == 1 else 5
        self.desc = desc
        self.o = orc.Oracle(desc)
        self.o.set_levels(tree.nrb, box, tree.leaves)
        self.ftab = smr.mesh_coords(tree, nx, ng, box)
        self.ctab = smr.mesh_coords(tree, nx, ng, box, coarse=True)
        self.g = orc.RefGeom(nx, nx, nx, ng, ng, 0)
        self.flux = [orc.RefItem(**d) for d in smr.flux_plan(tree)]
        nc = nx // 2 + 2 * ng
        self.A = dict(cu=np.zeros((nb, self.nh, nc, nc, nc)), cw=np.zeros((nb, self.nh, nc, nc, nc)),
                      cbcc=np.zeros((nb, 3, nc, nc, nc)),
                      cb=[np.zeros(s) for s in ((nb, nc, nc, nc + 1), (nb, nc, nc + 1, nc), (nb, nc + 1, nc, nc))])

    def set(self, name, arr):
        self.o.set(name, arr)

    def get(self, name):
        return self.o.get(name)

    def time(self):
        return self.o.time()

    def boundary(self):
        o, nh = self.o, self.nh
        # the oracle's hydro arrays hold 5 slots; the isothermal operators take 4
        A = dict(self.A, u=np.ascontiguousarray(o.view("u")[:, :nh]), w=np.ascontiguousarray(o.view("w")[:, :nh]),
                 bcc=o.view("bcc"), b=[o.view(n) for n in ("b1f", "b2f", "b3f")])
        R = gc.run_oracle_mhd(orc, self.tree, A, self.nx, self.ng, self.ftab, self.ctab, order="reference",
                              gamma=GAMMA, dfloor=self.desc.dfloor, pfloor=self.desc.pfloor, iso=self.iso)
        for k in ("u", "w"):
            o.view(k)[:, :nh] = R[k]
        o.view("bcc")[...] = R["bcc"]
        for c, n in enumerate(("b1f", "b2f", "b3f")):
            o.view(n)[...] = R["b"][c]
        self.A = {k: R[k] for k in ("cu", "cw", "cbcc", "cb")}

    def initialize(self, tlim=1e300):
        self.boundary()
        self.o.L.orc_cons2prim(self.o.h)
        t, _, n = self.o.time()
        self.o.set_time(t, 1.7976931348623157e308, n)
        self.o.L.orc_new_time_step(self.o.h, tlim)

    def stage(self, s):
        o = self.o
        o.L.orc_calculate_fluxes(o.h, 1 if s == 1 else self.xorder)
        if self.flux and not self.flux_skip:
            f = [np.ascontiguousarray(o.view(n)[:, :self.nh]) for n in ("fx1", "fx2", "fx3")]
            orc.ref_flux_from_finer(self.g, self.nh, *f, self.ftab, *f, self.flux)
            for n, x in zip(("fx1", "fx2", "fx3"), f):
                o.view(n)[:, :self.nh] = x
        o.L.orc_corner_e(o.h)
        E, _ = ec.run_oracle(orc, self.tree, [o.view(n) for n in ("e1", "e2", "e3")], self.nx, self.ng, self.ftab,
                             skip=self.emf_skip)
        for n, e in zip(("e1", "e2", "e3"), E):
            o.view(n)[...] = e
        o.L.orc_integrate(o.h, s)
        self.boundary()
        o.L.orc_cons2prim(o.h)

    def cycle(self, tlim, n=1):
        done = 0
        while done < n:
            t, dt, nc = self.o.time()
            if t >= tlim:
                break
            self.stage(1)
            self.stage(2)
            self.o.set_time(t + dt, dt, nc + 1)
            self.o.L.orc_new_time_step(self.o.h, tlim)
            done += 1
        return done

    def close(self):
        self.o.close()
