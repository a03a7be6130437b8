"""Geometry and arrays for the refinement-operator tests (host-side numpy only).

Blocks of nx^3 interior cells with ng ghosts; each block's coarse buffer has
nc = nx/2 + 2 cng cells with cng = ng, coarse interior from cis = ng
(mesh/meshblock.cpp:59-85: ncc1 = nx1/2 + 2*NGHOST, cis = NGHOST).  Coordinates
come from one face list per block and direction.  The fine faces are every face of the
list and the coarse faces every second one, aligned so that coarse cell cng starts at
fine cell ng.  The spacing may be stretched at random.
"""
import numpy as np


class Geo:
    def __init__(self, nx=(8, 8, 8), ng=2, nb=3, seed=0, stretch=0.0, h=None, origin=None):
        self.nx, self.ng, self.cng, self.nb = tuple(nx), ng, ng, nb
        self.n = tuple(x + 2 * ng for x in nx)
        self.nc = tuple(x // 2 + 2 * self.cng for x in nx)
        pad = 2 * self.cng - ng
        rng = np.random.default_rng(seed)
        self.fine = {k: [] for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f", "x1f", "x2f", "x3f")}
        self.coarse = {k: [] for k in self.fine}
        for b in range(nb):
            for d in range(3):
                hd = (h[b][d] if h is not None else 1.0 / nx[d])
                x0 = (origin[b][d] if origin is not None else float(b))
                m = 2 * self.nc[d] + 2
                sp = hd * (1.0 + stretch * rng.uniform(-1.0, 1.0, m))
                G = np.concatenate([[0.0], np.cumsum(sp)])
                G = G - G[2 * self.cng] + x0          # coarse face cng (= fine face ng) at x0
                xf = G[pad:pad + self.n[d] + 1]
                Xf = G[0:2 * self.nc[d] + 1:2]
                for tab, f in ((self.fine, xf), (self.coarse, Xf)):
                    tab[f"x{d + 1}f"].append(f)
                    tab[f"x{d + 1}v"].append(0.5 * (f[1:] + f[:-1]))
                    tab[f"dx{d + 1}f"].append(f[1:] - f[:-1])
        for tab in (self.fine, self.coarse):
            for k in list(tab):
                tab[k] = np.ascontiguousarray(np.array(tab[k]))

    def is_(self, d):
        return self.ng

    def ie(self, d):
        return self.ng + self.nx[d] - 1

    def cis(self, d):
        return self.cng

    def cie(self, d):
        return self.cng + self.nx[d] // 2 - 1

    # array shapes, (nb, [nvar,] d3, d2, d1)
    def cells(self, nvar):
        n1, n2, n3 = self.n
        return (self.nb, nvar, n3, n2, n1)

    def ccells(self, nvar):
        c1, c2, c3 = self.nc
        return (self.nb, nvar, c3, c2, c1)

    def faces(self, d, coarse=False):
        n1, n2, n3 = self.nc if coarse else self.n
        return (self.nb, n3 + (d == 2), n2 + (d == 1), n1 + (d == 0))

    def fluxes(self, d, nvar):
        n1, n2, n3 = self.n
        return (self.nb, nvar, n3 + (d == 2), n2 + (d == 1), n1 + (d == 0))

    def edges(self, d):
        n1, n2, n3 = self.n
        if d == 0:
            return (self.nb, n3 + 1, n2 + 1, n1)
        if d == 1:
            return (self.nb, n3 + 1, n2, n1 + 1)
        return (self.nb, n3, n2 + 1, n1 + 1)


def item(cls, fb, cb=0, si=0, ei=0, sj=0, ej=0, sk=0, ek=0, fid=0, fi1=0, fi2=0):
    return cls(fb, cb, si, ei, sj, ej, sk, ek, fid, fi1, fi2, 0)


def geom_struct(cls, g):
    return cls(g.nx[0], g.nx[1], g.nx[2], g.ng, g.cng, 0)
