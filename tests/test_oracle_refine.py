"""The oracle's refinement operators (oracle/agx_oracle_refine.c) against what the
mathematics fixes.  No golden data exists for them: the reference cannot be built here,
and its regression tests that use SMR need its Python harness and a build.

- restriction of a linear field gives the field at the coarse centre (volume and area
  weights, mesh_refinement.cpp:106-344);
- minmod prolongation reproduces a linear field and keeps the coarse value at a local
  extremum (mesh_refinement.cpp:386-458);
- prolongation followed by restriction returns the coarse data (the children's offsets
  cancel in the volume-weighted mean), also on stretched spacing;
- face prolongation: shared faces restrict back to the coarse face values, and with the
  interior faces (ProlongateInternalField, mesh_refinement.cpp:831-946) a divergence-free
  coarse field gives a divergence-free fine field;
- flux correction and EMF restriction put a linear field's value at the coarse face or
  edge centre into the quadrant and buffer position the reference's loops address
  (flux_correction_cc.cpp:77-283, flux_correction_fc.cpp:315-418).

A transposed index, a wrong sign of an offset, a dropped term or a wrong quadrant fails
one of these.
"""
import numpy as np
import pytest

from oracle import orc
from tests.refine_cases import Geo, geom_struct, item


def G(g):
    return geom_struct(orc.RefGeom, g)


def It(*a, **k):
    return item(orc.RefItem, *a, **k)


def lin(coef, x, y, z):
    return coef[0] + coef[1] * x + coef[2] * y + coef[3] * z


def centres(tab, b):
    return tab["x1v"][b], tab["x2v"][b], tab["x3v"][b]


@pytest.mark.parametrize("stretch", [0.0, 0.3])
def test_restrict_cc_linear_and_prolong_roundtrip(stretch):
    g = Geo(stretch=stretch, seed=1)
    nvar = 2
    coef = [(1.5, 0.3, -0.7, 0.2), (-2.0, 1.1, 0.4, -0.9)]
    fine = np.zeros(g.cells(nvar))
    for b in range(g.nb):
        x, y, z = centres(g.fine, b)
        for n in range(nvar):
            fine[b, n] = lin(coef[n], x[None, None, :], y[None, :, None], z[:, None, None])
    coarse = np.zeros(g.ccells(nvar))
    rng = [It(b, b, g.cis(0), g.cie(0), g.cis(1), g.cie(1), g.cis(2), g.cie(2)) for b in range(g.nb)]
    orc.ref_restrict_cc(G(g), nvar, fine, g.fine, coarse, rng)
    sl = np.s_[g.cis(2):g.cie(2) + 1, g.cis(1):g.cie(1) + 1, g.cis(0):g.cie(0) + 1]
    for b in range(g.nb):
        X, Y, Z = centres(g.coarse, b)
        for n in range(nvar):
            ref = lin(coef[n], X[None, None, :], Y[None, :, None], Z[:, None, None])[sl]
            np.testing.assert_allclose(coarse[b, n][sl], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # prolongation of random coarse data, then restriction, returns the data
    rs = np.random.default_rng(2)
    cdat = rs.normal(size=g.ccells(nvar))
    f2 = np.zeros(g.cells(nvar))
    orc.ref_prolong_cc(G(g), nvar, cdat, g.coarse, f2, g.fine, rng)
    back = np.zeros(g.ccells(nvar))
    orc.ref_restrict_cc(G(g), nvar, f2, g.fine, back, rng)
    np.testing.assert_allclose(back[:, :, sl[0], sl[1], sl[2]], cdat[:, :, sl[0], sl[1], sl[2]], rtol=0, atol=1e-14 * 8)


@pytest.mark.parametrize("stretch", [0.0, 0.4])
def test_prolong_cc_reproduces_linear_fields(stretch):
    g = Geo(stretch=stretch, seed=3)
    coef = (0.25, -1.3, 0.8, 2.1)
    coarse = np.zeros(g.ccells(1))
    for b in range(g.nb):
        X, Y, Z = centres(g.coarse, b)
        coarse[b, 0] = lin(coef, X[None, None, :], Y[None, :, None], Z[:, None, None])
    fine = np.full(g.cells(1), np.nan)
    items = [It(b, b, g.cis(0), g.cie(0), g.cis(1), g.cie(1), g.cis(2), g.cie(2)) for b in range(g.nb)]
    orc.ref_prolong_cc(G(g), 1, coarse, g.coarse, fine, g.fine, items)
    isl = np.s_[g.ng:g.ng + g.nx[2], g.ng:g.ng + g.nx[1], g.ng:g.ng + g.nx[0]]
    for b in range(g.nb):
        x, y, z = centres(g.fine, b)
        ref = lin(coef, x[None, None, :], y[None, :, None], z[:, None, None])[isl]
        np.testing.assert_allclose(fine[b, 0][isl], ref, rtol=0, atol=1e-13 * np.abs(ref).max())
    # only the prolonged region was written
    assert np.isnan(fine[:, 0, :g.ng]).all() and np.isnan(fine[:, 0, :, :, :g.ng]).all()


def test_prolong_cc_flat_at_extremum():
    g = Geo(nb=1, seed=4)
    rs = np.random.default_rng(5)
    coarse = rs.uniform(0.0, 1.0, g.ccells(1))
    k, j, i = g.cis(2) + 1, g.cis(1) + 2, g.cis(0) + 1
    coarse[0, 0, k, j, i] = 5.0           # local maximum in every direction
    fine = np.zeros(g.cells(1))
    orc.ref_prolong_cc(G(g), 1, coarse, g.coarse, fine, g.fine, [It(0, 0, i, i, j, j, k, k)])
    fk, fj, fi = (k - g.cng) * 2 + g.ng, (j - g.cng) * 2 + g.ng, (i - g.cng) * 2 + g.ng
    assert (fine[0, 0, fk:fk + 2, fj:fj + 2, fi:fi + 2] == 5.0).all()


def curl_faces(g, tab, b, A):
    """Face fields of B = curl A on a block's grid (A1 on x1-edges etc., given as
    functions of coordinates), as face-integrated fluxes divided by the face area,
    exactly divergence-free per cell up to round-off."""
    xf, yf, zf = tab["x1f"][b], tab["x2f"][b], tab["x3f"][b]
    xv, yv, zv = tab["x1v"][b], tab["x2v"][b], tab["x3v"][b]
    dx, dy, dz = tab["dx1f"][b], tab["dx2f"][b], tab["dx3f"][b]
    # edge-integrated potentials: A1 on (k face, j face, i centre) * dx, etc.
    E1 = A[0](xv[None, None, :], yf[None, :, None], zf[:, None, None]) * dx[None, None, :]
    E2 = A[1](xf[None, None, :], yv[None, :, None], zf[:, None, None]) * dy[None, :, None]
    E3 = A[2](xf[None, None, :], yf[None, :, None], zv[:, None, None]) * dz[:, None, None]
    b1 = ((E3[:, 1:, :] - E3[:, :-1, :]) - (E2[1:, :, :] - E2[:-1, :, :])) / (dy[None, :, None] * dz[:, None, None])
    b2 = ((E1[1:, :, :] - E1[:-1, :, :]) - (E3[:, :, 1:] - E3[:, :, :-1])) / (dx[None, None, :] * dz[:, None, None])
    b3 = ((E2[:, :, 1:] - E2[:, :, :-1]) - (E1[:, 1:, :] - E1[:, :-1, :])) / (dx[None, None, :] * dy[None, :, None])
    return b1, b2, b3


def test_face_prolongation_divergence_free_and_restricts_back():
    g = Geo(nb=2, seed=6)
    A = (lambda x, y, z: np.sin(2.0 * y + 0.3) * np.cos(1.5 * z),
         lambda x, y, z: np.cos(1.7 * x) * np.sin(2.2 * z + 0.1),
         lambda x, y, z: np.sin(1.9 * x + 0.5) * np.cos(2.4 * y))
    cb = [np.zeros(g.faces(d, coarse=True)) for d in range(3)]
    for b in range(g.nb):
        for d, arr in enumerate(curl_faces(g, g.coarse, b, A)):
            cb[d][b] = arr
    fb = [np.zeros(g.faces(d)) for d in range(3)]
    c0, c1 = [g.cis(d) for d in range(3)], [g.cie(d) for d in range(3)]
    for d in range(3):
        items = []
        for b in range(g.nb):
            hi = list(c1)
            hi[d] += 1    # faces: one more along the normal
            items.append(It(b, b, c0[0], hi[0], c0[1], hi[1], c0[2], hi[2]))
        orc.ref_prolong_fc_shared(G(g), d, cb[d], g.coarse, fb[d], g.fine, items)
    internal = [It(b, b, c0[0], c1[0], c0[1], c1[1], c0[2], c1[2]) for b in range(g.nb)]
    orc.ref_prolong_fc_internal(G(g), fb[0], fb[1], fb[2], g.fine, internal)
    s, e = g.ng, g.ng + g.nx[0]
    for b in range(g.nb):
        dx, dy, dz = g.fine["dx1f"][b], g.fine["dx2f"][b], g.fine["dx3f"][b]
        f1 = fb[0][b] * (dy[None, :, None] * dz[:, None, None])
        f2 = fb[1][b] * (dx[None, None, :] * dz[:, None, None])
        f3 = fb[2][b] * (dx[None, None, :] * dy[None, :, None])
        div = ((f1[s:e, s:e, s + 1:e + 1] - f1[s:e, s:e, s:e]) + (f2[s:e, s + 1:e + 1, s:e] - f2[s:e, s:e, s:e]) +
               (f3[s + 1:e + 1, s:e, s:e] - f3[s:e, s:e, s:e]))
        scale = np.abs(f1[s:e, s:e, s:e]).max()
        print(f"block {b}: max |div B| * V / max flux = {np.abs(div).max() / scale:.2e}", flush=True)
        assert np.abs(div).max() <= 1e-13 * scale
    # restricting the prolonged shared faces returns the coarse faces
    for d in range(3):
        hi = list(c1)
        hi[d] += 1
        back = np.zeros(g.faces(d, coarse=True))
        orc.ref_restrict_fc(G(g), d, fb[d], g.fine, back,
                            [It(b, b, c0[0], hi[0], c0[1], hi[1], c0[2], hi[2]) for b in range(g.nb)])
        sl = np.s_[:, c0[2]:hi[2] + 1, c0[1]:hi[1] + 1, c0[0]:hi[0] + 1]
        np.testing.assert_allclose(back[sl], cb[d][sl], rtol=0, atol=1e-14 * np.abs(cb[d][sl]).max() * 8)


@pytest.mark.parametrize("stretch", [0.0, 0.3])
def test_internal_faces_keep_fine_cells_divergence_free(stretch):
    """Outer faces of each 2x2x2 group from a divergence-free fine field (not from the
    separable shared-face prolongation, whose cross terms Uxyz, Vxyz, Wxyz vanish): the
    interior faces the operator computes leave every fine cell divergence-free."""
    g = Geo(nb=2, seed=8, stretch=stretch)
    A = (lambda x, y, z: np.sin(3.0 * y + 0.3) * np.cos(2.5 * z) * np.cos(2.0 * x),
         lambda x, y, z: np.cos(2.7 * x) * np.sin(3.2 * z + 0.1) * np.sin(1.3 * y),
         lambda x, y, z: np.sin(2.9 * x + 0.5) * np.cos(3.4 * y) * np.cos(1.1 * z))
    fb = [np.zeros(g.faces(d)) for d in range(3)]
    for b in range(g.nb):
        for d, arr in enumerate(curl_faces(g, g.fine, b, A)):
            fb[d][b] = arr
    orig = [f.copy() for f in fb]
    c0, c1 = [g.cis(d) for d in range(3)], [g.cie(d) for d in range(3)]
    orc.ref_prolong_fc_internal(G(g), fb[0], fb[1], fb[2], g.fine,
                                [It(b, b, c0[0], c1[0], c0[1], c1[1], c0[2], c1[2]) for b in range(g.nb)])
    s, e = g.ng, g.ng + g.nx[0]
    changed = 0
    for b in range(g.nb):
        dx, dy, dz = g.fine["dx1f"][b], g.fine["dx2f"][b], g.fine["dx3f"][b]
        f1 = fb[0][b] * (dy[None, :, None] * dz[:, None, None])
        f2 = fb[1][b] * (dx[None, None, :] * dz[:, None, None])
        f3 = fb[2][b] * (dx[None, None, :] * dy[None, :, None])
        div = ((f1[s:e, s:e, s + 1:e + 1] - f1[s:e, s:e, s:e]) + (f2[s:e, s + 1:e + 1, s:e] - f2[s:e, s:e, s:e]) +
               (f3[s + 1:e + 1, s:e, s:e] - f3[s:e, s:e, s:e]))
        scale = np.abs(f1[s:e, s:e, s:e]).max()
        print(f"block {b}: max |div B| * V / max flux = {np.abs(div).max() / scale:.2e}", flush=True)
        assert np.abs(div).max() <= 1e-13 * scale
        # the outer (even) faces are untouched, the interior (odd) ones were rewritten
        assert np.array_equal(fb[0][b][:, :, s:e + 1:2], orig[0][b][:, :, s:e + 1:2])
        changed += int((fb[0][b][s:e, s:e, s + 1:e:2] != orig[0][b][s:e, s:e, s + 1:e:2]).sum())
    assert changed > 0


def two_level_geo():
    """Block 0: a coarse block on [0,1]^3 (h = 1/8).  Block 1: a finer block (h = 1/16) on
    [1, 1.5] x [0.5, 1] x [0, 0.5], touching block 0's outer x1 face in the quadrant
    fi1 = 1 (upper half in x2), fi2 = 0 (lower half in x3)."""
    h = [(1 / 8,) * 3, (1 / 16,) * 3]
    return Geo(nb=2, h=h, origin=[(0.0, 0.0, 0.0), (1.0, 0.5, 0.0)])


def test_flux_from_finer_places_linear_flux():
    g = two_level_geo()
    nvar, coef = 2, [(0.7, 0.0, 1.3, -0.4), (-0.2, 0.0, 0.5, 2.0)]
    ff = [np.zeros(g.fluxes(d, nvar)) for d in range(3)]
    x, y, z = g.fine["x1f"][1], g.fine["x2v"][1], g.fine["x3v"][1]
    for n in range(nvar):
        ff[0][1, n] = lin(coef[n], x[None, None, :], y[None, :, None], z[:, None, None])
    cf = [np.full(g.fluxes(d, nvar), -99.0) for d in range(3)]
    orc.ref_flux_from_finer(G(g), nvar, ff[0], ff[1], ff[2], g.fine, cf[0], cf[1], cf[2],
                            [It(1, 0, fid=1, fi1=1, fi2=0)])
    ie1, js, je, ks = g.ie(0) + 1, g.ng, g.ie(1), g.ng
    h2, h3 = g.nx[1] // 2, g.nx[2] // 2
    Y, Z = g.fine["x2v"][0], g.fine["x3v"][0]    # block 0 face centres (its own grid)
    for n in range(nvar):
        got = cf[0][0, n, ks:ks + h3, js + h2:je + 1, ie1]
        ref = lin(coef[n], 1.0, Y[None, js + h2:je + 1], Z[ks:ks + h3, None])
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-14 * 4)
    written = cf[0][0, :, ks:ks + h3, js + h2:je + 1, ie1].size
    assert (cf[0] != -99.0).sum() == written and (cf[1] == -99.0).all() and (cf[2] == -99.0).all()


@pytest.mark.parametrize("fid", range(6))
def test_emf_to_coarser_buffer_order(fid):
    g = Geo(nb=1, seed=7, stretch=0.2)
    x1f, x2f, x3f = g.fine["x1f"][0], g.fine["x2f"][0], g.fine["x3f"][0]
    x1v, x2v, x3v = g.fine["x1v"][0], g.fine["x2v"][0], g.fine["x3v"][0]
    c1, c2, c3 = (0.3, 1.1, -0.6, 0.9), (-0.5, 0.4, 1.7, -1.2), (1.0, -0.8, 0.3, 0.6)
    e1 = lin(c1, x1v[None, None, :], x2f[None, :, None], x3f[:, None, None])[None].copy()
    e2 = lin(c2, x1f[None, None, :], x2v[None, :, None], x3f[:, None, None])[None].copy()
    e3 = lin(c3, x1f[None, None, :], x2f[None, :, None], x3v[:, None, None])[None].copy()
    stride = 200
    buf = np.full(stride, np.nan)
    orc.ref_emf_to_coarser(G(g), e1, e2, e3, g.fine, buf, stride, [It(0, fid=fid)])
    dir_, side = fid >> 1, fid & 1
    s = g.ng
    e = [g.ie(d) for d in range(3)]
    face = [s + side * (g.nx[d]) for d in range(3)]

    def mid(xv, dxf, i):   # length-weighted mean of two fine edge centres
        return (xv[i] * dxf[i] + xv[i + 1] * dxf[i + 1]) / (dxf[i] + dxf[i + 1])
    d1, d2, d3 = g.fine["dx1f"][0], g.fine["dx2f"][0], g.fine["dx3f"][0]
    exp = []
    if dir_ == 0:
        X = x1f[face[0]]
        exp += [lin(c2, X, mid(x2v, d2, j), x3f[k]) for k in range(s, e[2] + 2, 2) for j in range(s, e[1] + 1, 2)]
        exp += [lin(c3, X, x2f[j], mid(x3v, d3, k)) for k in range(s, e[2] + 1, 2) for j in range(s, e[1] + 2, 2)]
    elif dir_ == 1:
        Y = x2f[face[1]]
        exp += [lin(c1, mid(x1v, d1, i), Y, x3f[k]) for k in range(s, e[2] + 2, 2) for i in range(s, e[0] + 1, 2)]
        exp += [lin(c3, x1f[i], Y, mid(x3v, d3, k)) for k in range(s, e[2] + 1, 2) for i in range(s, e[0] + 2, 2)]
    else:
        Z = x3f[face[2]]
        exp += [lin(c1, mid(x1v, d1, i), x2f[j], Z) for j in range(s, e[1] + 2, 2) for i in range(s, e[0] + 1, 2)]
        exp += [lin(c2, x1f[i], mid(x2v, d2, j), Z) for j in range(s, e[1] + 1, 2) for i in range(s, e[0] + 2, 2)]
    exp = np.array(exp)
    from tests.conftest import load_agx
    agx = load_agx()
    assert len(exp) == agx.ref_emf_count(agx.RefGeom(g.nx[0], g.nx[1], g.nx[2], g.ng, g.cng, 0), fid)
    np.testing.assert_allclose(buf[:len(exp)], exp, rtol=0, atol=1e-13)
    assert np.isnan(buf[len(exp):]).all()


# ---------------------------------------------------------------- ghost prolongation
def test_prim_to_cons_inverts_cons2prim():
    """orc_ref_prim_to_cons (PrimitiveToConserved + uniform CalculateCellCenteredField)
    followed by the oracle's ConservedToPrimitive returns the primitives."""
    from oracle.orc import Oracle, make_desc
    g = Geo(nx=(8, 8, 8), ng=2, nb=1, seed=21)
    desc = make_desc(nx=(8, 8, 8), bnx=(8, 8, 8), box=((0.0, 1.0),) * 3, gamma=5.0 / 3.0)
    rs = np.random.default_rng(22)
    w = np.zeros(g.cells(5))
    w[:, 0] = rs.uniform(0.5, 2.0, w[:, 0].shape)
    w[:, 1:4] = rs.normal(size=w[:, 1:4].shape)
    w[:, 4] = rs.uniform(0.5, 2.0, w[:, 4].shape)
    A = (lambda x, y, z: np.sin(6.0 * y) * np.cos(5.0 * z), lambda x, y, z: np.cos(7.0 * x) * np.sin(4.0 * z),
         lambda x, y, z: np.sin(5.0 * x) * np.cos(6.0 * y))
    faces = [f[None].copy() for f in curl_faces(g, g.fine, 0, A)]
    bcc = np.zeros(g.cells(3))
    u = np.zeros(g.cells(5))
    c0, c1 = g.cis(0), g.cie(0)
    orc.ref_prim_to_cons(G(g), w, faces[0], faces[1], faces[2], bcc, u, 5.0 / 3.0, 0, [It(0, 0, c0, c1, c0, c1, c0, c1)])
    o = Oracle(desc)
    o.set("u", u)
    for n, f in zip(("b1f", "b2f", "b3f"), faces):
        o.set(n, f)
    o.L.orc_cons2prim(o.h)
    s = np.s_[:, :, g.ng:g.ng + 8, g.ng:g.ng + 8, g.ng:g.ng + 8]
    np.testing.assert_allclose(o.get("w")[s], w[s], rtol=1e-13, atol=1e-14)
    np.testing.assert_array_equal(o.get("bcc")[s], bcc[s])


def nblevel_cases(ox, rng):
    """Neighbourhoods with the neighbour at ox coarser: all coarser, and random mixes of
    coarser / same / finer elsewhere."""
    idx = ((ox[2] + 1) * 3 + ox[1] + 1) * 3 + ox[0] + 1
    cases = [[-1] * 27]
    for _ in range(3):
        lv = list(rng.integers(-1, 2, 27))
        lv[13] = 0
        lv[idx] = -1
        cases.append(lv)
    return cases


@pytest.mark.parametrize("ng", [2, 4])
def test_prolong_ranges_cover_the_ghost_slab(ng):
    """For every coarser-neighbour offset and block parity, the fine cells of the ranges
    stay inside the block array and cover the ghost slab facing that neighbour: along
    ox = -1 exactly the ng ghosts below, along +1 the ng above, along 0 the interior plus
    ng ghosts on the side the coarser neighbour extends to (the block's parity).  A face
    range drops its lower (upper) face exactly when the neighbour across it is at the
    same or a finer level.  The product's host function gives the same ranges."""
    import itertools
    from tests.conftest import load_agx
    agx = load_agx()
    g = Geo(nx=(8, 8, 8), ng=ng, nb=1)
    go, ga = G(g), agx.RefGeom(8, 8, 8, ng, ng, 0)
    rng = np.random.default_rng(23)
    is_, ie = g.ng, g.ng + 7
    for ox in itertools.product((-1, 0, 1), repeat=3):
        if ox == (0, 0, 0):
            continue
        for par in itertools.product((0, 1), repeat=3):
            for lv in nblevel_cases(ox, rng):
                r = orc.ref_prolong_ranges(go, ox, par, lv)
                assert r == agx.ref_prolong_ranges(ga, ox, par, lv)
                for d in range(3):
                    s, e = r[2 * d], r[2 * d + 1]
                    fs, fe = (s - g.cng) * 2 + g.ng, (e - g.cng) * 2 + g.ng + 1
                    assert 0 <= fs <= fe <= g.n[d] - 1
                    if ox[d] < 0:
                        assert (fs, fe) == (is_ - ng, is_ - 1)
                    elif ox[d] > 0:
                        assert (fs, fe) == (ie + 1, ie + ng)
                    else:
                        assert (fs, fe) == ((is_, ie + ng) if par[d] == 0 else (is_ - ng, ie))
                    lo_nb = [o + 1 for o in ox]
                    hi_nb = [o + 1 for o in ox]
                    lo_nb[d] = ox[d]
                    hi_nb[d] = ox[d] + 2
                    lev = lambda a: lv[(a[2] * 3 + a[1]) * 3 + a[0]]
                    assert r[6 + 2 * d] == s + (1 if ox[d] >= 0 and lev(lo_nb) >= 0 else 0)
                    assert r[7 + 2 * d] == e + 1 - (1 if ox[d] <= 0 and lev(hi_nb) >= 0 else 0)


def ghost_pipeline(g, ox, par, lv, coarse_w, coarse_b, fine_w, fine_b, gamma=5.0 / 3.0):
    """ProlongateGhostCells for one coarser neighbour, in the oracle: cell prolongation of
    w, shared faces over the face ranges, interior faces, then bcc and u."""
    r = orc.ref_prolong_ranges(G(g), ox, par, lv)
    cc = It(0, 0, r[0], r[1], r[2], r[3], r[4], r[5])
    orc.ref_prolong_cc(G(g), 5, coarse_w, g.coarse, fine_w, g.fine, [cc])
    fr = [It(0, 0, r[6], r[7], r[2], r[3], r[4], r[5]), It(0, 0, r[0], r[1], r[8], r[9], r[4], r[5]),
          It(0, 0, r[0], r[1], r[2], r[3], r[10], r[11])]
    for d in range(3):
        orc.ref_prolong_fc_shared(G(g), d, coarse_b[d], g.coarse, fine_b[d], g.fine, [fr[d]])
    orc.ref_prolong_fc_internal(G(g), fine_b[0], fine_b[1], fine_b[2], g.fine, [cc])
    bcc = np.zeros(g.cells(3))
    u = np.zeros(g.cells(5))
    orc.ref_prim_to_cons(G(g), fine_w, fine_b[0], fine_b[1], fine_b[2], bcc, u, gamma, 0, [cc])
    return r, bcc, u


@pytest.mark.parametrize("ox", [(-1, 0, 0), (0, 1, 0), (1, -1, 0), (-1, 1, 1)])
def test_ghost_prolongation_keeps_divb_zero(ox):
    """With every neighbour coarser, the prolonged ghost cells of a divergence-free coarse
    field are divergence-free, and a linear primitive field is reproduced there."""
    g = Geo(nx=(8, 8, 8), ng=2, nb=1, seed=24)
    A = (lambda x, y, z: np.sin(2.0 * y + 0.3) * np.cos(1.5 * z), lambda x, y, z: np.cos(1.7 * x) * np.sin(2.2 * z),
         lambda x, y, z: np.sin(1.9 * x + 0.5) * np.cos(2.4 * y))
    cb = [f[None].copy() for f in curl_faces(g, g.coarse, 0, A)]
    coef = [(1.0, 0.1, 0.2, -0.1), (0.3, 0.2, 0.0, 0.1), (0.0, -0.2, 0.1, 0.3), (0.1, 0.0, 0.0, 0.2),
            (1.5, 0.1, -0.1, 0.2)]
    X, Y, Z = centres(g.coarse, 0)
    cw = np.zeros(g.ccells(5))
    for n in range(5):
        cw[0, n] = lin(coef[n], X[None, None, :], Y[None, :, None], Z[:, None, None])
    fw = np.zeros(g.cells(5))
    fb = [np.zeros(g.faces(d)) for d in range(3)]
    r, bcc, u = ghost_pipeline(g, ox, (0, 1, 0), [-1] * 27, cw, cb, fw, fb)
    sl = [slice((r[2 * d] - g.cng) * 2 + g.ng, (r[2 * d + 1] - g.cng) * 2 + g.ng + 2) for d in range(3)]
    dx, dy, dz = g.fine["dx1f"][0], g.fine["dx2f"][0], g.fine["dx3f"][0]
    f1 = fb[0][0] * (dy[None, :, None] * dz[:, None, None])
    f2 = fb[1][0] * (dx[None, None, :] * dz[:, None, None])
    f3 = fb[2][0] * (dx[None, None, :] * dy[None, :, None])
    k, j, i = sl[2], sl[1], sl[0]
    ip = slice(i.start + 1, i.stop + 1)
    jp = slice(j.start + 1, j.stop + 1)
    kp = slice(k.start + 1, k.stop + 1)
    div = (f1[k, j, ip] - f1[k, j, i]) + (f2[k, jp, i] - f2[k, j, i]) + (f3[kp, j, i] - f3[k, j, i])
    assert np.abs(div).max() <= 1e-14 * np.abs(f1[k, j, i]).max()
    x, y, z = centres(g.fine, 0)
    for n in range(5):
        ref = lin(coef[n], x[None, None, :], y[None, :, None], z[:, None, None])[k, j, i]
        np.testing.assert_allclose(fw[0, n][k, j, i], ref, rtol=0, atol=1e-13)
    assert (u[0, 0][k, j, i] == fw[0, 0][k, j, i]).all()
