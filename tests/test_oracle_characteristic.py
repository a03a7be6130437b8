"""Pins of the oracle's characteristic projection (reconstruct/characteristic.cpp; the
"2c" and "3c" reconstructions of plm.cpp and ppm.cpp).

What fixes them, independently of the formulas restated in the oracle:
- the columns of R are right eigenvectors of the one-dimensional primitive-variable MHD
  Jacobian (written out below from the equations of ideal MHD), with eigenvalues
  vn -+ cf, vn -+ ca, vn -+ cs, vn in the order fast-, Alfven-, slow-, entropy, slow+,
  Alfven+, fast+, and L = R^-1 (the isothermal system has no entropy wave);
- a reconstruction limited in characteristic variables equals one built from numpy's own
  eigen-decomposition of that Jacobian (PLM's van Leer slope and PPM's limiter are odd and
  homogeneous of degree one in the data, so the eigenvectors' scaling drops out);
- a uniform state reconstructs to itself and linear data to the midpoint values.
"""
import numpy as np
import pytest

from oracle import orc
from oracle.orc import Oracle, make_desc

GAMMA = 5.0 / 3.0


def local_to_mesh(dir):
    """slot in the 7-slot mesh-frame vector of each local variable (rho, vn, vt1, vt2, P,
    Bt1, Bt2)"""
    return [0, 1 + dir, 1 + (dir + 1) % 3, 1 + (dir + 2) % 3, 4, 5, 6]


def jacobian(w, bn, dir, iso=False, cs=1.0):
    """A with W_t + A W_x = 0 for W in the mesh-frame slots, from the equations of ideal MHD
    (Heaviside-Lorentz units, B_n constant along the sweep)."""
    p = local_to_mesh(dir)
    rho, vn, p_ = w[0], w[p[1]], w[4]
    bt1, bt2 = w[5], w[6]
    a = np.zeros((7, 7))

    def put(r, c, v):
        a[p[r], p[c]] = v
    for r in range(7):
        put(r, r, vn)
    put(0, 1, rho)                                   # continuity
    put(1, 4, 1.0 / rho)                             # normal momentum: grad P + grad B_t^2/2
    put(1, 5, bt1 / rho)
    put(1, 6, bt2 / rho)
    put(2, 5, -bn / rho)                             # transverse momentum: -B_n dB_t/dx
    put(3, 6, -bn / rho)
    put(4, 1, GAMMA * p_)                            # pressure
    put(5, 1, bt1)                                   # induction
    put(5, 2, -bn)
    put(6, 1, bt2)
    put(6, 3, -bn)
    if iso:
        put(1, 0, cs * cs / rho)
        a[4, :] = 0.0
        a[:, 4] = 0.0
    return a


def speeds(w, bn, iso=False, cs=1.0):
    rho = w[0]
    a2 = cs * cs if iso else GAMMA * w[4] / rho
    b2 = (bn * bn + w[5] ** 2 + w[6] ** 2) / rho
    ca2 = bn * bn / rho
    disc = np.sqrt((a2 + b2) ** 2 - 4.0 * a2 * ca2)
    return np.sqrt(0.5 * (a2 + b2 + disc)), np.sqrt(ca2), np.sqrt(0.5 * (a2 + b2 - disc))


def random_state(rng, dir):
    w = np.zeros(7)
    w[0] = rng.uniform(0.5, 2.0)
    w[1:4] = rng.uniform(-1.0, 1.0, 3)
    w[4] = rng.uniform(0.3, 2.0)
    w[5:7] = rng.uniform(-1.5, 1.5, 2)
    bn = rng.uniform(0.2, 1.5) * rng.choice([-1.0, 1.0])
    return w, bn


@pytest.mark.parametrize("iso", [False, True])
@pytest.mark.parametrize("dir", [0, 1, 2])
def test_right_eigenvectors_of_the_jacobian_and_inverse(dir, iso):
    rng = np.random.default_rng(10 + dir + 3 * iso)
    for _ in range(20):
        w, bn = random_state(rng, dir)
        if iso:
            w[4] = 0.0
        L, R = orc.char_matrices(w, bn, dir=dir, gamma=GAMMA, iso=iso, iso_cs=0.9)
        A = jacobian(w, bn, dir, iso=iso, cs=0.9)
        cf, ca, cs = speeds(w, bn, iso=iso, cs=0.9)
        vn = w[1 + dir]
        lam = [vn - cf, vn - ca, vn - cs, vn, vn + cs, vn + ca, vn + cf]
        keep = list(range(7))
        if iso:
            lam = lam[:3] + lam[4:]
            keep = [0, 1, 2, 3, 5, 6]
        nw = len(lam)
        for r in range(nw):
            col = R[:, r]
            assert np.abs(col).max() > 0.0
            np.testing.assert_allclose(A @ col, lam[r] * col, rtol=0, atol=1e-12 * np.abs(col).max() * 4,
                                       err_msg=f"wave {r}")
        Lk, Rk = L[:nw][:, keep], R[keep][:, :nw]
        np.testing.assert_allclose(Lk @ Rk, np.eye(nw), atol=1e-12)
        np.testing.assert_allclose(Rk @ Lk, np.eye(nw), atol=1e-12)
        if iso:
            assert np.all(R[4] == 0.0) and np.all(L[:, 4] == 0.0)


def test_no_transverse_field_drops_the_transverse_families():
    """B_t = 0 with c_A < a: the transverse unit vector is taken as zero
    (characteristic.cpp:73-78) and the slow speed equals c_A, so alpha_s = 0 (:81-93): the
    Alfven and slow fields are zero and only the fast pair and the entropy field remain,
    inverse to each other."""
    w = np.array([1.2, 0.3, -0.2, 0.1, 0.8, 0.0, 0.0])
    L, R = orc.char_matrices(w, 0.7, dir=0, gamma=GAMMA)
    for r in (1, 2, 4, 5):
        assert np.all(L[r] == 0.0) and np.all(R[:, r] == 0.0)
    k = [0, 3, 6]
    np.testing.assert_allclose(L[k] @ R[:, k], np.eye(3), atol=1e-12)
    A = jacobian(w, 0.7, 0)
    cf, _, _ = speeds(w, 0.7)
    for r, lam in zip(k, (0.3 - cf, 0.3, 0.3 + cf)):
        np.testing.assert_allclose(A @ R[:, r], lam * R[:, r], atol=1e-12)


# ---------------------------------------------------------------- whole reconstructions
def mesh_with(fields, order, iso=False, char_proj=1):
    """one 8^3 MeshBlock (nghost 3) whose w and bcc are the given callables of (k, j, i)"""
    desc = make_desc(nx=(8, 8, 8), box=((0.0, 1.0),) * 3, nghost=3, xorder=order, char_proj=char_proj,
                     eos=1 if iso else 0, iso_cs=0.9, gamma=GAMMA)
    o = Oracle(desc)
    w = o.get("w")
    b = o.get("bcc")
    n3, n2, n1 = w.shape[2:]
    kk, jj, ii = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
    wv, bv = fields(kk, jj, ii)
    w[0] = wv
    b[0] = bv
    if iso:
        w[0, 4] = 0.0
    o.set("w", w)
    o.set("bcc", b)
    return o, w[0], b[0]


def stencil(w, b, dir, k, j, i, t):
    """7-slot mesh-frame vector at offset t along dir from (k, j, i)"""
    o = [0, 0, 0]
    o[2 - dir] = t
    kk, jj, ii = k + o[0], j + o[1], i + o[2]
    v = np.zeros(7)
    v[:5] = w[:, kk, jj, ii]
    v[5] = b[(dir + 1) % 3, kk, jj, ii]
    v[6] = b[(dir + 2) % 3, kk, jj, ii]
    return v, b[dir, kk, jj, ii]


def numpy_char_recon(w, b, dir, k, j, i, order, iso=False):
    """reconstruction through numpy's eigen-decomposition of the Jacobian at the cell"""
    wc, bn = stencil(w, b, dir, k, j, i, 0)
    keep = [0, 1, 2, 3, 5, 6] if iso else list(range(7))
    A = jacobian(wc, bn, dir, iso=iso, cs=0.9)[np.ix_(keep, keep)]
    ev, Rn = np.linalg.eig(A)
    assert np.abs(ev.imag).max() < 1e-12
    Rn = Rn.real
    Ln = np.linalg.inv(Rn)
    pts = range(-2, 3) if order == 3 else range(-1, 2)
    ch = np.array([Ln @ stencil(w, b, dir, k, j, i, t)[0][keep] for t in pts])   # (npts, nw)
    cm, cp = np.zeros(len(keep)), np.zeros(len(keep))
    for r in range(len(keep)):
        if order == 3:
            wl, wr = orc.ppm_pencil(ch[:, r])
            cm[r], cp[r] = wr[2], wl[3]
        else:
            wl, wr = orc.plm_pencil(ch[:, r])
            cm[r], cp[r] = wr[1], wl[2]
    wm, wp = np.zeros(7), np.zeros(7)
    wm[keep], wp[keep] = Rn @ cm, Rn @ cp
    if order == 2:   # the limited slope is added to the cell's primitive state
        wm[keep] = wc[keep] - 0.5 * (Rn @ (cp - cm))
        wp[keep] = wc[keep] + 0.5 * (Rn @ (cp - cm))
    return wm, wp


def smooth_fields(seed):
    rng = np.random.default_rng(seed)
    kx = rng.uniform(0.2, 0.6, (8, 3))
    ph = rng.uniform(0, 2 * np.pi, 8)
    base = [1.0, 0.2, -0.1, 0.15, 0.9]
    amp = [0.3, 0.4, 0.4, 0.4, 0.3]

    def f(kk, jj, ii):
        def mode(m):
            return np.sin(kx[m, 0] * ii + kx[m, 1] * jj + kx[m, 2] * kk + ph[m])
        wv = np.array([base[n] + amp[n] * mode(n) for n in range(5)])
        bv = np.array([0.6 + 0.4 * mode(5), -0.5 + 0.4 * mode(6), 0.4 + 0.4 * mode(7)])
        return wv, bv
    return f


@pytest.mark.parametrize("iso", [False, True])
@pytest.mark.parametrize("order", [2, 3])
def test_reconstruction_equals_numpy_eigen_decomposition(order, iso):
    o, w, b = mesh_with(smooth_fields(3 + order + 2 * iso), order, iso=iso)
    plain, _, _ = mesh_with(smooth_fields(3 + order + 2 * iso), order, iso=iso, char_proj=0)
    worst, apart = 0.0, 0.0
    for dir in range(3):
        for (k, j, i) in [(5, 5, 5), (4, 6, 3), (7, 3, 6), (3, 4, 8), (8, 8, 8), (6, 9, 4)]:
            wm, wp = o.recon_cell(dir, order, 0, k, j, i)
            em, ep = numpy_char_recon(w, b, dir, k, j, i, order, iso=iso)
            worst = max(worst, np.abs(wm - em).max(), np.abs(wp - ep).max())
            pm, pp = plain.recon_cell(dir, order, 0, k, j, i)
            apart = max(apart, np.abs(wm - pm).max(), np.abs(wp - pp).max())
    print(f"order {order} iso {iso}: max |oracle - numpy eigen| = {worst:.3e}; "
          f"characteristic vs primitive limiting differ by up to {apart:.3e}")
    assert worst < 1e-11
    assert apart > 1e-3   # the limiters act on this data, so the projection matters


@pytest.mark.parametrize("order", [2, 3])
def test_uniform_state_and_linear_data(order):
    wc = np.array([1.3, 0.2, -0.4, 0.1, 0.7])
    bc = np.array([0.5, -0.8, 0.3])
    o, w, b = mesh_with(lambda kk, jj, ii: (wc[:, None, None, None] + 0 * ii, bc[:, None, None, None] + 0 * ii),
                        order)
    for dir in range(3):
        wm, wp = o.recon_cell(dir, order, 0, 5, 5, 5)
        v, _ = stencil(w, b, dir, 5, 5, 5, 0)
        np.testing.assert_allclose(wm, v, rtol=1e-14, atol=1e-15)
        np.testing.assert_allclose(wp, v, rtol=1e-14, atol=1e-15)
    slope = np.array([0.01, 0.02, -0.015, 0.01, 0.012, 0.008, -0.01, 0.011])

    def lin(kk, jj, ii):
        s = ii + 0.7 * jj - 0.4 * kk
        return wc[:, None, None, None] + slope[:5, None, None, None] * s, \
            bc[:, None, None, None] + slope[5:, None, None, None] * s
    o, w, b = mesh_with(lin, order)
    for dir in range(3):
        wm, wp = o.recon_cell(dir, order, 0, 5, 5, 5)
        v0, _ = stencil(w, b, dir, 5, 5, 5, 0)
        vm, _ = stencil(w, b, dir, 5, 5, 5, -1)
        vp, _ = stencil(w, b, dir, 5, 5, 5, 1)
        np.testing.assert_allclose(wm, 0.5 * (v0 + vm), rtol=0, atol=1e-13)
        np.testing.assert_allclose(wp, 0.5 * (v0 + vp), rtol=0, atol=1e-13)
