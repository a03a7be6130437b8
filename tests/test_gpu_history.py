"""GPU history sums (agx_history, agx_history_sum) against the oracle (orc_history).

The device adds the same per-cell terms in a fixed tree order, the oracle (like the
reference) sequentially, so the bar is rounding of the sum order: |gpu - oracle| <=
1e-13 x (sum of the terms' magnitudes) per column.  Mesh and dense-array entry points
run the same kernels on the same cells, so they agree bit for bit.  At a larger size the
periodic run conserves mass and the momenta (property, no oracle).
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc

pytestmark = pytest.mark.gpu


def random_state(d, seed):
    rs = np.random.default_rng(seed)
    nb = (d.nx1 // d.bnx1) * (d.nx2 // d.bnx2) * (d.nx3 // d.bnx3)
    n = (d.bnx3 + 2 * d.nghost, d.bnx2 + 2 * d.nghost, d.bnx1 + 2 * d.nghost)
    u = rs.standard_normal((nb, 5) + n)
    u[:, 0] = 1.0 + rs.random((nb,) + n)
    u[:, 4] = 3.0 + rs.random((nb,) + n)
    bcc = rs.standard_normal((nb, 3) + n)
    phi = rs.standard_normal((nb,) + n)
    return u, bcc, phi


def scale(d, u, bcc, phi):
    """Sum of |term| per column: bounds the rounding of any summation order."""
    ng = d.nghost
    c = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    vol = (d.x1max - d.x1min) / d.nx1 * (d.x2max - d.x2min) / d.nx2 * (d.x3max - d.x3min) / d.nx3
    ui = [np.abs(u[(slice(None), n) + c[1:]]) for n in range(5)]
    s = [ui[0].sum()] + [x.sum() for x in ui[1:4]] + [(0.5 * x * x / ui[0]).sum() for x in ui[1:4]]
    s += [ui[4].sum(), (0.5 * ui[0] * np.abs(phi[c])).sum()]
    s += [(0.5 * bcc[(slice(None), n) + c[1:]] ** 2).sum() for n in range(3)]
    cols = list(range(7)) + ([7] if d.eos == 0 else []) + ([8] if d.self_gravity else []) + [9, 10, 11]
    return np.array(s)[cols] * vol


@pytest.mark.parametrize("precision,eos,sg", [(0, 0, 1), (1, 0, 1), (0, 1, 0), (1, 1, 1)])
def test_history_matches_oracle(agx, precision, eos, sg):
    d = make_desc(nx=(32, 16, 16), bnx=(8, 8, 8), box=((0.0, 2.0), (0.0, 1.0), (0.0, 1.0)), self_gravity=sg,
                  eos=eos, precision=precision)
    u, bcc, phi = random_state(d, 11 + precision + 2 * eos)
    if eos == 1:
        u[:, 4] = 0.0
    m = agx.Mesh(d)
    m.set("u", u[:, :4] if eos == 1 else u)   # the isothermal C ABI carries NHYDRO = 4
    m.set("bcc", bcc)
    if sg:
        m.set("phi", phi)
    bcc_dev = m.get("bcc")          # FP32-rounded in the mixed form
    o = Oracle(d)
    o.set("u", u)
    o.set("bcc", bcc_dev)
    if sg:
        o.set("phi", phi[:, None])
    h, ho = m.history(), o.history()
    assert h.shape == ho.shape
    sc = scale(d, u, bcc_dev, phi)
    err = np.abs(h - ho) / sc
    print(f"precision {precision} eos {eos} sg {sg}: max |gpu - oracle| / sum|terms| = {err.max():.2e}", flush=True)
    assert (err <= 1e-13).all()
    m.close()


def test_dense_arrays_equal_mesh_bitwise(agx):
    d = make_desc(nx=(32, 32, 16), bnx=(8, 8, 8), box=((0.0, 1.0), (0.0, 1.0), (0.0, 0.5)), self_gravity=1)
    u, bcc, phi = random_state(d, 3)
    m = agx.Mesh(d)
    m.set("u", u)
    m.set("bcc", bcc)
    m.set("phi", phi)
    h = m.history()
    nb = u.shape[0]
    vol = np.full(nb, (1.0 / 32) * (1.0 / 32) * (1.0 / 32))
    D = [agx.DeviceArray(x) for x in (u, bcc, phi[:, None], vol, np.zeros(16))]
    L = agx.lib()
    rc = L.agx_history_sum(D[0].ptr, D[1].ptr, D[2].ptr, nb, 5, 8, 8, 8, d.nghost, D[3].ptr, D[4].ptr, None)
    assert rc == 0, L.agx_last_error()
    hd = D[4].get()[:len(h)]
    print("mesh", h, "\ndense", hd, flush=True)
    assert np.array_equal(h, hd)
    # without phi: the gravitational column is left out, the others are unchanged
    rc = L.agx_history_sum(D[0].ptr, D[1].ptr, None, nb, 5, 8, 8, 8, d.nghost, D[3].ptr, D[4].ptr, None)
    assert rc == 0
    hn = D[4].get()[:11]
    assert np.array_equal(hn[:8], h[:8]) and np.array_equal(hn[8:], h[9:])
    # a contract violation is reported, not run
    assert L.agx_history_sum(D[0].ptr, D[1].ptr, None, nb, 6, 8, 8, 8, d.nghost, D[3].ptr, D[4].ptr, None) < 0
    for x in D:
        x.close()
    m.close()


def test_periodic_run_conserves_mass_and_momentum(agx):
    from tests.conftest import load_pgen
    pgen = load_pgen()
    d = agx.make_desc(nx=(64, 64, 64), bnx=(32, 32, 32), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0)
    U, b1, b2, b3 = pgen.turbulence(d, seed=2, mach=2.0, bx=0.1)
    m = agx.Mesh(d)
    m.set("u", agx.scatter_cells(d, U))
    for n, f in zip(("b1f", "b2f", "b3f"), pgen.scatter_faces(d, b1, b2, b3)):
        m.set(n, f)
    m.initialize()
    h0 = m.history()
    assert m.cycle(1e30, 4) == 4
    h1 = m.history()
    print("before", h0, "\nafter ", h1, flush=True)
    assert abs(h1[0] - h0[0]) <= 1e-13 * h0[0]
    for n in range(1, 4):   # momenta: conserved to round-off of the mass-weighted sums
        assert abs(h1[n] - h0[n]) <= 1e-13 * h0[0] * 2.0
    assert abs(h1[7] - h0[7]) <= 1e-12 * h0[7]   # total energy (adiabatic, no source)
    m.close()
