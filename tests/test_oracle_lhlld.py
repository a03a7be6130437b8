"""Pins of the oracle's low-dissipation HLLD (lhlld.cpp; Minoshima, Kitamura & Miyoshi
2021) and of the velocity differences that drive its shock detector
(calculate_velocity_differences.cpp).

What the mathematics fixes, none of it a retyped formula of the oracle:
- consistency F(U,U) = F_phys(U) for any detector values (the low-Mach terms multiply
  differences of the two states);
- an isolated stationary contact is resolved exactly (zero mass flux);
- mirror symmetry x -> -x for random states and detector values;
- with the detector off (theta = 1) and the low-Mach factor saturated (phi = 1, chi >= 1
  for flows whose kinetic or Alfvenic speed reaches the fast speed), LHLLD is the HLLD
  solver: the star pressure and S_M reduce to Miyoshi & Kusano's, so the two agree to
  rounding on random supersonic states (HLLD is pinned separately);
- theta scales the pressure-jump term of S_M only: across a small pure pressure jump
  (no flow, no field) the mass flux is proportional to theta, theta = th1^4 with th1 =
  (c_f - min(dvn, 0)) / (c_f - min(dvt, 0)) capped at 1 (the paper's eq. 31-32);
- phi = chi (2 - chi) scales the velocity-jump term of the star pressure: across a
  small pure velocity jump the momentum flux's excess over the ambient pressure is
  phi times HLLD's;
- the full scheme with LHLLD converges at second order on the linear fast wave and is no
  more diffusive than HLLD there;
- velocity differences: exact on linear velocity fields (dvn = the normal jump, dvt =
  the smallest transverse one-sided difference) for every direction.
"""
import numpy as np
import pytest

from oracle import orc as O
from oracle.orc import Oracle, make_desc
from tests.test_oracle_riemann_plm import GAMMA, STATES, mirror, phys_flux


def rand_state(rng, sup=False):
    d = rng.uniform(0.3, 2.0)
    v = rng.normal(size=3) * (3.0 if sup else 0.5)
    return np.array([d, *v, rng.uniform(0.2, 2.0), *rng.normal(size=2)])


@pytest.mark.parametrize("k", range(len(STATES)))
@pytest.mark.parametrize("dv", [(0.0, 0.0), (-0.7, 0.2), (0.3, -1.1)])
def test_consistency(k, dv):
    w, bx = STATES[k]
    f = O.riemann_lhlld(GAMMA, w, w, bx, *dv)
    np.testing.assert_allclose(f, phys_flux(w, bx), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("dv", [(0.0, 0.0), (-0.4, -0.9)])
def test_stationary_contact_exact(dv):
    wl = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    wr = np.array([0.125, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    f = O.riemann_lhlld(GAMMA, wl, wr, 0.75, *dv)
    assert abs(f[0]) < 1e-15
    np.testing.assert_allclose(f[1:], phys_flux(wl, 0.75)[1:], rtol=1e-13, atol=1e-14)


def test_mirror_symmetry():
    rng = np.random.default_rng(17)
    sign = np.array([-1, 1, -1, -1, -1, -1, -1])
    for _ in range(100):
        wl, wr, bx = rand_state(rng), rand_state(rng), rng.normal()
        dvn, dvt = rng.normal(scale=0.5), rng.normal(scale=0.5)
        f = O.riemann_lhlld(GAMMA, wl, wr, bx, dvn, dvt)
        g = O.riemann_lhlld(GAMMA, mirror(wr), mirror(wl), -bx, dvn, dvt)
        np.testing.assert_allclose(g, sign * f, rtol=1e-11, atol=1e-12)


def test_equals_hlld_when_detector_off_and_phi_saturated():
    rng = np.random.default_rng(3)
    n = 0
    for _ in range(400):
        wl, wr, bx = rand_state(rng, sup=True), rand_state(rng, sup=True), rng.normal()
        # chi >= 1 on both sides: the larger of the kinetic/Alfvenic speeds reaches c_f
        f = O.riemann_lhlld(GAMMA, wl, wr, bx, 0.5, 0.5)
        h = O.riemann(0, GAMMA, wl, wr, bx)
        cf = [np.sqrt(0.5 * ((GAMMA * w[4] + bx ** 2 + w[5] ** 2 + w[6] ** 2) / w[0] + np.sqrt(
            ((GAMMA * w[4] + bx ** 2 + w[5] ** 2 + w[6] ** 2) / w[0]) ** 2 - 4 * GAMMA * w[4] * bx ** 2 / w[0] ** 2)))
              for w in (wl, wr)]
        ke = [np.sqrt(w[1] ** 2 + w[2] ** 2 + w[3] ** 2 + (bx ** 2 + w[5] ** 2 + w[6] ** 2) / w[0]) for w in (wl, wr)]
        if max(ke) < max(cf) * 1.01:
            continue
        n += 1
        np.testing.assert_allclose(f, h, rtol=1e-9, atol=1e-11 * np.abs(h).max())
    assert n > 100


def test_theta_scales_pressure_jump_mass_flux():
    d, p0, dp = 1.0, 1.0, 1e-6
    wl = np.array([d, 0.0, 0.0, 0.0, p0 + dp, 0.0, 0.0])
    wr = np.array([d, 0.0, 0.0, 0.0, p0 - dp, 0.0, 0.0])
    cf = np.sqrt(GAMMA * p0 / d)
    f1 = O.riemann_lhlld(GAMMA, wl, wr, 0.0, 0.0, 0.0)[0]
    assert f1 > 0
    for dvn, dvt in [(0.0, -0.5), (0.2, -1.3), (-0.4, -0.8), (-2.0, -0.3)]:
        th1 = min(1.0, (cf - min(dvn, 0.0)) / (cf - min(dvt, 0.0)))
        f = O.riemann_lhlld(GAMMA, wl, wr, 0.0, dvn, dvt)[0]
        assert f / f1 == pytest.approx(th1 ** 4, rel=1e-5), (dvn, dvt)


def test_phi_scales_velocity_jump_pressure():
    """The star total pressure enters the energy flux through ulst.e (ptst * S_M): across a
    small velocity jump 2 du at mean velocity v0 (no field), HLLD's star pressure exceeds
    the ambient p0 by rho c_f du, LHLLD's by phi times that, so their energy fluxes differ
    by (phi - 1) rho c_f du * dF_e/dptst with dF_e/dptst = S_L S_M / (S_L - S_M)
    = v0 (c_f - v0) / c_f to first order in du."""
    d, p0, du = 1.0, 1.0, 1e-4
    cf = np.sqrt(GAMMA * p0 / d)
    for v0 in (0.1, 0.3):   # low Mach: chi < 1
        wl = np.array([d, v0 + du, 0.0, 0.0, p0, 0.0, 0.0])
        wr = np.array([d, v0 - du, 0.0, 0.0, p0, 0.0, 0.0])
        chi = min(1.0, (v0 + du) / cf)
        phi = chi * (2.0 - chi)
        fl = O.riemann_lhlld(GAMMA, wl, wr, 0.0)
        fh = O.riemann(0, GAMMA, wl, wr, 0.0)
        want = (phi - 1.0) * d * cf * du * v0 * (cf - v0) / cf
        print(f"v0 {v0}: F_e(LHLLD) - F_e(HLLD) = {fl[4] - fh[4]:.6e}, first-order prediction {want:.6e}",
              flush=True)
        assert (fl[4] - fh[4]) == pytest.approx(want, rel=0.03), v0
        np.testing.assert_allclose(fl[:4], fh[:4], rtol=1e-12)   # mass and momenta do not see ptst


def test_velocity_differences_on_linear_fields():
    """dvn/dvt through the full flux routine: with a uniform state except a linear
    velocity field, the LHLLD flux of a face equals the closed-form detector result."""
    d = make_desc(nx=(8, 8, 8), bnx=(8, 8, 8), box=((0, 1),) * 3, rsolver=3)
    o = Oracle(d)
    n = o.n[0]
    x = (np.arange(n) - d.nghost + 0.5) / 8
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    a = np.array([[0.3, -0.8, 0.2], [0.5, 0.1, -1.2], [-0.9, 0.4, 0.6]])   # dv_a / dx_b
    w = o.view("w")
    w[:, 0] = 1.0
    w[:, 4] = 0.6
    for c in range(3):
        w[:, 1 + c] = a[c, 0] * X + a[c, 1] * Y + a[c, 2] * Z
    o.view("bcc")[...] = 0.0
    for f in ("b1f", "b2f", "b3f"):
        o.view(f)[...] = 0.0
    o.set_time(0.0, 1e-3, 0)
    o.L.orc_calculate_fluxes(o.h, 1)
    h = 1.0 / 8
    for dirn, fx in enumerate(("fx1", "fx2", "fx3")):
        d1, d2 = (dirn + 1) % 3, (dirn + 2) % 3
        dvn = a[dirn, dirn] * h
        dvt = min(a[d1, d1] * h, a[d2, d2] * h)
        k = j = i = 5
        c = [i, j, k]
        cl = list(c)
        cl[dirn] -= 1
        wl = np.array([1.0, *[w[0, 1 + (dirn + q) % 3][cl[2], cl[1], cl[0]] for q in range(3)], 0.6, 0.0, 0.0])
        wr = np.array([1.0, *[w[0, 1 + (dirn + q) % 3][c[2], c[1], c[0]] for q in range(3)], 0.6, 0.0, 0.0])
        ref = O.riemann_lhlld(GAMMA, wl, wr, 0.0, dvn, dvt)
        got = o.get(fx)[0, :, k, j, i]
        perm = [0, 1 + dirn, 1 + d1, 1 + d2, 4]
        np.testing.assert_allclose(got[perm], ref[:5], rtol=1e-14, atol=1e-16)


def test_linear_wave_convergence_and_dissipation():
    from tests.test_oracle_mhd import run_linwave
    e = {}
    for rs in (0, 3):
        for nx in ((32, 16, 16), (64, 32, 32)):
            o, _ = run_linwave(nx, (16, 16, 16), rs=rs)
            e[(rs, nx[0])] = o.linear_wave_errors(0, 1e-6, 0.0)[8]
    print(f"fast wave RMS-L1: LHLLD {e[(3, 32)]:.4e} -> {e[(3, 64)]:.4e}, "
          f"HLLD {e[(0, 32)]:.4e} -> {e[(0, 64)]:.4e}", flush=True)
    assert e[(3, 64)] / e[(3, 32)] <= 0.4          # mhd_linwave.py:126 bar
    assert e[(3, 64)] <= 1.01 * e[(0, 64)]
