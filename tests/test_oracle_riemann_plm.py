"""Pins the oracle's Riemann solvers and PLM against what the mathematics fixes.

Checks (none retypes the oracle's formulas):
  - consistency: F(U,U) equals the physical MHD flux for HLLD and HLLE;
  - HLLD resolves an isolated stationary contact exactly (zero mass flux,
    Miyoshi & Kusano 2005, JCP 208, 315, sec. 4);
  - mirror symmetry x -> -x of both solvers;
  - PLM reproduces linear data exactly and flattens at extrema (van Leer limiter).
A dropped term, wrong sign or transposed operand in the solvers fails one of these.
Solvers: 0 = HLLD, 1 = HLLE, 2 = Roe (with its LLF fallback).
"""
import numpy as np
import pytest

GAMMA = 5.0 / 3.0


def phys_flux(w, bx, gamma=GAMMA):
    d, vx, vy, vz, p, by, bz = w
    b2 = bx * bx + by * by + bz * bz
    e = p / (gamma - 1) + 0.5 * d * (vx * vx + vy * vy + vz * vz) + 0.5 * b2
    pt = p + 0.5 * b2
    vb = vx * bx + vy * by + vz * bz
    return np.array([d * vx, d * vx * vx + pt - bx * bx, d * vx * vy - bx * by,
                     d * vx * vz - bx * bz, (e + pt) * vx - bx * vb,
                     by * vx - bx * vy, bz * vx - bx * vz])


STATES = [
    (np.array([1.0, 0.3, -0.2, 0.1, 0.6, 0.5, -0.3]), 0.7),
    (np.array([0.2, -1.5, 0.4, 0.0, 2.0, -1.1, 0.2]), -0.4),
    (np.array([3.0, 0.0, 0.0, 0.0, 1.0, 0.0, 0.0]), 0.0),
    (np.array([1.0, 2.0, 1.0, -1.0, 0.1, 2.0, 1.0]), 1.5),
]


@pytest.mark.parametrize("rs", [0, 1, 2])
@pytest.mark.parametrize("k", range(len(STATES)))
def test_consistency(orc, rs, k):
    w, bx = STATES[k]
    f = orc.riemann(rs, GAMMA, w, w, bx)
    ref = phys_flux(w, bx)
    np.testing.assert_allclose(f, ref, rtol=1e-13, atol=1e-13)


def test_hlld_stationary_contact_exact(orc):
    wl = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    wr = np.array([0.125, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    f = orc.riemann(0, GAMMA, wl, wr, 0.75)
    assert abs(f[0]) < 1e-15
    ref = phys_flux(wl, 0.75)
    np.testing.assert_allclose(f[1:], ref[1:], rtol=1e-13, atol=1e-14)


def mirror(w):
    d, vx, vy, vz, p, by, bz = w
    return np.array([d, -vx, vy, vz, p, by, bz])


@pytest.mark.parametrize("rs", [0, 1, 2])
def test_mirror_symmetry(orc, rs):
    rng = np.random.default_rng(7)
    for _ in range(50):
        wl = np.array([rng.uniform(0.2, 2), *rng.normal(size=3), rng.uniform(0.1, 2), *rng.normal(size=2)])
        wr = np.array([rng.uniform(0.2, 2), *rng.normal(size=3), rng.uniform(0.1, 2), *rng.normal(size=2)])
        bx = rng.normal()
        f = orc.riemann(rs, GAMMA, wl, wr, bx)
        # reflect x -> -x: left and right swap, vx and bx change sign
        g = orc.riemann(rs, GAMMA, mirror(wr), mirror(wl), -bx)
        # mass, transverse momentum, energy and transverse-B fluxes flip sign;
        # normal-momentum flux is even
        sign = np.array([-1, 1, -1, -1, -1, 1, 1])
        # induction: F(By) = By vx - Bx vy -> reflects to -(...) ... both vx, bx flip
        sign[5] = -1
        sign[6] = -1
        if rs == 2:
            # The reference's Roe energy flux uses (v1*bet3 - v3*bet2) for the
            # right-going Alfven wave (roe_mhd.cpp:481) where the eigenvector is
            # (v2*bet3 - v3*bet2) (pgen/linear_wave.cpp:753-757); the restatement keeps
            # it, so only the energy flux is exempt from the symmetry check.
            keep = [0, 1, 2, 3, 5, 6]
            np.testing.assert_allclose(g[keep], (sign * f)[keep], rtol=1e-9, atol=1e-11)
        else:
            np.testing.assert_allclose(g, sign * f, rtol=1e-9, atol=1e-11)


def test_hlld_upwind_supersonic(orc):
    w = np.array([1.0, 10.0, 0.1, 0.2, 0.5, 0.3, 0.1])
    wr = np.array([0.5, 9.0, 0.0, 0.1, 0.4, 0.2, 0.0])
    f = orc.riemann(0, GAMMA, w, wr, 0.2)
    np.testing.assert_allclose(f, phys_flux(w, 0.2), rtol=1e-14)


def test_plm_linear_exact(orc):
    q = 2.0 + 0.37 * np.arange(12)
    wl, wr = orc.plm_pencil(q)
    i = np.arange(1, 11)
    np.testing.assert_allclose(wl[i + 1], q[i] + 0.185, rtol=0, atol=1e-14)
    np.testing.assert_allclose(wr[i], q[i] - 0.185, rtol=0, atol=1e-14)


def test_plm_extremum_flat(orc):
    q = np.array([0.0, 1.0, 3.0, 1.0, 0.5, 0.5, 0.7])
    wl, wr = orc.plm_pencil(q)
    assert wl[3] == 3.0 and wr[2] == 3.0   # maximum at i=2
    assert wl[5] == 0.5 and wr[4] == 0.5   # plateau


def test_plm_monotone(orc):
    rng = np.random.default_rng(3)
    q = np.cumsum(rng.uniform(0, 1, 40))
    wl, wr = orc.plm_pencil(q)
    for i in range(1, 39):
        lo, hi = min(q[i - 1], q[i + 1]), max(q[i - 1], q[i + 1])
        assert lo - 1e-14 <= wl[i + 1] <= hi + 1e-14
        assert lo - 1e-14 <= wr[i] <= hi + 1e-14


def _cubic_cell_averages(n, h=0.1, c=(0.3, -1.2, 0.5, 0.7)):
    # cell averages of a cubic p(x) = c0 + c1 x + c2 x^2 + c3 x^3 over [i h, (i+1) h]
    x0 = np.arange(n) * h
    x1 = x0 + h
    P = lambda x: c[0] * x + c[1] * x ** 2 / 2 + c[2] * x ** 3 / 3 + c[3] * x ** 4 / 4
    p = lambda x: c[0] + c[1] * x + c[2] * x ** 2 + c[3] * x ** 3
    return (P(x1) - P(x0)) / h, p, x0


def test_ppm_cubic_interface_values(orc):
    """On cell averages of a monotone cubic, PPM's interface values are the exact point
    values (CW84 fourth-order interpolation) and the CS08 limiter stays inactive."""
    q, p, x0 = _cubic_cell_averages(16, h=0.05, c=(1.0, 2.0, 0.5, 0.3))
    wl, wr = orc.ppm_pencil(q)
    for i in range(2, 14):
        np.testing.assert_allclose(wl[i + 1], p(x0[i] + 0.05), rtol=0, atol=1e-13)
        np.testing.assert_allclose(wr[i], p(x0[i]), rtol=0, atol=1e-13)


def test_ppm_bounded_at_discontinuity(orc):
    q = np.array([1.0] * 8 + [0.1] * 8)
    wl, wr = orc.ppm_pencil(q)
    vals = np.concatenate([wl[3:14], wr[2:14]])
    assert vals.min() >= 0.1 - 1e-14 and vals.max() <= 1.0 + 1e-14


def test_roe_resolves_stationary_contact(orc):
    """A linearised (Roe) solver is exact for an isolated stationary contact."""
    wl = np.array([1.0, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    wr = np.array([0.125, 0.0, 0.0, 0.0, 1.0, 0.3, -0.2])
    f = orc.riemann(2, GAMMA, wl, wr, 0.75)
    assert abs(f[0]) < 1e-14
