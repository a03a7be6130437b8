"""Pins the oracle's multigrid solver.

  - brute force: the converged solution equals the FFT solution of the same discrete
    7-point periodic Poisson problem (rho - mean) to round-off;
  - closed form: pgen/poisson.cpp:85-93 (rho = 2 + sin sin sin) converges to the
    analytic potential at second order;
  - linearity and decomposition independence (one MeshBlock vs many);
  - fixed-iteration mode (SolveIterativeFixedTimes) reduces the defect each cycle.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_agx


def solve(N, B, rho, mode=0, threshold=0.0, niter=10, fpg=1.0):
    agx = load_agx()
    d = make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0, 1), (0, 1), (0, 1)), self_gravity=1,
                  mg_mode=mode, mg_threshold=threshold, mg_niter=niter, four_pi_G=fpg)
    o = Oracle(d)
    U = np.zeros((5, N, N, N))
    U[0] = rho
    o.set("u", agx.scatter_cells(d, U))
    o.mg_solve(0)
    return agx.gather_cells(d, o.get("phi"))[0], o.mg_last()


def fft_solution(rho, fpg=1.0):
    N = rho.shape[0]
    k = np.fft.fftfreq(N) * N
    lam = (2 * np.cos(2 * np.pi * k / N) - 2) * N * N
    L = lam[:, None, None] + lam[None, :, None] + lam[None, None, :]
    L[0, 0, 0] = 1.0
    f = np.fft.fftn(fpg * (rho - rho.mean()))
    ph = np.real(np.fft.ifftn(f / L))
    return ph - ph.mean()


def blobs(N, seed):
    rng = np.random.default_rng(seed)
    x = (np.arange(N) + 0.5) / N
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rho = np.ones((N, N, N))
    for _ in range(6):
        c = rng.uniform(0, 1, 3)
        s = rng.uniform(0.05, 0.12)
        d2 = 0
        for ax, cc in zip((X, Y, Z), c):
            dd = np.abs(ax - cc)
            dd = np.minimum(dd, 1 - dd)
            d2 = d2 + dd * dd
        rho += 10 ** rng.uniform(0, 1) * np.exp(-d2 / (2 * s * s))
    return rho


@pytest.mark.parametrize("N,B", [(32, 8), (32, 32), (16, 4)])
def test_matches_fft(N, B):
    rho = blobs(N, 1)
    phi, (norm, it) = solve(N, B, rho, fpg=0.7)
    ref = fft_solution(rho, 0.7)
    assert np.abs(phi - ref).max() < 1e-12 * np.abs(ref).max()
    assert it > 0


def test_closed_form_second_order():
    errs = []
    for N in (16, 32, 64):
        x = (np.arange(N) + 0.5) / N
        Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
        den = np.sin(2 * np.pi * X) * np.sin(2 * np.pi * Y) * np.sin(2 * np.pi * Z)
        phi, _ = solve(N, N // 2, den + 2.0)
        errs.append(np.abs(phi + den / (3 * (2 * np.pi) ** 2)).max())
    assert 3.6 < errs[0] / errs[1] < 4.4
    assert 3.6 < errs[1] / errs[2] < 4.4


def test_linearity_and_decomposition():
    r1, r2 = blobs(16, 2), blobs(16, 3)
    p1, _ = solve(16, 8, r1)
    p2, _ = solve(16, 8, r2)
    p12, _ = solve(16, 8, 2.0 * r1 - 0.5 * r2)
    np.testing.assert_allclose(p12, 2.0 * p1 - 0.5 * p2, atol=1e-12 * np.abs(p12).max())
    q1, _ = solve(16, 16, r1)
    np.testing.assert_allclose(q1, p1, atol=1e-13 * np.abs(p1).max())


def test_fixed_iterations_reduce_defect():
    rho = blobs(16, 4)
    norms = [solve(16, 8, rho, mode=1, threshold=-1.0, niter=n)[1][0] for n in (1, 2, 4)]
    assert norms[0] > norms[1] > norms[2]
