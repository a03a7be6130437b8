"""Pins the oracle's self-gravity source term (HydroSourceTerms::SelfGravity,
hydro/srcterms/self_gravity.cpp:35-125) and its coupling to the multigrid solve after
each VL2 stage (main.cpp:459-466).

1. Linear Jeans waves (pgen/jeans.cpp:53-230).  The reference's problem is hydro-only;
   here it runs in MHD with a uniform field along the wave vector, which leaves a
   longitudinal compressive wave untouched (the perturbed velocity is parallel to B, so
   no field line bends), and with no field at all.  Initial data and analytic
   solution are restated from jeans.cpp:59-150 and :166-203 (the error of w(IEN) is
   the pressure error, as the reference computes it).
   - stable, njeans = 0.5, 64x32x32 in 16^3 blocks, t = 1: RMS L1 error <= 1e-7
     (tst/regression/scripts/tests/grav/jeans_3d.py:100-104);
   - unstable, njeans = 1.5, t = 0.04: error <= 1e-7 at 32x16x16 and 64x32x32 and
     second-order convergence, error(64) <= 1.5 error(32) / 4
     (grav/unstable_jeans_3d_mg.py:54-80).
   The growth or oscillation comes only from the gravitational force: a sign error or a
   missing factor in the momentum source breaks these bounds.
2. Work done by a uniform field g = -grad(phi).  On a uniform state with velocity v
   and a linear potential phi = g.x, every flux divergence vanishes, so one stage
   changes momentum by -w rho g and total energy by -w rho (v.g) (w = stage weight
   times dt): Newton's second law and the work-energy theorem.  The energy source of
   self_gravity.cpp (mass fluxes times face potential differences) is second order
   in the Jeans amplitude and invisible to test 1; this closed form pins it, per axis.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc

GAMMA = 1.666666666667   # inputs/hydro/athinput.jeans_3d <hydro>/gamma
BOX = ((0.0, 1.0), (0.0, 0.5), (0.0, 0.5))


def jeans_params(njeans):
    """Mesh::InitUserMeshData of jeans.cpp:59-120 for the jeans_3d box."""
    L = [b[1] - b[0] for b in BOX]
    ang3 = np.arctan(L[0] / L[1])
    s3, c3 = np.sin(ang3), np.cos(ang3)
    ang2 = np.arctan(0.5 * (L[0] * c3 + L[1] * s3) / L[2])
    s2, c2 = np.sin(ang2), np.cos(ang2)
    lam = min(L[0] * c2 * c3, L[1] * c2 * s3, L[2] * s2)
    d0, p0 = 1.0, 1.0 / GAMMA
    cs2 = GAMMA * p0 / d0
    gconst = cs2 * np.pi * njeans ** 2 / (d0 * lam ** 2)
    k = 2.0 * np.pi / lam
    om2 = k * k * cs2 * (1.0 - njeans ** 2)
    return dict(s2=s2, c2=c2, s3=s3, c3=c3, k=k, om2=om2, om=np.sqrt(abs(om2)),
                four_pi_G=4.0 * np.pi * gconst, d0=d0, p0=p0)


def cell_centres(o):
    return [[o.coords(g, d)[1] for d in range(3)] for g in range(o.nb)]


def jeans_phase(P, X1, X2, X3):
    return P["c2"] * (X1 * P["c3"] + X2 * P["s3"]) + X3 * P["s2"]


def setup_jeans(n, njeans, amp, b0):
    P = jeans_params(njeans)
    desc = make_desc(nx=(n, n // 2, n // 2), bnx=(n // 4, n // 4, n // 4), box=BOX, cfl=0.3,
                     gamma=GAMMA, self_gravity=1, four_pi_G=P["four_pi_G"], mg_mode=0,
                     mg_threshold=0.0)
    o = Oracle(desc)
    u = o.view("u")
    kh = np.array([P["c2"] * P["c3"], P["c2"] * P["s3"], P["s2"]])   # wave vector direction
    B = b0 * kh
    for g, (x1, x2, x3) in enumerate(cell_centres(o)):
        X3, X2, X1 = np.meshgrid(x3, x2, x1, indexing="ij")
        x = jeans_phase(P, X1, X2, X3)
        sk, ck = np.sin(x * P["k"]), np.cos(x * P["k"])
        d = P["d0"] * (1.0 + amp * sk + amp * amp * np.sin(X1 * P["k"]))
        m = P["d0"] * (P["om"] / P["k"]) * amp * ck if P["om2"] < 0 else np.zeros_like(d)
        u[g, 0] = d
        for a in range(3):
            u[g, 1 + a] = m * kh[a]
        e = P["p0"] / (GAMMA - 1.0) * (1.0 + GAMMA * amp * sk)
        for a in range(3):
            e = e + 0.5 * u[g, 1 + a] ** 2 / d
        u[g, 4] = e + 0.5 * float(B @ B)
    for a, name in enumerate(("b1f", "b2f", "b3f")):
        o.view(name)[...] = B[a]
    return o, P


def jeans_error(o, P, amp):
    """RMS of the L1 errors of rho, M1..M3 and P (jeans.cpp:166-240, NFIELD = 0 there)."""
    t = o.time()[0]
    if P["om2"] < 0:
        sinot, cosot = -np.exp(P["om"] * t), np.exp(P["om"] * t)
    else:
        sinot, cosot = np.sin(P["om"] * t), np.cos(P["om"] * t)
    kh = np.array([P["c2"] * P["c3"], P["c2"] * P["s3"], P["s2"]])
    ng = o.desc.nghost
    sl = (slice(ng, -ng),) * 3
    u, w = o.get("u"), o.get("w")
    l1 = np.zeros(5)
    for g, (x1, x2, x3) in enumerate(cell_centres(o)):
        X3, X2, X1 = np.meshgrid(x3[ng:-ng], x2[ng:-ng], x1[ng:-ng], indexing="ij")
        x = jeans_phase(P, X1, X2, X3)
        sk, ck = np.sin(x * P["k"]), np.cos(x * P["k"])
        den = P["d0"] * (1.0 + amp * sk * cosot)
        m = -den * (P["om"] / P["k"]) * amp * ck * sinot
        l1[0] += np.abs(den - u[g, 0][sl]).sum()
        for a in range(3):
            l1[1 + a] += np.abs(m * kh[a] - u[g, 1 + a][sl]).sum()
        e0 = P["p0"] * (1.0 + GAMMA * amp * sk * cosot)
        l1[4] += np.abs(e0 - w[g, 4][sl]).sum()
    l1 /= o.desc.nx1 * o.desc.nx2 * o.desc.nx3
    return float(np.sqrt((l1 ** 2).sum()))


@pytest.mark.parametrize("b0", [0.0, 0.3])
def test_stable_jeans_wave(b0):
    amp = 1e-6
    o, P = setup_jeans(64, 0.5, amp, b0)
    o.initialize()
    o.run(1.0)
    err = jeans_error(o, P, amp)
    print(f"stable Jeans wave, B0 = {b0}: RMS L1 error {err:.4e} after {o.time()[2]} cycles",
          flush=True)
    assert err <= 1e-7
    # the wave has moved: a run without gravity (pure sound wave) misses the bound
    assert o.time()[2] > 50


@pytest.mark.parametrize("b0", [0.0, 0.3])
def test_unstable_jeans_growth_second_order(b0):
    amp = 1e-6
    errs = []
    for n in (32, 64):
        o, P = setup_jeans(n, 1.5, amp, b0)
        o.initialize()
        o.run(0.04)
        errs.append(jeans_error(o, P, amp))
    print(f"unstable Jeans, B0 = {b0}: RMS L1 {errs[0]:.4e} (32), {errs[1]:.4e} (64), "
          f"ratio {errs[1] / errs[0]:.3f}", flush=True)
    assert max(errs) <= 1e-7
    assert errs[1] <= 1.5 * errs[0] / 4.0


def test_jeans_without_gravity_fails_bound():
    """Discrimination check: the same unstable state evolved without self-gravity
    misses the growth."""
    amp = 1e-6
    o, P = setup_jeans(32, 1.5, amp, 0.0)
    o2 = Oracle(make_desc(nx=(32, 16, 16), bnx=(8, 8, 8), box=BOX, cfl=0.3, gamma=GAMMA))
    for n in ("u", "b1f", "b2f", "b3f"):
        o2.set(n, o.get(n))
    o2.initialize()
    o2.run(0.04)
    assert jeans_error(o2, P, amp) > 1e-7


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("stage", [1, 2])
def test_uniform_field_work_energy(axis, stage):
    rho, p = 1.4, 0.7
    v = np.array([0.3, -0.2, 0.15])
    B = np.array([0.2, 0.1, -0.3])
    g = np.zeros(3)
    g[axis] = 0.8 if axis != 1 else -1.1
    o = Oracle(make_desc(nx=(16, 8, 8), bnx=(8, 8, 8), box=((0.0, 1.0), (0.0, 0.5), (0.0, 0.25)),
                         cfl=0.3, gamma=5.0 / 3.0, self_gravity=1))
    u = o.view("u")
    u[:, 0] = rho
    for a in range(3):
        u[:, 1 + a] = rho * v[a]
    u[:, 4] = p / (5.0 / 3.0 - 1.0) + 0.5 * rho * float(v @ v) + 0.5 * float(B @ B)
    for a, name in enumerate(("b1f", "b2f", "b3f")):
        o.view(name)[...] = B[a]
    o.L.orc_bvals_cc(o.h, 0)
    o.L.orc_bvals_fc(o.h, 0)
    o.L.orc_cons2prim(o.h)
    dt = 0.01
    o.set_time(0.0, dt, 0)
    if stage == 2:
        # u1 holds the stage-1 state; a zero potential leaves it equal to u
        o.L.orc_calculate_fluxes(o.h, 1)
        o.L.orc_corner_e(o.h)
        o.view("phi")[...] = 0.0
        o.L.orc_integrate(o.h, 1)
    phi = o.view("phi")
    for gid, (x1, x2, x3) in enumerate(cell_centres(o)):
        X3, X2, X1 = np.meshgrid(x3, x2, x1, indexing="ij")
        phi[gid, 0] = -(g[0] * X1 + g[1] * X2 + g[2] * X3)   # g = -grad(phi)
    u0 = o.get("u")
    o.L.orc_calculate_fluxes(o.h, 2)
    o.L.orc_corner_e(o.h)
    o.L.orc_integrate(o.h, stage)
    du = o.get("u") - u0
    wght = (0.5 if stage == 1 else 1.0) * dt
    ng = 2
    sl = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    np.testing.assert_allclose(du[:, 0][sl], 0.0, atol=1e-15)
    for a in range(3):
        np.testing.assert_allclose(du[:, 1 + a][sl], wght * rho * g[a], rtol=1e-10, atol=1e-15)
    np.testing.assert_allclose(du[:, 4][sl], wght * rho * float(v @ g), rtol=1e-10, atol=1e-15)


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_gravity_source_on_flowing_state(axis):
    """The same linear potential on a state whose velocity varies along the axis: the
    source is the difference between a stage with the potential and one without.
    Momentum gains w rho g with rho the density the stage started from (the primitive
    w, not the partly updated u, which the flux divergence has changed here), and the
    energy gains w g times the mean of the cell's two face mass fluxes along g (force
    times mass transport through the cell)."""
    rho, p = 1.4, 0.7
    B = np.array([0.2, 0.1, -0.3])
    g = np.zeros(3)
    g[axis] = 0.8
    box = ((0.0, 1.0), (0.0, 0.5), (0.0, 0.25))
    runs = []
    for with_phi in (False, True):
        o = Oracle(make_desc(nx=(16, 8, 8), bnx=(8, 8, 8), box=box, cfl=0.3, gamma=5.0 / 3.0,
                             self_gravity=1))
        u = o.view("u")
        phi = o.view("phi")
        for gid, (x1, x2, x3) in enumerate(cell_centres(o)):
            X3, X2, X1 = np.meshgrid(x3, x2, x1, indexing="ij")
            X = (X1, X2, X3)[axis] / (box[axis][1] - box[axis][0])
            v = [0.1 + 0 * X1, -0.2 + 0 * X1, 0.15 + 0 * X1]
            v[axis] = 0.3 + 0.2 * np.sin(2.0 * np.pi * X)
            u[gid, 0] = rho
            for a in range(3):
                u[gid, 1 + a] = rho * v[a]
            u[gid, 4] = p / (5.0 / 3.0 - 1.0) + 0.5 * rho * sum(x * x for x in v) + 0.5 * float(B @ B)
            phi[gid, 0] = -(g[0] * X1 + g[1] * X2 + g[2] * X3) if with_phi else 0.0
        for a, name in enumerate(("b1f", "b2f", "b3f")):
            o.view(name)[...] = B[a]
        o.L.orc_bvals_cc(o.h, 0)
        o.L.orc_bvals_fc(o.h, 0)
        o.L.orc_cons2prim(o.h)
        o.set_time(0.0, 0.01, 0)
        o.L.orc_calculate_fluxes(o.h, 2)
        o.L.orc_corner_e(o.h)
        o.L.orc_integrate(o.h, 1)
        runs.append((o.get("u"), o.get("fx%d" % (axis + 1))[:, 0]))
    wght = 0.5 * 0.01
    du = runs[1][0] - runs[0][0]
    ng = 2
    sl = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
    F = runs[0][1]
    # mass flux through the cell's lower and upper faces along the axis (8^3 blocks)
    lo = [slice(None)] + [slice(ng, ng + 8)] * 3
    up = list(lo)
    up[3 - axis] = slice(ng + 1, ng + 9)
    Fl, Fu = F[tuple(lo)], F[tuple(up)]
    for a in range(3):
        np.testing.assert_allclose(du[:, 1 + a][sl], wght * rho * g[a], rtol=1e-9, atol=1e-13)
    np.testing.assert_allclose(du[:, 4][sl], wght * g[axis] * 0.5 * (Fl + Fu), rtol=1e-9, atol=1e-13)
    assert np.abs(Fu - Fl).max() > 1e-3   # the two face fluxes differ
