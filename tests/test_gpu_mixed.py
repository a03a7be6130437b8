"""Mixed-precision form (AGX_PREC_MIXED) against the FP64 oracle.

The mixed form reconstructs and solves the Riemann problem in FP32 from FP32 copies of
w and bcc, and keeps the conserved-variable update, C2P with floors, the corner EMFs and
CT in FP64.  Bars (stated per test): density and pressure within 1e-5 of the FP64 oracle
relative to their maximum over the mesh; the face field's div B at FP64 round-off,
because CT preserves it for any EMF values; the time step within 1e-5.  With PPM the
max-norm bar is twice the FP64 oracle's own response to FP32 rounding of its primitives
after every C2P (PPM's limiter is discontinuous at its sign tests, so that response is
1e-5..4e-5 here, set by a few cells, while the mean stays at 1e-7); the mixed form runs
the PPM arithmetic itself in FP64, so only the FP32 storage of w and bcc perturbs it.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, oracle_self_sensitivity, turbulent_state

pytestmark = pytest.mark.gpu
RHO_P_TOL = 1e-5


def rel_err(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def mixed_pair(agx, desc, state):
    o = Oracle(desc)
    load_state(o, state)
    dm = type(desc).from_buffer_copy(desc)
    dm.precision = 1
    m = agx.Mesh(dm)
    load_state(m, state)
    return o, m


def interior(a, ng):
    return a[:, :, ng:-ng, ng:-ng, ng:-ng]


@pytest.mark.parametrize("rsolver,order", [(0, 2), (1, 2), (0, 3), (3, 2)])
def test_mixed_fluxes(agx, rsolver, order):
    desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=rsolver,
                     xorder=order, nghost=3 if order == 3 else 2)
    o, m = mixed_pair(agx, desc, turbulent_state(desc, seed=11, mach=1.5))
    o.initialize()
    m.initialize()
    assert abs(m.time()[1] - o.time()[1]) <= 1e-5 * o.time()[1]
    o.L.orc_calculate_fluxes(o.h, order)
    m.calculate_fluxes(order)
    for n in ("fx1", "fx2", "fx3", "e3x1", "e2x1", "e1x2", "e3x2", "e1x3", "e2x3", "wx1", "wx2", "wx3"):
        e = rel_err(m.get(n), o.get(n))
        print(f"{n}: max rel err {e:.3e}", flush=True)
        # the CT weight is 0.5 + clamp(1024 dt F_rho / (dx (rho_l + rho_r))) (hydro.cpp:154-158):
        # its argument amplifies the FP32 flux rounding by 1024, so its bar is 1e-4 absolute
        # on a quantity in [0, 1]
        assert e <= (1e-4 if n.startswith("wx") else 1e-5), n


@pytest.mark.parametrize("rsolver,shape", [(0, "cube"), (1, "cube"), (0, "ragged"), (1, "ragged"),
                                            (0, "cube_ppm")])
def test_mixed_turbulence_run(agx, rsolver, shape):
    if shape == "cube":
        desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=rsolver)
    elif shape == "cube_ppm":
        desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=rsolver,
                         xorder=3, nghost=3)
    else:
        desc = make_desc(nx=(24, 16, 8), bnx=(8, 4, 8), box=((0, 1.5), (0, 1), (0, 0.5)), rsolver=rsolver)
    o, m = mixed_pair(agx, desc, turbulent_state(desc, seed=11, mach=1.5))
    tlim = 0.04
    o.initialize()
    m.initialize(tlim)
    o.run(tlim)
    m.run(tlim)
    assert m.time()[0] == pytest.approx(tlim, rel=1e-14) and o.time()[0] == pytest.approx(tlim, rel=1e-14)
    print(f"cycles: mixed {m.time()[2]}, fp64 {o.time()[2]}", flush=True)
    ng = desc.nghost
    wm, wo = interior(m.get("w"), ng), interior(o.get("w"), ng)
    erho = rel_err(wm[:, 0], wo[:, 0])
    ep = rel_err(wm[:, 4], wo[:, 4])
    um, uo = interior(m.get("u"), ng), interior(o.get("u"), ng)
    print(f"rho {erho:.3e}  P {ep:.3e}  E {rel_err(um[:, 4], uo[:, 4]):.3e}", flush=True)
    bar = {"rho": RHO_P_TOL, "P": RHO_P_TOL, "b1f": RHO_P_TOL, "b2f": RHO_P_TOL, "b3f": RHO_P_TOL}
    if shape == "cube_ppm":
        # PPM's extremum limiter (ppm.cpp:190-250) is discontinuous at its sign tests: the
        # FP64 oracle itself moves by 1e-5..4e-5 in max norm at a few cells when its
        # primitives are rounded to FP32, and no form that stores them in FP32 can be held
        # below that.  Bars: twice the oracle's own response (at least 1e-5) in max norm,
        # 1e-6 in the mean.
        sens = oracle_self_sensitivity(desc, turbulent_state(desc, seed=11, mach=1.5), lambda x: x.run(tlim))
        print("FP64 oracle's response to FP32-rounded primitives: " + ", ".join(f"{k} {v:.3e}" for k, v in sens.items()),
              flush=True)
        bar = {k: max(RHO_P_TOL, 2.0 * v) for k, v in sens.items()}
        l1 = lambda a, b: np.abs(a - b).mean() / np.abs(b).mean()
        print(f"L1: rho {l1(wm[:, 0], wo[:, 0]):.3e}  P {l1(wm[:, 4], wo[:, 4]):.3e}", flush=True)
        assert l1(wm[:, 0], wo[:, 0]) <= 1e-6 and l1(wm[:, 4], wo[:, 4]) <= 1e-6
    assert erho <= bar["rho"] and ep <= bar["P"]
    for n in ("b1f", "b2f", "b3f"):
        sl = (slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
        a, b = m.get(n)[sl], o.get(n)[sl]
        e = rel_err(a, b)
        print(f"{n}: {e:.3e} (bar {bar[n]:.3e})", flush=True)
        assert e <= bar[n]
    o2 = Oracle(desc)
    for n in ("b1f", "b2f", "b3f"):
        o2.set(n, m.get(n))
    assert o2.max_divb() < 1e-13   # absolute, B per unit length


@pytest.mark.parametrize("N,B", [(32, 8), (64, 16), (256, 64)])
def test_mixed_mg_reaches_fp64_accuracy(agx, N, B):
    """FP32 V-cycles inside the FP64 defect correction converge to the FP64 solution:
    with threshold 0 (iterate until the defect stalls, multigrid_driver.cpp:738-745) the
    potential matches the discrete FFT solution to 1e-10 of its maximum, as the FP64
    path does to 1e-11."""
    from tests.test_oracle_mg import blobs, fft_solution
    rho = blobs(N, 1)
    d = make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0, 1),) * 3, self_gravity=1, mg_mode=0,
                  mg_threshold=0.0, four_pi_G=0.7, precision=1)
    U = np.zeros((5, N, N, N))
    U[0] = rho
    m = agx.Mesh(d)
    m.set("u", agx.scatter_cells(d, U))
    m.mg_solve(0)
    phi = agx.gather_cells(d, m.get("phi"))[0]
    norm, iters = m.mg_last()
    ref = fft_solution(rho, 0.7)
    err = np.abs(phi - ref).max() / np.abs(ref).max()
    print(f"N={N} B={B}: {iters} correction cycles, defect {norm:.3e}, max rel err {err:.3e}", flush=True)
    assert err <= 1e-10
    assert iters <= 30


def test_mixed_mg_threshold_matches_oracle(agx):
    """Absolute threshold 1e-10 x ||source||: the mixed solve stops below the threshold and
    its potential agrees with the FP64 oracle's within the error the threshold allows."""
    from tests.test_oracle_mg import blobs
    N, B = 32, 8
    rho = blobs(N, 3)
    src_norm = float(np.sqrt(np.mean((rho - rho.mean()) ** 2)))
    kw = dict(nx=(N, N, N), bnx=(B, B, B), box=((0, 1),) * 3, self_gravity=1, mg_mode=0,
              mg_threshold=1e-10 * src_norm, four_pi_G=1.0)
    U = np.zeros((5, N, N, N))
    U[0] = rho
    o = Oracle(make_desc(**kw))
    ub = agx.scatter_cells(o.desc, U)
    o.set("u", ub)
    o.mg_solve(0)
    m = agx.Mesh(make_desc(precision=1, **kw))
    m.set("u", ub)
    m.mg_solve(0)
    norm, iters = m.mg_last()
    assert norm <= 1e-10 * src_norm
    a, b = m.get("phi"), o.get("phi")
    err = np.abs(a - b).max() / np.abs(b).max()
    print(f"mixed iters {iters} (oracle {o.mg_last()[1]}), norm {norm:.3e}, rel diff {err:.3e}", flush=True)
    assert err <= 1e-8


def test_mixed_selfgravity_cycles(agx):
    """Coupled MHD + multigrid gravity in the mixed form against the FP64 oracle."""
    kw = dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, self_gravity=1,
              four_pi_G=4.0, mg_mode=0, mg_threshold=0.0)
    d = make_desc(**kw)
    st = turbulent_state(d, seed=21, mach=1.0, drho=0.4)
    o = Oracle(d)
    m = agx.Mesh(make_desc(precision=1, **kw))
    load_state(o, st)
    load_state(m, st)
    tlim = 0.02
    o.initialize()
    m.initialize(tlim)
    o.run(tlim)
    m.run(tlim)
    ng = d.nghost
    wm, wo = interior(m.get("w"), ng), interior(o.get("w"), ng)
    erho, ep = rel_err(wm[:, 0], wo[:, 0]), rel_err(wm[:, 4], wo[:, 4])
    g1 = slice(ng, -ng)
    ephi = rel_err(m.get("phi")[:, :, g1, g1, g1], o.get("phi")[:, :, g1, g1, g1])
    print(f"cycles {m.time()[2]}/{o.time()[2]}: rho {erho:.3e} P {ep:.3e} phi {ephi:.3e}", flush=True)
    assert erho <= RHO_P_TOL and ep <= RHO_P_TOL and ephi <= RHO_P_TOL
