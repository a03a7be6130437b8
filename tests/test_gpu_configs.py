"""The benchmark workloads' own inputs and settings on the GPU, against the oracle.

- Linear fast wave as specified: ONE 64x32x32 MeshBlock on the periodic 3x1.5x1.5 box,
  gamma = 5/3, amplitude 1e-6, VL2 + PLM + HLLD + CT, CFL 0.4, run to t = 1.  The plain
  FP64 GPU path equals the oracle (bar 1e-12 relative; bitwise in practice), with the
  same cycle count and the same linear-wave error row (linear_wave.cpp:395-424).
- Decaying turbulence from its own generator (pgen.turbulence: k^-4 on 2 <= |k| <= 8,
  rms Mach 2, rho = 1, P = 0.6/gamma, Bx = 0.1) at 32^3 in 8 MeshBlocks: FP64 bitwise
  (1e-12) against the oracle; the mixed form within 1e-5 of it in rho and P (max norm,
  relative to the maximum); div B at round-off in both.
- The large run's state at reduced size (pgen.turbulence_lognormal: sigma_lnrho = 1,
  rms Mach 5, beta = 1, 4 pi G = 1) with PPM + HLLD + CT and multigrid self-gravity
  (FMG + V-cycles to 1e-10 x rms(rho - <rho>) each stage) at 32^3 in 8 MeshBlocks of 16^3
  (nghost 3): FP64 within 1e-12 of the oracle (potential 1e-10); the mixed form within
  1e-5 of the FP64 oracle in rho and P (max norm; or twice the FP64 oracle's own response
  to FP32 rounding of its primitives, if larger), 1e-6 in the potential and the means,
  1e-4 in bcc (the CT weight's amplification of FP32 mass-flux rounding, stated in the test).

div B bars: absolute max |div B| (B per unit length) is held to 1e-13 where dx is at most
1/32; the dimensionless max |div B| dx / max|B| (the round-off measure that does not
grow with resolution) to 1e-14.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_pgen
from tests.states import oracle_self_sensitivity

pytestmark = pytest.mark.gpu


def close(a, b, what, tol=1e-12):
    err = np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
    print(f"  {what}: rel diff {err:.3e}", flush=True)
    assert err <= tol, what


def rel_err(a, b):
    return float(np.abs(a - b).max() / np.abs(b).max())


def interior(a, ng):
    return a[:, :, ng:-ng, ng:-ng, ng:-ng]


def load_global(agx, targets, desc, U, b1, b2, b3):
    pgen = load_pgen()
    ub = agx.scatter_cells(desc, U)
    faces = pgen.scatter_faces(desc, b1, b2, b3)
    for t in targets:
        t.set("u", ub)
        for n, f in zip(("b1f", "b2f", "b3f"), faces):
            t.set(n, f)
    return ub, faces


def divb(desc, m):
    o = Oracle(desc)
    for n in ("b1f", "b2f", "b3f"):
        o.set(n, m.get(n))
    d = o.max_divb()
    dx = min((desc.x1max - desc.x1min) / desc.nx1, (desc.x2max - desc.x2min) / desc.nx2,
             (desc.x3max - desc.x3min) / desc.nx3)
    bmax = max(np.abs(m.get(n)).max() for n in ("b1f", "b2f", "b3f"))
    return d, d * dx / bmax


def test_linear_fast_wave_one_block_cfl04_t1(agx):
    desc = make_desc(nx=(64, 32, 32), bnx=(64, 32, 32), cfl=0.4, gamma=5.0 / 3.0)
    o = Oracle(desc)
    o.pgen_linear_wave(0, 1e-6, 0.0)
    m = agx.Mesh(desc)
    for n in ("u", "b1f", "b2f", "b3f"):
        m.set(n, o.get(n))
    o.initialize()
    m.initialize(1.0)
    no = o.run(1.0)
    nm = m.run(1.0)
    print(f"cycles to t = 1: oracle {no}, GPU {nm}; t {m.time()[0]!r}", flush=True)
    assert nm == no and m.time()[0] == o.time()[0] == 1.0
    for n in ("u", "w", "bcc", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), n)
    o2 = Oracle(desc)
    o2.set("u", m.get("u"))
    o2.set("bcc", m.get("bcc"))
    eg = o2.linear_wave_errors(0, 1e-6, 0.0)
    eo = o.linear_wave_errors(0, 1e-6, 0.0)
    print(f"  RMS-L1 error {eg[8]:.6e} (oracle {eo[8]:.6e}), max/L1 {eg[17]:.3f}", flush=True)
    np.testing.assert_allclose(eg, eo, rtol=1e-9)
    assert eg[17] <= 8.0                       # mhd_linwave.py:108-109
    dabs, drel = divb(desc, m)
    print(f"  max |div B| {dabs:.3e} (x dx / max|B|: {drel:.3e})", flush=True)
    assert drel <= 1e-14
    m.close()


def test_turbulence_generator_fp64_and_mixed(agx):
    pgen = load_pgen()
    desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0)
    U, b1, b2, b3 = pgen.turbulence(desc, seed=2, mach=2.0, bx=0.1)
    dmix = type(desc).from_buffer_copy(desc)
    dmix.precision = 1
    o, m, mx = Oracle(desc), agx.Mesh(desc), agx.Mesh(dmix)
    load_global(agx, (o, m, mx), desc, U, b1, b2, b3)
    o.initialize()
    m.initialize()
    mx.initialize()
    ncyc = 10
    for _ in range(ncyc):
        o.cycle(1e30)
    assert m.cycle(1e30, ncyc) == ncyc and mx.cycle(1e30, ncyc) == ncyc
    for n in ("u", "w", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), f"fp64 {n}")
    ng = desc.nghost
    wo, wx = interior(o.get("w"), ng), interior(mx.get("w"), ng)
    er, ep = rel_err(wx[:, 0], wo[:, 0]), rel_err(wx[:, 4], wo[:, 4])
    eb = rel_err(interior(mx.get("bcc"), ng), interior(o.get("bcc"), ng))
    print(f"  mixed vs FP64 oracle after {ncyc} cycles: rho {er:.3e}, P {ep:.3e}, bcc {eb:.3e}", flush=True)
    assert er <= 1e-5 and ep <= 1e-5 and eb <= 1e-5
    for tag, mm in (("fp64", m), ("mixed", mx)):
        dabs, drel = divb(desc, mm)
        print(f"  {tag} max |div B| {dabs:.3e} (x dx / max|B| {drel:.3e})", flush=True)
        assert dabs <= 1e-13 and drel <= 1e-14
    m.close()
    mx.close()


def test_lognormal_ppm_selfgravity_fp64_and_mixed(agx):
    pgen = load_pgen()
    base = dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0, xorder=3,
                nghost=3, rsolver=0, self_gravity=1, four_pi_G=1.0, mg_mode=0)
    d0 = make_desc(**base)
    U, b1, b2, b3 = pgen.turbulence_lognormal(d0, seed=5, mach=5.0, sigma_lnrho=1.0, beta=1.0)
    thr = 1e-10 * float(np.sqrt(np.mean((U[0] - U[0].mean()) ** 2)))
    desc = make_desc(mg_threshold=thr, **base)
    dmix = make_desc(mg_threshold=thr, precision=1, **base)
    o, m, mx = Oracle(desc), agx.Mesh(desc), agx.Mesh(dmix)
    ub, faces = load_global(agx, (o, m, mx), desc, U, b1, b2, b3)
    o.initialize()
    m.initialize()
    mx.initialize()
    ncyc = 3
    for _ in range(ncyc):
        o.cycle(1e30)
    assert m.cycle(1e30, ncyc) == ncyc and mx.cycle(1e30, ncyc) == ncyc
    for n in ("u", "w", "b1f", "b2f", "b3f"):
        close(m.get(n), o.get(n), f"fp64 {n}")
    close(m.get("phi"), o.get("phi"), "fp64 phi", tol=1e-10)
    ng = desc.nghost
    wo, wx = interior(o.get("w"), ng), interior(mx.get("w"), ng)
    po, px = interior(o.get("phi"), ng), interior(mx.get("phi"), ng)
    err = {"rho": rel_err(wx[:, 0], wo[:, 0]), "P": rel_err(wx[:, 4], wo[:, 4]),
           "bcc": rel_err(interior(mx.get("bcc"), ng), interior(o.get("bcc"), ng)), "phi": rel_err(px, po)}
    mean = {"rho": float(np.abs(wx[:, 0] - wo[:, 0]).mean() / np.abs(wo[:, 0]).mean()),
            "P": float(np.abs(wx[:, 4] - wo[:, 4]).mean() / np.abs(wo[:, 4]).mean())}
    print(f"  mixed vs FP64 oracle after {ncyc} cycles, max norm: "
          + ", ".join(f"{k} {v:.3e}" for k, v in err.items())
          + "; mean: " + ", ".join(f"{k} {v:.3e}" for k, v in mean.items()), flush=True)
    # bars: the potential and the means to 1e-6; the max norm of rho, P, bcc to 1e-5, or to
    # twice the FP64 oracle's own response to FP32 rounding of its primitives if that is
    # larger (PPM's extremum limiter is discontinuous at its sign tests, so that
    # response can reach 1e-5..1e-4 at a few cells of a Mach-5, 1000:1 density field)
    sens = oracle_self_sensitivity(desc, (ub,) + tuple(faces), lambda x: [x.cycle(1e30) for _ in range(ncyc)])
    print("  FP64 oracle's response to FP32-rounded primitives: "
          + ", ".join(f"{k} {sens[k]:.3e}" for k in ("rho", "P", "bcc")), flush=True)
    assert err["phi"] <= 1e-6 and mean["rho"] <= 1e-6 and mean["P"] <= 1e-6
    for k in ("rho", "P"):
        assert err[k] <= max(1e-5, 2.0 * sens[k]), k
    # bcc: 1e-4.  The CT upwind weight is 0.5 + clamp(1024 dt F_rho / (dx (rho_l + rho_r)),
    # -0.5, 0.5) (hydro.cpp:154-158): it multiplies the FP32 rounding of the mass flux by
    # ~1024 CFL, an absolute weight error of ~2e-5, which the corner EMFs carry into B
    # where the face and cell-centred EMFs differ most (the Mach-5 shocks here)
    assert err["bcc"] <= 1e-4
    for tag, mm in (("fp64", m), ("mixed", mx)):
        dabs, drel = divb(desc, mm)
        print(f"  {tag} max |div B| {dabs:.3e} (x dx / max|B| {drel:.3e})", flush=True)
        assert dabs <= 1e-13 and drel <= 1e-14
    m.close()
    mx.close()
