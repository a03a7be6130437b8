"""GPU multigrid with multipole boundaries and FAS against the oracle.

The GPU keeps one grid per level with a ghost frame; the oracle keeps the reference's
MeshBlock levels and root grid with explicit exchanges and boundary fills.  Cell
coordinates agree to rounding (the GPU forms them from the mesh extent, the reference
per MeshBlock) and sums run in another order, so the bar is a relative tolerance:
- fixed V-cycle counts (same iterate sequence): potential within 1e-9 of max|phi|,
  and defect norms within 1e-6 relative;
- threshold 0 (iterate until the defect stalls): potential within 1e-10 of max|phi|.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.test_oracle_mg_multipole import BOX, gaussian

pytestmark = pytest.mark.gpu


def run_pair(agx, n, bnx, rho, nx=None, **kw):
    nx = nx or (n, n, n)
    box = kw.pop("box", BOX)
    d = make_desc(nx=nx, bnx=bnx, box=box, self_gravity=1, mg_bc=1, mg_mode=0, four_pi_G=1.0, **kw)
    U = np.zeros((5, nx[2], nx[1], nx[0]))
    U[0] = rho
    ub = agx.scatter_cells(d, U)
    o = Oracle(d)
    o.set("u", ub)
    o.mg_solve(0)
    m = agx.Mesh(d)
    m.set("u", ub)
    m.mg_solve(0)
    return o, m


@pytest.mark.parametrize("order,auto,nodip", [(4, 1, 0), (2, 1, 0), (4, 0, 0), (4, 0, 1), (2, 0, 1)])
def test_fixed_cycles_match(agx, order, auto, nodip):
    n = 32
    rho = gaussian(n, 0.05, c=(0.08, -0.05, 0.04))
    o, m = run_pair(agx, n, (16, 16, 16), rho, mporder=order, mp_auto_origin=auto, mp_nodipole=nodip,
                    mp_origin=(0.01, 0.02, -0.03), mg_threshold=-1.0, mg_niter=3)
    a, b = m.get("phi"), o.get("phi")
    err = np.abs(a - b).max() / np.abs(b).max()
    (nm, im), (no, io) = m.mg_last(), o.mg_last()
    print(f"order {order} auto {auto} nodipole {nodip}: phi rel diff {err:.3e}; defect {nm:.6e} vs {no:.6e}", flush=True)
    assert err <= 1e-9
    assert im == io == 3
    assert abs(nm - no) <= 1e-6 * no


def test_converged_potential_matches(agx):
    n = 32
    rho = gaussian(n, 0.06, c=(-0.05, 0.03, 0.02))
    o, m = run_pair(agx, n, (8, 8, 8), rho, mporder=4, mg_threshold=0.0)
    a, b = m.get("phi"), o.get("phi")
    err = np.abs(a - b).max() / np.abs(b).max()
    print(f"threshold 0: phi rel diff {err:.3e}; cycles {m.mg_last()} vs {o.mg_last()}", flush=True)
    assert err <= 1e-10


def test_noncubic_root_grid(agx):
    """4 x 2 x 2 MeshBlocks: the root grid's coarsest level is 2 x 1 x 1 cells and is
    swept twice (SolveCoarsestGrid, ni = 2)."""
    nx = (64, 32, 32)
    box = ((-1.0, 1.0), (-0.5, 0.5), (-0.5, 0.5))
    x1 = -1.0 + (np.arange(64) + 0.5) / 32
    x2 = -0.5 + (np.arange(32) + 0.5) / 32
    Z, Y, X = np.meshgrid(x2, x2, x1, indexing="ij")
    rho = np.exp(-((X - 0.2) ** 2 + Y ** 2 + (Z + 0.1) ** 2) / 0.06 ** 2)
    o, m = run_pair(agx, None, (16, 16, 16), rho, nx=nx, box=box, mporder=4, mg_threshold=-1.0, mg_niter=4)
    a, b = m.get("phi"), o.get("phi")
    err = np.abs(a - b).max() / np.abs(b).max()
    print(f"non-cubic root: phi rel diff {err:.3e}; defect {m.mg_last()} vs {o.mg_last()}", flush=True)
    assert err <= 1e-9


@pytest.mark.parametrize("bc", [dict(mg_bc=0), dict(mg_bc=1, mporder=4, mp_auto_origin=0)])
def test_source_mask_matches(agx, bc):
    """agx_mg_set_source_mask_sphere against the oracle's mask, and against a density
    that is already zero outside the sphere (bitwise on the GPU)"""
    from tests.test_oracle_mg_srcmask import C, N, RC, inside
    rho = 1.0 + gaussian(N, 0.1, c=(0.1, 0.0, -0.05), amp=5.0)
    d = make_desc(nx=(N, N, N), bnx=(16, 16, 16), box=BOX, self_gravity=1, mg_mode=0, four_pi_G=1.0,
                  mg_threshold=-1.0, mg_niter=4, **bc)
    phis = []
    for dens, mask in ((rho, True), (np.where(inside(N), rho, 0.0), False)):
        U = np.zeros((5, N, N, N))
        U[0] = dens
        ub = agx.scatter_cells(d, U)
        m = agx.Mesh(d)
        m.set("u", ub)
        if mask:
            m.mg_set_source_mask_sphere(*C, RC)
            o = Oracle(d)
            o.set("u", ub)
            o.mg_set_source_mask_sphere(*C, RC)
            o.mg_solve(0)
        m.mg_solve(0)
        phis.append(m.get("phi"))
        if mask:
            ref = o.get("phi")
            err = np.abs(phis[0] - ref).max() / np.abs(ref).max()
            print(f"{bc}: masked phi rel diff vs oracle {err:.3e}", flush=True)
            assert err <= 1e-9
    assert np.array_equal(phis[0], phis[1])


@pytest.mark.parametrize("mode,thr,nit", [(0, -1.0, 3), (0, 0.0, 0), (1, -1.0, 2)])
@pytest.mark.parametrize("precision", [0, 1])
def test_coarse_kernel_identical(agx, mode, thr, nit, precision):
    """the coarse end of the multipole V-cycle in one workgroup (k_mb_coarse) against one
    launch per step: the same potential bit for bit (FMG + fixed cycles, threshold 0,
    MGI), on a 4x2x2 root grid of 16^3 blocks; FP64, and the mixed form, whose FP32
    correction cycles use the kernel's homogeneous-boundary FP32 instance"""
    n = (64, 32, 32)
    x = [np.linspace(-0.5, 0.5, m, endpoint=False) + 0.5 / m for m in n]
    Z, Y, X = np.meshgrid(x[2], x[1], x[0], indexing="ij")
    rho = np.exp(-((X - 0.1) ** 2 + Y ** 2 + (Z + 0.05) ** 2) / 0.02)
    d = make_desc(nx=n, bnx=(16, 16, 16), box=((-0.5, 0.5),) * 3, self_gravity=1, mg_bc=1, mg_mode=mode,
                  four_pi_G=1.0, mporder=4, mg_threshold=thr, mg_niter=nit, precision=precision)
    U = np.zeros((5, n[2], n[1], n[0]))
    U[0] = rho
    ub = agx.scatter_cells(d, U)
    out = []
    for on in (1, 0):
        m = agx.Mesh(d)
        m.set("u", ub)
        m.set_mg_coarse(on)
        m.mg_solve(0)
        if mode == 1:
            m.mg_solve(0)      # MGI starts from the previous potential
        out.append((m.get("phi"), m.mg_last()))
        m.close()
    print(f"mode {mode}: cycles {out[0][1]} vs {out[1][1]}", flush=True)
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1]


@pytest.mark.parametrize("thr", ["stall", "absolute"])
def test_mixed_precision_matches(agx, thr):
    """Mixed precision with multipole boundaries: FP64 FMG, then FP32 V(1,1) correction
    cycles (FAS, the correction's ghosts by the homogeneous multipole chain) inside the
    FP64 defect correction.  The fixed point is the FP64 discrete solution with the
    multipole boundary values, so the potential agrees with the FP64 oracle's to the
    accuracy of the stopping rule: threshold 0 (iterate until the defect stalls) within
    1e-9 of the potential's maximum; an absolute threshold of 1e-10 x ||source||, reached
    by both, within 1e-8 (the bar of the periodic mixed form)."""
    n = 32
    rho = gaussian(n, 0.06, c=(-0.05, 0.03, 0.02))
    src_norm = float(np.sqrt(np.mean(rho ** 2)))
    eps = 0.0 if thr == "stall" else 1e-10 * src_norm
    kw = dict(nx=(n, n, n), bnx=(8, 8, 8), box=BOX, self_gravity=1, mg_bc=1, mg_mode=0, four_pi_G=1.0,
              mporder=4, mg_threshold=eps)
    U = np.zeros((5, n, n, n))
    U[0] = rho
    d64 = make_desc(**kw)
    ub = agx.scatter_cells(d64, U)
    o = Oracle(d64)
    o.set("u", ub)
    o.mg_solve(0)
    m = agx.Mesh(make_desc(precision=1, **kw))
    m.set("u", ub)
    m.mg_solve(0)
    a, b = m.get("phi"), o.get("phi")
    err = np.abs(a - b).max() / np.abs(b).max()
    (nm, im), (no, io) = m.mg_last(), o.mg_last()
    print(f"{thr}: phi rel diff {err:.3e}; mixed {im} cycles, defect {nm:.3e}; FP64 oracle {io} cycles, "
          f"defect {no:.3e}", flush=True)
    assert im <= io + 3   # FP32 corrections cost at most a few more cycles than FP64 V-cycles
    if thr == "stall":
        assert err <= 1e-9
    else:
        assert nm <= eps and no <= eps
        assert err <= 1e-8


def test_vcycles_need_a_solve_first(agx):
    """Multipole V-cycles continue a solve (its source, boundary values and FAS levels):
    before any agx_mg_solve the entry point refuses with -22 instead of cycling an unloaded
    problem; after one it runs."""
    d = make_desc(nx=(16, 16, 16), bnx=(8, 8, 8), box=BOX, self_gravity=1, mg_bc=1, mg_mode=0, four_pi_G=1.0,
                  mporder=4)
    m = agx.Mesh(d)
    U = np.zeros((5, 16, 16, 16))
    U[0] = gaussian(16, 0.08)
    m.set("u", agx.scatter_cells(d, U))
    with pytest.raises(agx.AgxError, match="-22"):
        m.mg_vcycles(1)
    m.mg_solve(0)
    m.mg_vcycles(2)
    m.sync()
    m.close()
