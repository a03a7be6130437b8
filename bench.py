"""Benchmark: MHD zone-cycles/s and multigrid V-cycles/s at 256^3 on one MI355X.

Workload (value): 256^3 decaying MHD turbulence, 64 MeshBlocks of 64^3, VL2 + PLM +
HLLD + CT, plain FP64, synthetic seeded initial data.  One step = one full VL2 cycle
(both stages, ghost exchange, cons->prim, time-step reduction), inputs resident in HBM.
The same workload in the mixed-precision form (FP32 reconstruction and Riemann fluxes,
FP64 update, C2P and CT) is reported under "mixed".
Also reported: V(1,1) cycles/s of the multigrid (FP64, mixed, multipole boundaries) on
the 256^3 32-blob self-gravity problem; the coupled 256^3 MHD + multigrid self-gravity
cycle (VL2+PLM+HLLD+CT, lognormal density) in FP64 and in the mixed form, with the
mixed-vs-FP64 ratio and differences ("self_gravity"); the dominant kernel's roofline;
the refinement operators' bandwidth on 512 blocks of 32^3 ("refinement"), and whole-mesh
SMR steps on a 120-block two-level mesh (cell ghost exchange, EMF correction, the MHD
boundary step; "refinement.smr"); and the CPU oracle timed on a bounded sample of the
MHD workload.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK = 8.0e12   # MI355X HBM3E, B/s


def load_pkg():
    name = "athena_gamma_amd"
    spec = importlib.util.spec_from_file_location(name, os.path.join(ROOT, "athena-gamma_amd", "__init__.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    pspec = importlib.util.spec_from_file_location("agx_pgen", os.path.join(ROOT, "athena-gamma_amd", "pgen.py"))
    pgen = importlib.util.module_from_spec(pspec)
    pspec.loader.exec_module(pgen)
    return mod, pgen


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# Algorithmic HBM bytes per unit for each kernel class: compulsory words x word size,
# for the plain FP64 form and the mixed form (w, bcc, fluxes, face EMFs and CT weights
# stored in FP32; conserved state and face field FP64).  Unit counts per launch come from
# the mesh geometry below.
#   flux_x?    per face: read 5 prim + 2 transverse bcc (TW) + 1 normal face field (FP64),
#              write 5 fluxes + 2 face EMFs + 1 CT weight (TW)
#   integrate  per interior cell (k_stage): read 15 fluxes + 6 face EMFs + 3 weights +
#              3 v + 3 bcc (TW) + 5 conserved + 3 CT-source faces (FP64); write 5 conserved
#              + 3 faces (FP64) + 5 prim + 3 bcc (TW)
#   vl2        per interior cell (k_vl2, the whole stage in one launch; fluxes stay in
#              LDS): read 5 prim + 3 bcc (TW), 5 conserved + 3 face-field words for the
#              Riemann problems and CT (FP64) and, in the second stage, the 3 CT-source
#              faces (another register); write 5 conserved + 3 faces (FP64) + 5 prim +
#              3 bcc (TW).  Stage 1: 32 words, stage 2: 35 (FP64); a launch is counted
#              as the mean of the two (268 B FP64, 204 B mixed)
BYTES_PER_UNIT = {
    "fp64": {"flux_x1": 128, "flux_x2": 128, "flux_x3": 128, "integrate": 432, "vl2": (256 + 280) / 2},
    "mixed": {"flux_x1": 68, "flux_x2": 68, "flux_x3": 68, "integrate": 280, "vl2": (192 + 216) / 2},
}
STAGE_FORMS = ("16x16", "32x8", "64x4", "64x8", "one launch 16x16", "one launch 32x8")
VL2_FORM = 4
# whole VL2 cycle of a fully fused form, per zone (85 words of compulsory traffic):
# 680 B in FP64, 552 B with w and bcc in FP32
CYCLE_BYTES = {"fp64": 680.0, "mixed": 552.0}


def units_per_launch(desc, cls):
    nb = (desc.nx1 // desc.bnx1) * (desc.nx2 // desc.bnx2) * (desc.nx3 // desc.bnx3)
    b1, b2, b3 = desc.bnx1, desc.bnx2, desc.bnx3
    if cls == "flux_x1":
        return nb * (b1 + 1) * (b2 + 2) * (b3 + 2)
    if cls == "flux_x2":
        return nb * (b1 + 2) * (b2 + 1) * (b3 + 2)
    if cls == "flux_x3":
        return nb * (b1 + 2) * (b2 + 2) * (b3 + 1)
    if cls in ("integrate", "vl2"):
        return nb * b1 * b2 * b3
    raise KeyError(cls)


def max_divb(desc, m):
    """max |div B| over the interior cells from the downloaded face field (host-side
    numpy), as (absolute in B per unit length, dimensionless max|div B| * dx / max|B|):
    the round-off measure of constrained transport."""
    ng, b = desc.nghost, desc.bnx1
    dx = [(desc.x1max - desc.x1min) / desc.nx1, (desc.x2max - desc.x2min) / desc.nx2,
          (desc.x3max - desc.x3min) / desc.nx3]
    b1, b2, b3 = m.get("b1f"), m.get("b2f"), m.get("b3f")
    c = slice(ng, ng + b)
    d = (b1[:, c, c, ng + 1:ng + b + 1] - b1[:, c, c, c]) / dx[0]
    d += (b2[:, c, ng + 1:ng + b + 1, c] - b2[:, c, c, c]) / dx[1]
    d += (b3[:, ng + 1:ng + b + 1, c, c] - b3[:, c, c, c]) / dx[2]
    bmax = max(np.abs(b1[:, c, c, c]).max(), np.abs(b2[:, c, c, c]).max(), np.abs(b3[:, c, c, c]).max())
    dmax = float(np.abs(d).max())
    return dmax, dmax * min(dx) / float(bmax)


KCLASSES = ("flux_x1", "flux_x2", "flux_x3", "corner_e", "integrate", "bvals", "c2p", "newdt")


def mhd_bench(agx, pgen, args, precision):
    N, B = args.n, args.block
    desc = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, cfl=0.4,
                         gamma=5.0 / 3.0, xorder=2, rsolver=0,
                         precision=1 if precision == "mixed" else 0)
    t0 = time.time()
    U, b1, b2, b3 = pgen.turbulence(desc, seed=2, mach=2.0, bx=0.1)
    m = agx.Mesh(desc)
    m.set("u", agx.scatter_cells(desc, U))
    del U
    f1, f2, f3 = pgen.scatter_faces(desc, b1, b2, b3)
    m.set("b1f", f1)
    m.set("b2f", f2)
    m.set("b3f", f3)
    del f1, f2, f3, b1, b2, b3
    m.initialize()
    tile = m.autotune()
    log(f"[bench] {precision} MHD {N}^3 in {desc.nx1 // B}^3 blocks of {B}^3 set up in {time.time() - t0:.1f}s, "
        f"device {m.device_bytes() / 1e9:.2f} GB, stage form {STAGE_FORMS[tile]}")
    m.cycle(1e30, args.warmup)
    m.sync()
    t0 = time.perf_counter()
    done = m.cycle(1e30, args.steps)
    m.sync()
    t1 = time.perf_counter()
    assert done == args.steps
    sec = t1 - t0
    zc = float(N) ** 3 * args.steps / sec
    log(f"[bench] {precision}: {args.steps} cycles in {sec * 1e3:.1f} ms -> {zc:.4e} zone-cycles/s")
    # the same cycles replayed as one hipGraph launch per cycle
    m.set_graph(True)
    m.cycle(1e30, 2)
    m.sync()
    t0 = time.perf_counter()
    m.cycle(1e30, args.steps)
    m.sync()
    gsec = time.perf_counter() - t0
    m.set_graph(False)
    log(f"[bench] {precision} hipGraph replay: {args.steps} cycles in {gsec * 1e3:.1f} ms -> "
        f"{float(N) ** 3 * args.steps / gsec:.4e} zone-cycles/s")
    # instrumented pass: HIP events on the mesh stream around every launch
    m.timing(True)
    m.timing_reset()
    m.cycle(1e30, 2)
    m.sync()
    per = {}
    for cls in KCLASSES:
        ms, n = m.timing_get(cls)
        per[cls] = (ms, n)
    m.timing(False)
    state = None
    if args.compare:
        ng = desc.nghost
        sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
        state = (m.time(), m.get("w")[sl].astype(np.float64), max_divb(desc, m))
        log(f"[bench] {precision} max |div B| after {int(state[0][2])} cycles: {state[2][0]:.3e} "
            f"(x dx / max|B|: {state[2][1]:.3e})")
    m.close()
    mhd_bench.tile = tile
    return desc, zc, sec * 1e3 / args.steps, per, state


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "mhd_turb256_plm_hlld_fp64_mixed_hbm_traffic.txt")


def kernel_name(precision, tile, order=2, push=True):
    """rocprof name of the stage kernel of this form (adiabatic, HLLD).  push: the form
    that writes ghost u, w and bcc itself, which runs while the two copies of every shared
    face agree (always the case for a state uploaded as global arrays, as here)."""
    t = "double" if precision == "fp64" else "float"
    p = "true" if push else "false"
    if tile >= VL2_FORM:
        ti = "16, 16" if tile == 4 else "32, 8"
        return f"k_vl2<{ti}, {order}, 0, {t}, false, {p}>"
    return f"k_stage<{('16, 16', '32, 8', '64, 4', '64, 8')[tile]}, {t}, false, {p}>"


def measured_traffic(desc, precision, tile):
    """HBM bytes per launch of the stage kernel of this form and precision, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE runs (already scaled by the copy-kernel
    factors), when the workload is the one profiled (256^3 in 64^3 blocks); else None.
    For the one-launch form the profile lists the PLM stage (order 2)."""
    if (desc.nx1, desc.bnx1) != (256, 64) or not os.path.exists(TRAFFIC_PROFILE):
        return None
    name = kernel_name(precision, tile)
    for line in open(TRAFFIC_PROFILE):
        if line.startswith(name):
            f = line[len(name):].split()
            return {"bytes": (float(f[1]) + float(f[2])) * 1e9, "read_GB": float(f[1]), "write_GB": float(f[2]),
                    "kernel": name,
                    "source": f"{os.path.relpath(TRAFFIC_PROFILE, ROOT)} (rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE; scaled by the 1 GiB copy factors)"}
    return None


VALU_PROFILE = os.path.join(ROOT, "profiles", "mhd_turb256_plm_hlld_stage_kernels_valu.txt")


def measured_valu_issue(desc, kernel):
    """Wave-level VALU instructions of the kernel over what 1024 SIMDs issue in its time
    (one 64-lane instruction per 4 cycles), from the committed rocprofv3 SQ-counter
    summary of the 256^3 workload; None for other workloads or kernels."""
    if (desc.nx1, desc.bnx1) != (256, 64) or not os.path.exists(VALU_PROFILE):
        return None
    for line in open(VALU_PROFILE):
        if line.startswith(kernel):
            f = line[len(kernel):].split()
            return {"frac": float(f[2]), "insts_valu": float(f[1]), "cycles": float(f[0]),
                    "source": os.path.relpath(VALU_PROFILE, ROOT) + " (rocprofv3 --pmc SQ_INSTS_VALU, GRBM_GUI_ACTIVE)"}
    return None


def roofline(desc, per, precision, tile=None):
    """Dominant kernel (among those with an algorithmic byte count) and its HBM rate.
    tile: the stage form agx_autotune picked (index into STAGE_FORMS)."""
    table = dict(BYTES_PER_UNIT[precision])
    # the integrate timer class holds k_stage (split forms) or k_vl2 (one-launch forms)
    vl2 = tile is not None and tile >= VL2_FORM
    table.pop("integrate" if vl2 else "vl2")
    dom = max(table, key=lambda c: per["integrate" if c == "vl2" else c][0])
    ms_tot, n_l = per["integrate" if dom == "vl2" else dom]
    avg_s = ms_tot / max(n_l, 1) / 1e3
    alg = table[dom] * units_per_launch(desc, dom)
    achieved = alg / avg_s
    tr = measured_traffic(desc, precision, tile) if (dom in ("integrate", "vl2") and tile is not None) else None
    out = {"bound": "hbm",
           "kernel": kernel_name(precision, tile, desc.xorder) if dom in ("integrate", "vl2") else dom,
           "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9,
           "unit": "GB/s", "frac": achieved / HBM_PEAK, "traffic": tr["bytes"] if tr else None,
           "avg_launch_ms": avg_s * 1e3, "alg_bytes_per_launch": alg}
    vi = measured_valu_issue(desc, out["kernel"])
    if vi is not None:
        out["valu_issue"] = vi
    if tr:
        out["traffic_detail"] = {"read_GB": tr["read_GB"], "write_GB": tr["write_GB"], "source": tr["source"],
                                 "traffic_rate_GBs": tr["bytes"] / avg_s / 1e9}
    return out


def mg_bench(agx, pgen, args, precision="fp64"):
    """precision "fp64" or "mixed" (periodic, the 32-blob problem), or "multipole" /
    "multipole_mixed" (isolated: the blobs without the uniform background, multipole
    order 4 boundaries about the centre of mass, FAS cycles; FP64, or FP32 correction
    cycles inside the FP64 defect correction)."""
    N, B = args.n, args.block
    mp = precision.startswith("multipole")
    desc = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, self_gravity=1,
                         four_pi_G=1.0, mg_mode=0, mg_threshold=0.0,
                         precision=1 if precision.endswith("mixed") else 0, mg_bc=1 if mp else 0, mporder=4)
    rho = pgen.blobs(desc, seed=3)
    if mp:
        rho = rho - 1.0
    U = np.zeros((5,) + rho.shape)
    U[0] = rho
    m = agx.Mesh(desc)
    m.set("u", agx.scatter_cells(desc, U))
    del U
    # reference tolerance: 1e-10 relative to ||4 pi G (rho - mean)|| in the solver's norm
    src_norm = float(np.sqrt(np.mean((rho - rho.mean()) ** 2)))
    m.mg_set_mode(0, 1e-10 * src_norm, -1)
    t0 = time.perf_counter()
    m.mg_solve(0)
    m.sync()
    t_solve = time.perf_counter() - t0
    norm, iters = m.mg_last()
    log(f"[bench] {precision} MG FMG+{iters} V-cycles to {norm:.3e} (target {1e-10 * src_norm:.3e}) "
        f"in {t_solve * 1e3:.1f} ms")
    m.mg_vcycles(2)
    m.sync()
    nv = args.mg_cycles
    t0 = time.perf_counter()
    m.mg_vcycles(nv)
    m.sync()
    sec = time.perf_counter() - t0
    vps = nv / sec
    m.timing(True)
    m.timing_reset()
    m.mg_vcycles(3)
    m.sync()
    per = {c: m.timing_get(c) for c in ("mg_smooth", "mg_restrict", "mg_prolong", "mg_other")}
    m.timing(False)
    m.close()
    log(f"[bench] {precision} MG {nv} V-cycles in {sec * 1e3:.1f} ms -> {vps:.1f} V-cycles/s")
    return vps, iters, norm, t_solve, per


def sg_run(agx, pgen, N, B, xorder, mach, steps, tag, bx=None, sensitivity=False):
    """MHD + multigrid self-gravity at N^3 in MeshBlocks of B^3: lognormal density
    (sigma_lnrho = 1), rms Mach `mach`, uniform Bx (plasma beta 1, or Bx = `bx`), 4 pi G = 1, VL2 + PLM (xorder 2) or PPM
    (xorder 3) + HLLD + CT, FMG + V-cycles to 1e-10 x rms(rho - <rho>) at every stage.  The
    plain FP64 path and the mixed form (FP32 reconstruction/Riemann and FP32 V-cycles in
    FP64 defect correction) run `steps` cycles each from the same state; returns both
    rates and the mixed-vs-FP64 differences after those cycles.  sensitivity: also run the
    FP64 path from the state with its conserved variables multiplied by (1 + 3e-8 N(0,1))
    and report how far that moves it (the discretization's own response to a perturbation
    below FP32 rounding, the scale the mixed-vs-FP64 differences are read against)."""
    ng = 3 if xorder == 3 else 2
    base = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0,
                         xorder=xorder, nghost=ng, rsolver=0)
    t0 = time.time()
    U, b1, b2, b3 = pgen.turbulence_lognormal(base, seed=5, mach=mach, sigma_lnrho=1.0, beta=1.0, bx=bx)
    src_norm = float(np.sqrt(np.mean((U[0] - U[0].mean()) ** 2)))
    ub = agx.scatter_cells(base, U)
    del U
    faces = pgen.scatter_faces(base, b1, b2, b3)
    del b1, b2, b3
    log(f"[{tag}] {N}^3 lognormal turbulence built in {time.time() - t0:.1f}s")
    res = {}
    for prec in ("fp64", "mixed") + (("fp64_perturbed",) if sensitivity else ()):
        desc = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, cfl=0.4,
                             gamma=5.0 / 3.0, xorder=xorder, nghost=ng, rsolver=0, self_gravity=1,
                             four_pi_G=1.0, mg_mode=0, mg_threshold=1e-10 * src_norm,
                             precision=1 if prec == "mixed" else 0)
        m = agx.Mesh(desc)
        if prec == "fp64_perturbed":
            m.set("u", ub * (1.0 + 3e-8 * np.random.default_rng(1).standard_normal(ub.shape)))
        else:
            m.set("u", ub)
        for n, f in zip(("b1f", "b2f", "b3f"), faces):
            m.set(n, f)
        m.initialize()
        tile = m.autotune()
        m.calculate_fluxes(xorder)
        m.sync()
        t0 = time.perf_counter()
        done = m.cycle(1e30, steps)
        m.sync()
        sec = time.perf_counter() - t0
        assert done == steps
        sl = (slice(None), slice(None), slice(ng, -ng), slice(ng, -ng), slice(ng, -ng))
        w = m.get("w")[sl][:, [0, 4]].astype(np.float64)
        bcc = m.get("bcc")[sl].astype(np.float64)
        phi = m.get("phi")[sl].astype(np.float64)
        divb = max_divb(desc, m)
        vc = m.mg_last()[1]   # V-cycles after FMG of the last solve
        m.timing(True)
        m.timing_reset()
        m.cycle(1e30, 2)
        m.sync()
        per = {c: m.timing_get(c) for c in KCLASSES + ("mg_smooth", "mg_restrict", "mg_prolong", "mg_other")}
        m.timing(False)
        # whole-cycle byte model: the MHD cycle (680 B FP64 / 552 B mixed per zone) and two
        # solves, each load + FMG (about 8/7 of a finest-level V-cycle) + vc V-cycles at
        # 148.6 B (FP64) or 114.3 B (mixed correction cycle) per zone, + retrieve (16 B)
        key = "mixed" if prec == "mixed" else "fp64"
        vbytes = 114.3 if key == "mixed" else 148.6
        bcyc = CYCLE_BYTES[key] + 2 * (24.0 + vbytes * (8.0 / 7.0 + vc) + 16.0)
        res[prec] = dict(zc=float(N) ** 3 * steps / sec, ms=sec * 1e3 / steps, w=w, bcc=bcc, phi=phi, divb=divb,
                         per=per, tile=tile, bytes=m.device_bytes(), vc=vc, bcyc=bcyc,
                         rf=roofline(desc, per, key, tile))
        log(f"[{tag}] {prec}: {steps} cycles in {sec:.2f}s -> {res[prec]['zc']:.4e} zone-cycles/s, "
            f"device {m.device_bytes() / 1e9:.1f} GB, stage form {STAGE_FORMS[tile]}; per cycle: "
            + ", ".join(f"{k} {v[0] / 2:.2f} ms" for k, v in per.items() if v[0] > 0))
        m.close()
    a, b = res["mixed"], res["fp64"]
    err = lambda x, y: float(np.abs(x - y).max() / np.abs(y).max())
    l1 = lambda x, y: float(np.abs(x - y).mean() / np.abs(y).mean())
    out = {
        "workload": f"{N}^3 self-gravitating MHD turbulence (lognormal rho, sigma_lnrho=1, rms Mach {mach:g}, "
                    f"{'beta=1' if bx is None else f'Bx={bx:g}'}, 4piG=1), {(N // B) ** 3} MeshBlocks of {B}^3, "
                    f"VL2+{'PPM' if xorder == 3 else 'PLM'}+HLLD+CT, FMG + V-cycles to 1e-10 x rms(rho - <rho>) "
                    f"each stage",
        "mixed": {"value": a["zc"], "ms_per_step": a["ms"], "roofline": a["rf"], "max_divb": a["divb"][0],
                  "max_divb_dx_over_B": a["divb"][1], "vcycles_per_solve": a["vc"],
                  "cycle_bytes_per_zone_model": a["bcyc"], "cycle_roofline_frac": a["zc"] * a["bcyc"] / HBM_PEAK,
                  "kernel_ms_per_cycle": {k: v[0] / 2 for k, v in a["per"].items()}},
        "fp64": {"value": b["zc"], "ms_per_step": b["ms"], "roofline": b["rf"], "max_divb": b["divb"][0],
                 "max_divb_dx_over_B": b["divb"][1], "vcycles_per_solve": b["vc"],
                 "cycle_bytes_per_zone_model": b["bcyc"], "cycle_roofline_frac": b["zc"] * b["bcyc"] / HBM_PEAK,
                 "kernel_ms_per_cycle": {k: v[0] / 2 for k, v in b["per"].items()}},
        "vs_fp64": a["zc"] / b["zc"],
        "rel_err_rho": err(a["w"][:, 0], b["w"][:, 0]),
        "rel_err_p": err(a["w"][:, 1], b["w"][:, 1]),
        "rel_err_bcc": err(a["bcc"], b["bcc"]),
        "rel_err_phi": err(a["phi"], b["phi"]),
        "rel_l1_rho": l1(a["w"][:, 0], b["w"][:, 0]),
        "rel_l1_p": l1(a["w"][:, 1], b["w"][:, 1]),
        "rel_l1_bcc": l1(a["bcc"], b["bcc"]),
        "error_note": f"max-norm (rel_err_*) and mean (rel_l1_*) differences relative to the FP64 path's "
                      f"max and mean magnitude after the same {steps} cycles",
    }
    if sensitivity:
        c = res["fp64_perturbed"]
        out["fp64_self_sensitivity"] = {
            "perturbation": "conserved variables x (1 + 3e-8 N(0,1)), seed 1",
            "rel_err_rho": err(c["w"][:, 0], b["w"][:, 0]), "rel_err_p": err(c["w"][:, 1], b["w"][:, 1]),
            "rel_err_bcc": err(c["bcc"], b["bcc"]), "rel_err_phi": err(c["phi"], b["phi"]),
            "rel_l1_rho": l1(c["w"][:, 0], b["w"][:, 0]), "rel_l1_p": l1(c["w"][:, 1], b["w"][:, 1])}
        log(f"[{tag}] FP64 vs FP64 from the perturbed state: " +
            ", ".join(f"{k} {v:.3e}" for k, v in out["fp64_self_sensitivity"].items() if k != "perturbation"))
    log(f"[{tag}] mixed vs fp64: {out['vs_fp64']:.3f}x; after {steps} cycles max rho {out['rel_err_rho']:.3e}, "
        f"P {out['rel_err_p']:.3e}, bcc {out['rel_err_bcc']:.3e}, phi {out['rel_err_phi']:.3e}; "
        f"mean rho {out['rel_l1_rho']:.3e}, P {out['rel_l1_p']:.3e}, bcc {out['rel_l1_bcc']:.3e}")
    return out


def sg_graph_bench(agx, pgen, N, B, steps, niter, bx=0.1, mach=2.0):
    """The self-gravity workload of sg_run (lognormal density, rms Mach `mach`, Bx, PLM +
    HLLD + CT, FP64), `steps` cycles eager and `steps` cycles replayed from hipGraphs, in
    two solver forms: a fixed V-cycle count per solve (niteration = `niter`, FMG first; one
    graph per cycle with both solves recorded) and sg_run's threshold, 1e-10 x rms(rho -
    <rho>) (three graph segments per cycle, each solve's loop recorded with the last
    count and gated on the device by the stopping rule).  Returns the rates."""
    base = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0,
                         xorder=2, nghost=2, rsolver=0)
    U, b1, b2, b3 = pgen.turbulence_lognormal(base, seed=5, mach=mach, sigma_lnrho=1.0, beta=1.0, bx=bx)
    src_norm = float(np.sqrt(np.mean((U[0] - U[0].mean()) ** 2)))
    ub = agx.scatter_cells(base, U)
    del U
    faces = pgen.scatter_faces(base, b1, b2, b3)
    del b1, b2, b3
    out = {}
    for form, eps, nit in (("fixed", -1.0, niter), ("threshold", 1e-10 * src_norm, -1)):
        desc = agx.make_desc(nx=(N, N, N), bnx=(B, B, B), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0,
                             xorder=2, nghost=2, rsolver=0, self_gravity=1, four_pi_G=1.0, mg_mode=0,
                             mg_threshold=eps, mg_niter=nit)
        for graph in (False, True):
            m = agx.Mesh(desc)
            m.set("u", ub)
            for n, f in zip(("b1f", "b2f", "b3f"), faces):
                m.set(n, f)
            m.initialize()
            m.autotune()
            m.set_graph(graph)
            m.cycle(1e30, 2)
            m.sync()
            t0 = time.perf_counter()
            assert m.cycle(1e30, steps) == steps
            m.sync()
            sec = time.perf_counter() - t0
            out[(form, graph)] = float(N) ** 3 * steps / sec
            if form == "threshold":
                out["iters"] = m.mg_last()[1]
            m.close()
    res = {"workload": f"{N}^3 self-gravitating MHD (lognormal rho, rms Mach {mach:g}, Bx={bx:g}), "
                       f"{(N // B) ** 3} MeshBlocks of {B}^3, VL2+PLM+HLLD+CT FP64, FMG + {niter} V-cycles per solve",
           "eager": out[("fixed", False)], "graph": out[("fixed", True)],
           "graph_vs_eager": out[("fixed", True)] / out[("fixed", False)],
           "threshold": {"workload": "the same with FMG + V-cycles to 1e-10 x rms(rho - <rho>) per solve",
                         "eager": out[("threshold", False)], "graph": out[("threshold", True)],
                         "graph_vs_eager": out[("threshold", True)] / out[("threshold", False)],
                         "vcycles_last_solve": out["iters"]},
           "unit": "zone-cycles/s", "steps": steps}
    log(f"[sg-graph] fixed {niter} V-cycles per solve: eager {res['eager']:.4e}, graph {res['graph']:.4e} "
        f"zone-cycles/s ({res['graph_vs_eager']:.3f}x)")
    th = res["threshold"]
    log(f"[sg-graph] threshold ({th['vcycles_last_solve']} V-cycles in the last solve): eager {th['eager']:.4e}, "
        f"graph {th['graph']:.4e} zone-cycles/s ({th['graph_vs_eager']:.3f}x)")
    return res


def large_bench(agx, pgen, args):
    """The large single-GPU run: 512^3 periodic MHD + multigrid self-gravity in 512
    MeshBlocks of 64^3, VL2+PPM+HLLD+CT, lognormal density (sigma_lnrho = 1), rms Mach 5,
    beta = 1, 4 pi G = 1; the mixed form is the headline value."""
    N, B = args.large_n, 64
    r = sg_run(agx, pgen, N, B, 3, 5.0, args.steps, "large", sensitivity=True)
    a = r["mixed"]
    out = {
        "metric": "MHD+MG zone-cycles/s at 512^3 on 1 MI355X, mixed precision; % of HBM roofline",
        "value": a["value"], "unit": "zone-cycles/s", "n_gpus": 1, "steps": args.steps, "warmup": 0,
        "ms_per_step": a["ms_per_step"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 reconstruction+Riemann and MG smoothing, f64 update/C2P/CT/defect",
        "data": "synthetic",
        "config": {"workload": r["workload"], "mesh": [N] * 3, "meshblock": [B] * 3, "nghost": 3},
        "roofline": a["roofline"],
        "fp64": {k: r["fp64"][k] for k in ("value", "ms_per_step", "roofline", "max_divb", "max_divb_dx_over_B")},
        "max_divb": a["max_divb"], "max_divb_dx_over_B": a["max_divb_dx_over_B"],
        "kernel_ms_per_cycle": a["kernel_ms_per_cycle"],
        "cycle_roofline_frac": a["cycle_roofline_frac"],
        "cycle_bytes_per_zone_model": a["cycle_bytes_per_zone_model"],
    }
    out.update({k: v for k, v in r.items()
                if k.startswith("rel_") or k in ("vs_fp64", "error_note", "fp64_self_sensitivity")})
    print(json.dumps(out), flush=True)


def smr_cycle_bench(agx, pgen, args):
    """Static mesh refinement, three levels: a 4^3 root grid of 32^3 MeshBlocks on the
    periodic unit cube, the central 2^3 root blocks refined and one inner level-1 block
    refined again (127 blocks: 56 + 63 + 8, 4.16e6 cells), a smooth divergence-free state
    (pgen.divfree_blocks), VL2+PLM+HLLD+CT in FP64 through SMRMesh: the task-level launches
    plus flux correction, EMF correction and the ghost exchange with prolongation across
    levels, every step on the mesh's stream.  Zone-cycles/s over K cycles; the per-class
    kernel timers of two more cycles split the cycle into the MHD kernels and the rest
    (refinement operators and launch gaps)."""
    smr = agx._load_smr()
    nx, ng, box = args.smr_block, 2, ((0.0, 1.0),) * 3
    inner = [(x, y, z) for z in (1, 2) for y in (1, 2) for x in (1, 2)]
    leaves = [(0, x, y, z) for z in range(4) for y in range(4) for x in range(4) if (x, y, z) not in inner]
    for (x, y, z) in inner:
        leaves += [(1, 2 * x + a, 2 * y + b, 2 * z + c) for c in range(2) for b in range(2) for a in range(2)
                   if (2 * x + a, 2 * y + b, 2 * z + c) != (3, 3, 3)]
    leaves += [(2, 6 + a, 6 + b, 6 + c) for c in range(2) for b in range(2) for a in range(2)]
    tree = smr.SMRTree((4, 4, 4), leaves, periodic=(True, True, True))
    t0 = time.time()
    u, b1, b2, b3 = pgen.divfree_blocks(smr.face_positions(tree, nx, ng, box), box, amp=0.2)
    m = agx.SMRMesh(tree, nx, ng, box, cfl=0.3)
    try:
        for name, a in zip(("u", "b1f", "b2f", "b3f"), (u, b1, b2, b3)):
            m.set(name, a)
        del u, b1, b2, b3
        m.initialize()
        m.cycle(1e30, 2)
        m.sync()
        log(f"[smr] {len(tree.leaves)} blocks of {nx}^3 on 3 levels set up in {time.time() - t0:.1f}s")
        steps = args.smr_steps
        t1 = time.perf_counter()
        m.cycle(1e30, steps)
        m.sync()
        sec = time.perf_counter() - t1
        m.m.timing(True)
        m.m.timing_reset()
        t2 = time.perf_counter()
        m.cycle(1e30, 2)
        m.sync()
        sec2 = (time.perf_counter() - t2) / 2
        per = {c: m.m.timing_get(c)[0] / 2 for c in KCLASSES}
        m.m.timing(False)
    finally:
        m.close()
    ncell = len(tree.leaves) * nx ** 3
    kern = sum(per.values())
    out = {"workload": f"{len(tree.leaves)} MeshBlocks of {nx}^3 on 3 refinement levels (4^3 root grid, central "
                       f"2^3 refined, one inner block refined again), periodic, smooth MHD state, "
                       f"VL2+PLM+HLLD+CT FP64, flux and EMF correction",
           "value": ncell * steps / sec, "unit": "zone-cycles/s", "ms_per_step": sec * 1e3 / steps,
           "cells": ncell, "steps": steps,
           "mhd_kernel_ms_per_cycle": {k: v for k, v in per.items() if v > 0},
           "refinement_and_gaps_ms_per_cycle": sec2 * 1e3 - kern}
    log(f"[smr] {steps} cycles in {sec:.3f}s -> {out['value']:.4e} zone-cycles/s; per cycle (timed run) "
        f"{sec2 * 1e3:.2f} ms: MHD kernels {kern:.2f} ms, refinement operators and gaps "
        f"{out['refinement_and_gaps_ms_per_cycle']:.2f} ms")
    return out


def ref_bench(agx, args):
    """Refinement operators (static mesh refinement) on 512 blocks of 32^3 (ng 2, 5
    variables, coarse buffers 16^3 + 2 ng): whole-block cell restriction and prolongation,
    and the interior face prolongation; algorithmic bytes per launch over the mean launch
    time of `reps` launches (device-resident inputs)."""
    nb, nx, ng, nvar, reps = 512, 32, 2, 5, 20
    n, nc = nx + 2 * ng, nx // 2 + 2 * ng
    g = agx.RefGeom(nx, nx, nx, ng, ng, 0)
    # uniform coordinates per block (the operators read per-block tables)
    h = 1.0 / 256
    tab_f = {k: np.tile((np.arange(n) - ng + 0.5) * h, (nb, 1)) for k in ("x1v", "x2v", "x3v")}
    tab_f.update({k: np.full((nb, n), h) for k in ("dx1f", "dx2f", "dx3f")})
    tab_c = {k: np.tile((np.arange(nc) - ng + 0.5) * 2 * h, (nb, 1)) for k in ("x1v", "x2v", "x3v")}
    tab_c.update({k: np.full((nb, nc), 2 * h) for k in ("dx1f", "dx2f", "dx3f")})
    fc, keep_f = agx.device_coords(tab_f)
    cc, keep_c = agx.device_coords(tab_c)
    rs = np.random.default_rng(1)
    fine = agx.DeviceArray(rs.random((nb, nvar, n, n, n)))
    coarse = agx.DeviceArray(rs.random((nb, nvar, nc, nc, nc)))
    b = [agx.DeviceArray(rs.random((nb, n + (d == 2), n + (d == 1), n + (d == 0)))) for d in range(3)]
    lo, hi = ng, ng + nx // 2 - 1
    items = [agx.RefItem(q, q, lo, hi, lo, hi, lo, hi, 0, 0, 0, 0) for q in range(nb)]
    ncoarse = nb * (nx // 2) ** 3
    cases = [("restrict_cc", lambda L: L.restrict_cc(nvar, fine.ptr, fc, coarse.ptr),
              ncoarse * nvar * (8 * 8 + 8)),
             ("prolong_cc", lambda L: L.prolong_cc(nvar, coarse.ptr, cc, fine.ptr, fc),
              ncoarse * nvar * (8 + 8 * 8)),
             # 24 outer faces read (each shared by two groups: 12 per group), 12 written
             ("prolong_fc_int", lambda L: L.prolong_fc_internal(b[0].ptr, b[1].ptr, b[2].ptr, fc),
              ncoarse * (12 + 12) * 8)]
    out = {"workload": f"{nb} blocks of {nx}^3 (nghost {ng}), {nvar} variables, whole-block ranges"}
    for name, call, alg in cases:
        L = agx.RefList(g, name, items, nb, nb)
        call(L)
        agx.RefList.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            call(L)
        agx.RefList.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        L.close()
        gbs = alg / (ms * 1e-3) / 1e9
        out[name] = {"ms_per_launch": ms, "alg_bytes": alg, "achieved_GBs": gbs, "frac": gbs * 1e9 / HBM_PEAK}
        log(f"[ref] {name}: {ms:.3f} ms per launch, {alg / 1e6:.0f} MB algorithmic -> {gbs:.0f} GB/s "
            f"({100 * gbs * 1e9 / HBM_PEAK:.0f}% of 8 TB/s)")
    for a in [fine, coarse] + b + keep_f + keep_c:
        a.close()
    out["smr"] = smr_bench(agx, reps)
    return out


def _box_cells(d):
    return (d["ei"] - d["si"] + 1) * (d["ej"] - d["sj"] + 1) * (d["ek"] - d["sk"] + 1)


def smr_bench(agx, reps):
    """Whole-mesh SMR steps on a two-level mesh: 4^3 root blocks of 32^3 cells (ng 2) on
    the unit cube with periodic boundaries, the central 2^3 root blocks refined (56 + 64
    = 120 blocks).  Times one cell-centred ghost exchange with prolongation (5 variables;
    same-level copies, restriction and copy to coarser blocks, copy to finer blocks'
    coarse buffers, same-level ghost restriction, prolongation: 6 launches) and one EMF
    correction (pack, set, clear, set from finer, average: 6 launches).  Algorithmic
    bytes: every value moved is read and written once (8 B each), restriction reads 8
    fine values per coarse one, prolongation writes 8 fine values per coarse one."""
    import importlib.util
    spec = importlib.util.spec_from_file_location("agx_smr", os.path.join(ROOT, "athena-gamma_amd", "smr.py"))
    smr = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(smr)
    nx, ng, nvar = 32, 2, 5
    refine = {(x, y, z) for x in (1, 2) for y in (1, 2) for z in (1, 2)}
    leaves = [(0, x, y, z) for z in range(4) for y in range(4) for x in range(4) if (x, y, z) not in refine]
    for (x, y, z) in sorted(refine):
        leaves += [(1, 2 * x + a, 2 * y + b, 2 * z + c) for c in range(2) for b in range(2) for a in range(2)]
    t = smr.SMRTree((4, 4, 4), leaves)
    nb = len(leaves)
    n, nc = nx + 2 * ng, nx // 2 + 2 * ng
    g = agx.RefGeom(nx, nx, nx, ng, ng, 0)
    fc, k1 = agx.device_coords(smr.block_coords(t, nx, ng))
    cc, k2 = agx.device_coords(smr.block_coords(t, nx, ng, coarse=True))
    rs = np.random.default_rng(5)
    u = agx.DeviceArray(rs.random((nb, nvar, n, n, n)))
    cu = agx.DeviceArray(rs.random((nb, nvar, nc, nc, nc)))
    e = [agx.DeviceArray(rs.standard_normal(s)) for s in
         ((nb, n + 1, n + 1, n), (nb, n + 1, n, n + 1), (nb, n, n + 1, n + 1))]
    keep = k1 + k2 + [u, cu] + e
    out = {"workload": f"{nb} blocks of {nx}^3 (ng {ng}) on two levels, 4^3 root blocks with the central 2^3 refined"}

    def timed(run, sync=agx.RefList.sync):
        run()
        sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            run()
        sync()
        return (time.perf_counter() - t0) * 1e3 / reps

    # cell-centred ghost exchange
    p = smr.cc_plan(t, nx, ng)
    L = lambda kind, items: agx.RefList(g, kind, [agx.RefItem(**d) for d in items], nb, nb)   # noqa: E731
    ls = [L("copy_cc_bb", p.copy_same), L("restrict_cc", p.restrict_send), L("copy_cc_cb", p.copy_from_finer),
          L("copy_cc_bc", p.copy_to_finer), L("restrict_cc", p.restrict_same), L("prolong_cc", p.prolong)]

    def cc_run():
        ls[0].copy_cc(nvar, u.ptr, u.ptr)
        ls[1].restrict_cc(nvar, u.ptr, fc, cu.ptr)
        ls[2].copy_cc(nvar, cu.ptr, u.ptr)
        ls[3].copy_cc(nvar, u.ptr, cu.ptr)
        ls[4].restrict_cc(nvar, u.ptr, fc, cu.ptr)
        ls[5].prolong_cc(nvar, cu.ptr, cc, u.ptr, fc)
    ms = timed(cc_run)
    copies = sum(_box_cells(d) for d in p.copy_same + p.copy_from_finer + p.copy_to_finer)
    restr = sum(_box_cells(d) for d in p.restrict_send + p.restrict_same)
    prol = sum(_box_cells(d) for d in p.prolong)
    alg = nvar * 8 * (2 * copies + 9 * restr + 9 * prol)
    ghosts = nvar * (copies - sum(_box_cells(d) for d in p.copy_to_finer) + 8 * prol)
    out["cc_exchange"] = {"ms": ms, "launches": 6, "alg_bytes": alg, "achieved_GBs": alg / (ms * 1e-3) / 1e9,
                          "frac": alg / (ms * 1e-3) / HBM_PEAK, "ghost_values": ghosts}
    log(f"[smr] cell ghost exchange + prolongation, {nb} blocks x {nvar} vars: {ms:.3f} ms "
        f"({alg / 1e6:.0f} MB algorithmic, {alg / (ms * 1e-3) / 1e9:.0f} GB/s)")
    for x in ls:
        x.close()
    # EMF correction
    q = smr.emf_plan(t)
    ss = max(agx.ref_emf_count_same_level(g, d["fid"]) for d in q.same_level)
    sc = max(agx.ref_emf_count(g, d["fid"]) for d in q.to_coarser)
    bs = agx.DeviceArray(np.zeros((len(q.same_level), ss)))
    bc = agx.DeviceArray(np.zeros((len(q.to_coarser), sc)))
    keep += [bs, bc]
    le = [L("emf_same_level", q.same_level), L("emf_to_coarser", q.to_coarser),
          agx.RefList(g, "emf_set_same", [agx.RefItem(**d) for d in q.set_same], len(q.same_level), nb),
          L("emf_clear", q.clear),
          agx.RefList(g, "emf_set_finer", [agx.RefItem(**d) for d in q.set_finer], len(q.to_coarser), nb),
          L("emf_average", q.average)]
    ep = [x.ptr for x in e]

    def emf_run():
        le[0].emf_same_level(*ep, bs.ptr, ss)
        le[1].emf_to_coarser(*ep, fc, bc.ptr, sc)
        le[2].emf_set(*ep, bs.ptr, ss)
        le[3].emf_clear(*ep)
        le[4].emf_set(*ep, bc.ptr, sc)
        le[5].emf_average(*ep)
    ms = timed(emf_run)
    nsame = sum(agx.ref_emf_count_same_level(g, d["fid"]) for d in q.same_level)
    ncoarse = sum(agx.ref_emf_count(g, d["fid"]) for d in q.to_coarser)
    face_in = 2 * (nx - 1) * nx          # EMFs inside one face, not on its own edges
    nclear = sum(face_in if d["fid"] < 6 else nx for d in q.clear)
    navg = sum((face_in if d["fi1"] == 0 else 2 * nx) if d["fid"] < 6 else nx for d in q.average)
    # pack same level: read + write; pack to coarser: 2 reads + write; set: buffer read,
    # EMF read + write; clear: write; average: read + write (8 B each)
    alg = 8 * (2 * nsame + 3 * ncoarse + 3 * (nsame + ncoarse) + nclear + 2 * navg)
    out["emf_correction"] = {"ms": ms, "launches": "6 (the set lists run as one launch per class and quadrant)",
                             "values_exchanged": nsame + ncoarse, "alg_bytes": alg,
                             "achieved_GBs": alg / (ms * 1e-3) / 1e9, "frac": alg / (ms * 1e-3) / HBM_PEAK}
    log(f"[smr] EMF correction, {nb} blocks: {ms:.3f} ms ({nsame + ncoarse} values exchanged)")
    for x in le:
        x.close()
    for a in keep:
        a.close()
    # the whole MHD boundary step through the package's driver: exchanges of u and b
    # across levels, then ProlongateBoundaries per slot of coarser neighbours
    sb = agx.SMRBoundary(t, nx, ng, 5.0 / 3.0, 1e-8, 1e-8)
    A = dict(u=rs.random((nb, 5, n, n, n)) + 1.0, w=np.zeros((nb, 5, n, n, n)), bcc=np.zeros((nb, 3, n, n, n)),
             cu=np.ones((nb, 5, nc, nc, nc)), cw=np.zeros((nb, 5, nc, nc, nc)), cbcc=np.zeros((nb, 3, nc, nc, nc)))
    D = {k: agx.DeviceArray(v) for k, v in A.items()}
    b = [agx.DeviceArray(rs.standard_normal(s_)) for s_ in ((nb, n, n, n + 1), (nb, n, n + 1, n), (nb, n + 1, n, n))]
    cb = [agx.DeviceArray(np.zeros(s_)) for s_ in ((nb, nc, nc, nc + 1), (nb, nc, nc + 1, nc), (nb, nc + 1, nc, nc))]
    P = {k: v.ptr for k, v in D.items()}
    args_ = (P["u"], [x.ptr for x in b], P["w"], P["bcc"], P["cu"], [x.ptr for x in cb], P["cw"], P["cbcc"])
    ms = timed(lambda: sb.boundary_step(*args_), sb.sync)
    ms_unfused = timed(lambda: sb.boundary_step(*args_, fused=False), sb.sync)
    sb.capture(*args_)
    msg = timed(sb.replay, sb.sync)
    out["mhd_boundary_step"] = {"ms": ms, "ms_one_launch_per_list": ms_unfused, "ms_graph_replay": msg,
                                "slots": len(sb.slots),
                                "what": "u (5 variables) and b exchanged across levels, then ProlongateBoundaries "
                                        "(restriction, cons->prim on the coarse buffer, prolongation, prim->cons)"}
    log(f"[smr] MHD boundary step (exchanges + prolongation, {len(sb.slots)} slots), {nb} blocks: {ms:.3f} ms "
        f"(fused slot launches; {ms_unfused:.3f} ms with one launch per list), {msg:.3f} ms as one hipGraph replay")
    sb.close()
    for a in list(D.values()) + b + cb:
        a.close()
    out["three_levels"] = smr_three_levels(agx, smr, reps, timed)
    return out


def smr_three_levels(agx, smr, reps, timed):
    """The MHD boundary step and the EMF correction on a three-level mesh of 16^3 blocks:
    8^3 root blocks, the central 2^3 refined, the central 2^3 of those refined again
    (624 blocks; levels 0, 1, 2), through SMRBoundary."""
    nx, ng = 16, 2
    r1 = {(x, y, z) for x in (3, 4) for y in (3, 4) for z in (3, 4)}
    r2 = {(x, y, z) for x in (7, 8) for y in (7, 8) for z in (7, 8)}
    leaves = [(0, x, y, z) for z in range(8) for y in range(8) for x in range(8) if (x, y, z) not in r1]
    for (x, y, z) in sorted(r1):
        leaves += [(1, 2 * x + a, 2 * y + b, 2 * z + c) for c in range(2) for b in range(2) for a in range(2)
                   if (2 * x + a, 2 * y + b, 2 * z + c) not in r2]
    for (x, y, z) in sorted(r2):
        leaves += [(2, 2 * x + a, 2 * y + b, 2 * z + c) for c in range(2) for b in range(2) for a in range(2)]
    t0 = time.time()
    t = smr.SMRTree((8, 8, 8), leaves)
    nb = len(leaves)
    sb = agx.SMRBoundary(t, nx, ng, 5.0 / 3.0, 1e-8, 1e-8)
    setup = time.time() - t0
    n, nc = nx + 2 * ng, nx // 2 + 2 * ng
    rs = np.random.default_rng(7)
    D = dict(u=agx.DeviceArray(rs.random((nb, 5, n, n, n)) + 1.0), w=agx.DeviceArray(np.zeros((nb, 5, n, n, n))),
             bcc=agx.DeviceArray(np.zeros((nb, 3, n, n, n))), cu=agx.DeviceArray(np.ones((nb, 5, nc, nc, nc))),
             cw=agx.DeviceArray(np.zeros((nb, 5, nc, nc, nc))), cbcc=agx.DeviceArray(np.zeros((nb, 3, nc, nc, nc))))
    b = [agx.DeviceArray(rs.standard_normal(s_)) for s_ in ((nb, n, n, n + 1), (nb, n, n + 1, n), (nb, n + 1, n, n))]
    cb = [agx.DeviceArray(np.zeros(s_)) for s_ in ((nb, nc, nc, nc + 1), (nb, nc, nc + 1, nc), (nb, nc + 1, nc, nc))]
    e = [agx.DeviceArray(rs.standard_normal(s_)) for s_ in ((nb, n + 1, n + 1, n), (nb, n + 1, n, n + 1),
                                                           (nb, n, n + 1, n + 1))]
    P = {k: v.ptr for k, v in D.items()}
    args_ = (P["u"], [x.ptr for x in b], P["w"], P["bcc"], P["cu"], [x.ptr for x in cb], P["cw"], P["cbcc"])
    ms = timed(lambda: sb.boundary_step(*args_), sb.sync)
    ms_emf = timed(lambda: sb.correct_emf(*[x.ptr for x in e]), sb.sync)
    out = {"workload": f"{nb} blocks of {nx}^3 (ng {ng}) on three levels: 8^3 root blocks, the central 2^3 "
                       f"refined, the central 2^3 of those refined again", "blocks": nb,
           "mhd_boundary_step_ms": ms, "emf_correction_ms": ms_emf, "slots": len(sb.slots),
           "host_setup_s": setup}
    log(f"[smr] three levels, {nb} blocks of {nx}^3: MHD boundary step {ms:.3f} ms, EMF correction "
        f"{ms_emf:.3f} ms (tree, plans and lists built in {setup:.1f} s)")
    sb.close()
    for a in list(D.values()) + b + cb + e:
        a.close()
    return out


def host_cpu():
    """The host's CPU as lscpu reports it: model, sockets x physical cores per socket x
    threads per core, and the logical CPUs this process may use."""
    info = {}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            info[k.strip()] = v.strip()
    except Exception:   # lscpu missing: report the logical count only
        pass
    usable = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return {"model": info.get("Model name"), "sockets": info.get("Socket(s)"),
            "cores_per_socket": info.get("Core(s) per socket"), "threads_per_core": info.get("Thread(s) per core"),
            "logical_cpus": os.cpu_count(), "usable_cpus": usable}


def cpu_baseline(args):
    """The CPU oracle (FP64 restatement, OpenMP over MeshBlocks) on a bounded sample of
    the same workload: a 64^3 periodic sub-box of the turbulence problem in 16
    MeshBlocks of 32x32x16, timed over a fixed number of VL2 cycles."""
    sys.path.insert(0, ROOT)
    from oracle.orc import Oracle, make_desc
    agx, pgen = load_pkg()
    threads = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(threads))
    desc = make_desc(nx=(64, 64, 64), bnx=(32, 32, 16), box=((0.0, 0.25),) * 3, cfl=0.4,
                     gamma=5.0 / 3.0)
    U, b1, b2, b3 = pgen.turbulence(desc, seed=2, mach=2.0, bx=0.1)
    o = Oracle(desc)
    o.set("u", agx.scatter_cells(desc, U))
    f1, f2, f3 = pgen.scatter_faces(desc, b1, b2, b3)
    o.set("b1f", f1)
    o.set("b2f", f2)
    o.set("b3f", f3)
    o.initialize()
    o.cycle(1e30)
    ncyc = args.cpu_cycles
    t0 = time.perf_counter()
    for _ in range(ncyc):
        o.cycle(1e30)
    sec = time.perf_counter() - t0
    zc = 64.0 ** 3 * ncyc / sec
    used = int(os.environ.get("OMP_NUM_THREADS", threads))
    host = host_cpu()
    log(f"[bench] CPU oracle: {ncyc} cycles of 64^3 in {sec:.1f}s -> {zc:.3e} zone-cycles/s ({used} threads; "
        f"host {host})")
    return {"value": zc, "unit": "zone-cycles/s", "cores": used, "host": host,
            "build": "gcc -O3 -fopenmp -ffp-contract=off (oracle/Makefile)",
            "kind": "port",
            "sample": f"{ncyc} VL2+PLM+HLLD+CT cycles of a 64^3 periodic sub-box in 16 MeshBlocks of 32x32x16 "
                      f"(same turbulence generator), oracle/agx_oracle.c FP64, OpenMP over MeshBlocks"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--block", type=int, default=64)
    ap.add_argument("--mg-cycles", type=int, default=20)
    ap.add_argument("--cpu-cycles", type=int, default=600)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mg", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-sg", action="store_true",
                    help="skip the 256^3 MHD + self-gravity leg (FP64 and mixed)")
    ap.add_argument("--sg-steps", type=int, default=10)
    ap.add_argument("--no-ref", action="store_true", help="skip the refinement-operator leg")
    ap.add_argument("--no-smr", action="store_true", help="skip the three-level static-refinement MHD leg")
    ap.add_argument("--smr-block", type=int, default=32)
    ap.add_argument("--smr-steps", type=int, default=10)
    ap.add_argument("--mg-only", action="store_true",
                    help="only the 256^3 multigrid legs (periodic and multipole, FP64 and mixed)")
    ap.add_argument("--smr-only", action="store_true",
                    help="run only the static-refinement leg and print its JSON object (for profiling)")
    ap.add_argument("--no-compare", dest="compare", action="store_false",
                    help="skip downloading w and B after both runs (mixed-vs-FP64 difference, div B)")
    ap.add_argument("--large", action="store_true",
                    help="run the 512^3 PPM + self-gravity mixed-precision case instead")
    ap.add_argument("--large-n", type=int, default=512)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device three times (scale for rocprofv3 "
                         "FETCH_SIZE / WRITE_SIZE: bytes per dispatch of a known copy)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        x = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        for _ in range(3):
            y.copy_(x)
        torch.cuda.synchronize()
        log("[bench] copy probe: 3 x 1 GiB device-to-device copies")
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (one MI355X per run); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    agx, pgen = load_pkg()
    if args.smr_only:
        print(json.dumps(smr_cycle_bench(agx, pgen, args)), flush=True)
        return
    if args.mg_only:
        res = {}
        for form in ("fp64", "mixed", "multipole", "multipole_mixed"):
            vps, iters, norm, t_solve, per = mg_bench(agx, pgen, args, form)
            res[form] = {"vcycles_per_s": vps, "solve_ms_to_1e-10": t_solve * 1e3, "cycles_after_fmg": iters,
                         "final_defect": norm, "kernel_ms_per_cycle": {k: v[0] / 3 for k, v in per.items()}}
        print(json.dumps({"mg": res}), flush=True)
        return
    if args.large:
        large_bench(agx, pgen, args)
        return
    desc, zc, ms_step, per, st64 = mhd_bench(agx, pgen, args, "fp64")
    rf = roofline(desc, per, "fp64", mhd_bench.tile)
    log("[bench] fp64 per-cycle kernel time: " + ", ".join(f"{k} {v[0] / 2:.2f} ms" for k, v in per.items())
        + f" (sum {sum(v[0] for v in per.values()) / 2:.2f} ms/cycle)")
    out = {
        "metric": "MHD zone-cycles/s + MG V-cycles/s at 256^3 on 1 MI355X; % of HBM roofline",
        "value": zc,
        "unit": "zone-cycles/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{args.n}^3 decaying MHD turbulence (rms Mach 2, Bx=0.1, k^-4 on 2<=k<=8), "
                               f"{(args.n // args.block) ** 3} MeshBlocks of {args.block}^3, "
                               f"VL2+PLM+HLLD+CT, plain FP64, one step = one VL2 cycle",
                   "mesh": [args.n] * 3, "meshblock": [args.block] * 3, "nghost": 2},
        "cycle_roofline_frac": zc * CYCLE_BYTES["fp64"] / HBM_PEAK,
        "stage_form": STAGE_FORMS[mhd_bench.tile],
        "roofline": rf,
        "kernel_ms_per_cycle": {k: v[0] / 2 for k, v in per.items()},
    }
    if not args.no_mixed:
        mdesc, mzc, mms, mper, stmx = mhd_bench(agx, pgen, args, "mixed")
        mtile = mhd_bench.tile
        log("[bench] mixed per-cycle kernel time: " + ", ".join(f"{k} {v[0] / 2:.2f} ms" for k, v in mper.items())
            + f" (sum {sum(v[0] for v in mper.values()) / 2:.2f} ms/cycle)")
        mixed = {"value": mzc, "unit": "zone-cycles/s", "ms_per_step": mms, "vs_fp64": mzc / zc,
                 "dtype": "f32 reconstruction+Riemann, f64 update/C2P/CT",
                 "cycle_roofline_frac": mzc * CYCLE_BYTES["mixed"] / HBM_PEAK,
                 "stage_form": STAGE_FORMS[mtile],
                 "roofline": roofline(mdesc, mper, "mixed", mtile),
                 "kernel_ms_per_cycle": {k: v[0] / 2 for k, v in mper.items()}}
        if st64 is not None and stmx is not None:
            (t64, w64, db64), (tmx, wmx, dbmx) = st64, stmx
            out["max_divb"], out["max_divb_dx_over_B"] = db64
            mixed["max_divb"], mixed["max_divb_dx_over_B"] = dbmx
            mixed["same_cycle_count"] = bool(t64[2] == tmx[2])
            mixed["rel_err_rho"] = float(np.abs(wmx[:, 0] - w64[:, 0]).max() / np.abs(w64[:, 0]).max())
            mixed["rel_err_p"] = float(np.abs(wmx[:, 4] - w64[:, 4]).max() / np.abs(w64[:, 4]).max())
            mixed["compare_note"] = (f"max-norm relative difference of rho and P vs the FP64 GPU path "
                                     f"(bitwise equal to the oracle) after {int(t64[2])} cycles")
            log(f"[bench] mixed vs fp64 after {int(t64[2])} cycles: rho {mixed['rel_err_rho']:.3e}, "
                f"P {mixed['rel_err_p']:.3e}")
        out["mixed"] = mixed
    if not args.no_mg:
        vps, iters, norm, t_solve, mper = mg_bench(agx, pgen, args)
        out["mg_vcycles_per_s"] = vps
        out["mg_roofline_frac"] = vps * float(args.n) ** 3 * 148.6 / HBM_PEAK
        out["mg_solve_ms_to_1e-10"] = t_solve * 1e3
        out["mg_solve_vcycles_after_fmg"] = iters
        out["mg_kernel_ms_per_vcycle"] = {k: v[0] / 3 for k, v in mper.items()}
        if not args.no_mixed:
            vps32, it32, norm32, ts32, mper32 = mg_bench(agx, pgen, args, "mixed")
            out["mg_mixed"] = {
                "vcycles_per_s": vps32, "vs_fp64": vps32 / vps,
                "note": "one cycle = FP64 defect + FP32 V(1,1) + FP64 correction",
                "roofline_frac": vps32 * float(args.n) ** 3 * 114.3 / HBM_PEAK,
                "solve_ms_to_1e-10": ts32 * 1e3, "cycles_after_fmg": it32, "final_defect": norm32,
                "kernel_ms_per_cycle": {k: v[0] / 3 for k, v in mper32.items()}}
        vmp, itmp, nmp, tsmp, mpper = mg_bench(agx, pgen, args, "multipole")
        out["mg_multipole"] = {
            "vcycles_per_s": vmp, "solve_ms_to_1e-10": tsmp * 1e3, "cycles_after_fmg": itmp,
            "final_defect": nmp,
            "note": "isolated blobs, multipole order-4 boundaries about the centre of mass, FAS V(1,1), FP64",
            "kernel_ms_per_vcycle": {k: v[0] / 3 for k, v in mpper.items()}}
        vmq, itmq, nmq, tsmq, mqper = mg_bench(agx, pgen, args, "multipole_mixed")
        out["mg_multipole_mixed"] = {
            "vcycles_per_s": vmq, "vs_fp64": vmq / vmp, "solve_ms_to_1e-10": tsmq * 1e3,
            "cycles_after_fmg": itmq, "final_defect": nmq,
            "note": "the same problem; one cycle = FP64 ghosts + defect, FP32 FAS V(1,1) on the "
                    "correction, FP64 correction",
            "kernel_ms_per_cycle": {k: v[0] / 3 for k, v in mqper.items()}}
    if not args.no_ref:
        out["refinement"] = ref_bench(agx, args)
    if not args.no_smr:
        out["smr_cycle"] = smr_cycle_bench(agx, pgen, args)
    if not args.no_sg:
        out["self_gravity"] = sg_run(agx, pgen, args.n, args.block, 2, 2.0, args.sg_steps, "sg", bx=0.1)
        nv = max(1, int(np.ceil(out["self_gravity"]["fp64"]["vcycles_per_solve"])))
        out["self_gravity_graph"] = sg_graph_bench(agx, pgen, args.n, args.block, args.sg_steps, nv)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
