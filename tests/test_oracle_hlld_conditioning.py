"""Conditioning of the reference HLLD solver near its degenerate star states (FP64 oracle).

A uniform Bx at plasma beta 1 over a lognormal density (sigma_lnrho = 1) puts the
Alfven speed above the sound speed in most cells.  HLLD's star-state denominators
rho (S - u)(S - S_M) - Bx^2 (hydro/rsolvers/mhd/hlld.cpp:187,215) then come close to
zero wherever the transverse field is weak.  There the solver amplifies an input
perturbation at the FP32 rounding level (the initial state rounded to FP32 once, about
3e-8 relative) to more than 1e-5 in the max norm within a few cycles. HLLE on the same
state keeps it at the rounding level, and so does the mean over cells for HLLD.

This is a property of the reference algorithm in FP64. It bounds the cell-by-cell
agreement that any form reading FP32 inputs can reach on such states, which is why
the mixed form's error on them is stated as a mean over cells.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_agx, load_pgen


def run(desc, ub, faces, cycles):
    o = Oracle(desc)
    o.set("u", ub)
    for n, f in zip(("b1f", "b2f", "b3f"), faces):
        o.set(n, f)
    o.initialize()
    for _ in range(cycles):
        o.cycle(1e30)
    ng = desc.nghost
    return o.get("w")[:, :, ng:-ng, ng:-ng, ng:-ng]


@pytest.mark.parametrize("rsolver", [0, 1])
def test_hlld_amplifies_fp32_rounding_of_the_state(rsolver):
    agx, pgen = load_agx(), load_pgen()
    desc = make_desc(nx=(32, 32, 32), bnx=(32, 32, 32), box=((0.0, 1.0),) * 3, cfl=0.4, gamma=5.0 / 3.0,
                     xorder=2, rsolver=rsolver)
    U, b1, b2, b3 = pgen.turbulence_lognormal(desc, seed=5, mach=2.0, sigma_lnrho=1.0, beta=1.0)
    ub = agx.scatter_cells(desc, U)
    faces = pgen.scatter_faces(desc, b1, b2, b3)
    a = run(desc, ub, faces, 4)
    b = run(desc, ub.astype(np.float32).astype(np.float64), faces, 4)
    for name, v in (("rho", 0), ("P", 4)):
        diff = np.abs(a[:, v] - b[:, v])
        emax = diff.max() / np.abs(a[:, v]).max()
        emean = diff.mean() / np.abs(a[:, v]).mean()
        print(f"rsolver {rsolver} {name}: max {emax:.2e} mean {emean:.2e}", flush=True)
        assert emean <= 1e-6
        if name == "P":
            if rsolver == 0:   # HLLD
                assert emax >= 1e-5
            else:              # HLLE
                assert emax <= 1e-6
