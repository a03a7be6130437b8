"""The one-launch VL2 stage (k_vl2: fluxes in LDS) against the split form (three flux
launches + the stage kernel), which the other parity tests hold to the oracle bit for
bit.  Both forms form every value with the same operations in the same order, so every
array must be identical bit for bit after whole cycles: all Riemann solvers and
reconstruction orders, both tile shapes, both EOS, FP64 and the mixed form, with
self-gravity, on ragged MeshBlocks (tiles cut by the block edge) and on blocks smaller
than a tile."""
import numpy as np
import pytest

from oracle.orc import make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu

FIELDS = ("u", "w", "bcc", "b1f", "b2f", "b3f")


def run(agx, desc, state, tile, ncyc=3, phi=False):
    m = agx.Mesh(desc)
    u, f1, f2, f3 = state
    # the isothermal C ABI carries NHYDRO = 4 conserved variables
    load_state(m, (u[:, :4] if desc.eos == 1 else u, f1, f2, f3))
    m.set_tile(tile)
    m.initialize()
    assert m.cycle(1e9, ncyc) == ncyc
    out = {f: m.get(f) for f in FIELDS + (("phi",) if phi else ())}
    out["time"] = m.time()
    m.close()
    return out


def same(a, b, what):
    ok = np.array_equal(a, b)
    if not ok:
        d = np.abs(a - b)
        print(f"{what}: max |diff| {d.max():.3e} at {np.unravel_index(d.argmax(), d.shape)}", flush=True)
    assert ok, what


CASES = [
    # (name, desc kwargs, state kwargs)
    ("plm_hlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4), dict(mach=1.5)),
    ("plm_hlle", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=1), dict(mach=1.5)),
    ("plm_roe", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=2), dict(mach=1.5)),
    ("plm_lhlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, rsolver=3), dict(mach=1.5)),
    ("ppm_lhlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, xorder=3, nghost=3,
                       rsolver=3), dict(mach=1.5)),
    ("dc_hlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, xorder=1), dict(mach=1.5)),
    ("ppm_hlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, xorder=3, nghost=3),
     dict(mach=1.5)),
    ("ragged", dict(nx=(24, 16, 8), bnx=(8, 4, 8), box=((0, 1.5), (0, 1), (0, 0.5))), dict(mach=1.5)),
    ("wide_blocks", dict(nx=(96, 40, 16), bnx=(48, 20, 8), box=((0, 2.4), (0, 1), (0, 0.4)), cfl=0.4),
     dict(mach=1.5)),
    ("iso_hlld", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, eos=1, iso_cs=0.8),
     dict(mach=1.5, iso=True)),
    ("iso_hlle", dict(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, eos=1, iso_cs=0.8, rsolver=1),
     dict(mach=1.5, iso=True)),
    ("floors", dict(nx=(16, 16, 16), bnx=(8, 8, 8), box=((0, 1),) * 3, cfl=0.3),
     dict(p0=1e-19, vamp=0.3, drho=0.5)),
]


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("name,dk,sk", CASES, ids=[c[0] for c in CASES])
def test_one_launch_stage_equals_split(agx, name, dk, sk, precision):
    desc = make_desc(precision=precision, **dk)
    st = turbulent_state(desc, seed=17, **sk)
    ref = run(agx, desc, st, 1)          # k_flux x 3 + k_stage<32, 8>
    for tile in (4, 5):                  # k_vl2<16, 16>, k_vl2<32, 8>
        got = run(agx, desc, st, tile)
        assert got["time"] == ref["time"]
        for f in FIELDS:
            same(got[f], ref[f], f"{name} precision {precision} tile {tile}: {f}")


@pytest.mark.parametrize("precision", [0, 1])
def test_one_launch_stage_with_self_gravity(agx, precision):
    desc = make_desc(nx=(32, 32, 32), bnx=(16, 16, 16), box=((0, 1),) * 3, cfl=0.4, self_gravity=1,
                     four_pi_G=1.0, mg_mode=0, mg_threshold=1e-9, precision=precision, xorder=3, nghost=3)
    st = turbulent_state(desc, seed=21, mach=1.5, drho=0.5)
    ref = run(agx, desc, st, 1, phi=True)
    got = run(agx, desc, st, 5, phi=True)
    assert got["time"] == ref["time"]
    for f in FIELDS + ("phi",):
        same(got[f], ref[f], f"self-gravity precision {precision}: {f}")
