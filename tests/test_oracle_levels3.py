"""The SMR pins on a three-level mesh (4^3 root blocks of 8^3, the central 2^3 refined,
one inner level-1 block refined again: 127 blocks, levels 0, 1, 2, 2:1 balanced),
through the host module's plans (athena-gamma_amd/smr.py) and the oracle:
  - EMF correction: one value per shared edge on every level, and every edge covered by
    edges of the next finer level equals their mean;
  - cell ghost exchange with prolongation: a linear field exact in every ghost cell,
    every prolongation stencil filled;
  - face ghost exchange with prolongation: a linear divergence-free field exact on every
    face;
  - MHD ghost prolongation: per-slot batches equal the reference's neighbour order bit
    for bit with floors active.
"""
import numpy as np

from oracle import orc
from tests import emf_cases as ec
from tests import ghost_cases as gc
from tests.conftest import load_smr

NX, NG = 8, 2


def mesh():
    t = gc.tree(*gc.THREE)
    assert len(t.leaves) == 127 and t.maxlev == 2
    return t, gc.coords(t, NX, NG), gc.coords(t, NX, NG, coarse=True)


def test_emf_consistent_three_levels():
    t, ftab, _ = mesh()
    E = ec.random_emfs(len(t.leaves), NX, NG, seed=31)
    out, _ = ec.run_oracle(orc, t, E, NX, NG, load_smr().block_coords(t, NX, NG))
    spread, err, ncmp = ec.check_consistency(t, out, NX, NG)
    print(f"spread {spread:.2e}, coarse vs fine {err:.2e} over {ncmp} edges", flush=True)
    # level 0 over the 2^3 refined roots (4 x 4 faces of 2 x 9 x 8 edges less shared
    # ones) and level 1 over the level-2 block
    assert ncmp > 2000 and spread < 4e-15 and err < 4e-15


def test_cell_ghosts_linear_three_levels():
    t, ftab, ctab = mesh()
    exact = gc.linear(t, ftab, 2)
    u = exact.copy()
    gc.poison_exchanged_ghosts(t, u, NX, NG)
    cu = np.full((len(t.leaves), 2) + (NX // 2 + 2 * NG,) * 3, np.nan)
    out, _ = gc.run_oracle(orc, t, u, cu, NX, NG, ftab, ctab, 2)
    assert not gc.run_oracle.uncovered.any()
    assert not np.isnan(out).any() and np.array_equal(out, exact)


def test_face_ghosts_linear_three_levels():
    t, ftab, ctab = mesh()
    exact = gc.linear_faces(t, NX, NG, ftab)
    f = [a.copy() for a in exact]
    gc.poison_face_ghosts(t, f, NX, NG)
    cf = [np.full(a.shape, np.nan) for a in gc.linear_faces(t, NX, NG, ctab, coarse=True)]
    cov = []
    out, _ = gc.run_oracle_fc(orc, t, f, cf, NX, NG, ftab, ctab, coverage=cov)
    assert not any(cov)
    for c in range(3):
        assert not np.isnan(out[c]).any() and np.array_equal(out[c], exact[c]), c


def test_mhd_slot_order_three_levels():
    t, ftab, ctab = mesh()
    A = gc.mhd_state(t, NX, NG, seed=32)
    a = gc.run_oracle_mhd(orc, t, A, NX, NG, ftab, ctab, order="reference")
    b = gc.run_oracle_mhd(orc, t, A, NX, NG, ftab, ctab, order="batched")
    for k in a:
        x, y = (a[k], b[k]) if isinstance(a[k], list) else ([a[k]], [b[k]])
        for xx, yy in zip(x, y):
            assert np.array_equal(xx, yy), k
