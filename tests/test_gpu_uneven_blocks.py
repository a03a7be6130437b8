"""MeshBlocks of uneven sizes on one level (agx_upload_blocks / agx_download_blocks).

A 32^3 periodic mesh is held on the device as the uniform lattice of 8^3 blocks.  The
host side describes it as four 16^3 blocks (lower half in x3) and thirty-two 8^3 blocks
(upper half), each with its own ghost zones.  Cycles run from the uneven upload.  The
downloaded blocks, ghosts included, must equal bit for bit the oracle's run on the
uniform 8^3 decomposition mapped onto the same blocks.  In exact arithmetic the update
of a cell does not depend on how the mesh is cut into blocks, and the GPU path keeps that
in FP64.
"""
import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.states import load_state, turbulent_state

pytestmark = pytest.mark.gpu
FDIR = {"u": None, "w": None, "bcc": None, "b1f": 0, "b2f": 1, "b3f": 2}


def layout():
    blocks = [((i, j, 0), (16, 16, 16)) for j in (0, 16) for i in (0, 16)]
    blocks += [((i, j, k), (8, 8, 8)) for k in (16, 24) for j in range(0, 32, 8) for i in range(0, 32, 8)]
    return blocks


def to_block(lat, desc, origin, size, fdir):
    """Host block (ghosts included) from lattice block arrays, by global index with
    periodic wrap; a face on a lattice-block boundary is taken from the upper block."""
    ng = desc.nghost
    bn = (desc.bnx1, desc.bnx2, desc.bnx3)
    N = (desc.nx1, desc.nx2, desc.nx3)
    nl = [N[d] // bn[d] for d in range(3)]
    lb, lc = [], []
    for d in range(3):
        n = size[d] + 2 * ng + (1 if fdir == d else 0)
        I = (origin[d] + np.arange(n) - ng) % N[d]
        lb.append(I // bn[d])
        lc.append(I % bn[d] + ng)
    B = (lb[2][:, None, None] * nl[1] + lb[1][None, :, None]) * nl[0] + lb[0][None, None, :]
    K, J, I = lc[2][:, None, None], lc[1][None, :, None], lc[0][None, None, :]
    if lat.ndim == 5:
        out = lat[B, :, K, J, I]                  # (n3, n2, n1, nvar)
        return np.ascontiguousarray(np.moveaxis(out, -1, 0))
    return np.ascontiguousarray(lat[B, K, J, I])


def test_uneven_blocks_match_uniform_lattice(agx):
    desc = make_desc(nx=(32, 32, 32), bnx=(8, 8, 8), box=((0, 1),) * 3, cfl=0.4)
    st = turbulent_state(desc, seed=9, mach=1.5)
    o = Oracle(desc)
    load_state(o, st)
    o.initialize()
    blocks = layout()
    m = agx.Mesh(desc)
    for name in ("u", "b1f", "b2f", "b3f"):
        lat = o.get(name)
        m.set_blocks(name, blocks, [to_block(lat, desc, org, sz, FDIR[name]) for org, sz in blocks])
    # round trip before any update: interiors back exactly, ghosts from the neighbours
    shapes = {n: [to_block(o.get(n), desc, org, sz, FDIR[n]).shape for org, sz in blocks] for n in FDIR}
    got = m.get_blocks("u", blocks, shapes["u"])
    for (org, sz), a in zip(blocks, got):
        assert np.array_equal(a, to_block(o.get("u"), desc, org, sz, None))
    m.initialize()
    for _ in range(3):
        o.cycle(1e9)
    assert m.cycle(1e9, 3) == 3
    assert m.time() == o.time()
    for name in FDIR:
        lat = o.get(name)
        got = m.get_blocks(name, blocks, shapes[name])
        nbad = 0
        for (org, sz), a in zip(blocks, got):
            ref = to_block(lat, desc, org, sz, FDIR[name])
            nbad += int(not np.array_equal(a, ref))
        print(f"{name}: {len(blocks) - nbad}/{len(blocks)} blocks bitwise equal", flush=True)
        assert nbad == 0, name


def test_uneven_blocks_rejects_misaligned(agx):
    desc = make_desc(nx=(32, 32, 32), bnx=(8, 8, 8), box=((0, 1),) * 3)
    m = agx.Mesh(desc)
    with pytest.raises(agx.AgxError):
        m.set_blocks("u", [((4, 0, 0), (8, 8, 8))], [np.zeros((5, 12, 12, 12))])
    with pytest.raises(agx.AgxError):
        m.set_blocks("u", [((24, 0, 0), (16, 8, 8))], [np.zeros((5, 12, 12, 20))])
    with pytest.raises(agx.AgxError):
        m.set_blocks("fx1", [((0, 0, 0), (8, 8, 8))], [np.zeros((5, 12, 12, 13))])
