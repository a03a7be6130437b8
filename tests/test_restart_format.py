"""Restart files in the reference's layout (outputs/restart.cpp:37-207; read back as
mesh/mesh.cpp:551-700 does), through the library's host functions (no GPU needed).

The file is parsed here independently with struct/numpy from the layout of
restart.cpp: the parameter text up to "<par_end>\\n"; int nbtotal, int root_level,
RegionSize (9 doubles, 3 ints, padded to 88 bytes), double time, double dt, int ncycle,
uint64 datasize; nbtotal x (LogicalLocation: 3 int64 + int padded to 32 bytes, double
cost); then per MeshBlock u, b.x1f, b.x2f, b.x3f.  Checks: every field, the Z order of
the blocks (MeshBlockTree::GetMeshBlockList: Morton order with x1 the fastest bit, which
for a root grid that is not a power of two per side skips the missing leaves), the data
of each block, the round trip, and the refusals (another mesh, truncated file).
"""
import struct

import numpy as np
import pytest

from tests.conftest import load_agx


def desc_of(nx, bnx, eos=0, ng=2):
    agx = load_agx()
    return agx.make_desc(nx=nx, bnx=bnx, box=((-1.0, 1.0), (0.0, 0.5), (0.0, 0.75)), eos=eos, nghost=ng, cfl=0.3)


def shapes(d):
    nb = (d.nx1 // d.bnx1) * (d.nx2 // d.bnx2) * (d.nx3 // d.bnx3)
    n1, n2, n3 = (d.bnx1 + 2 * d.nghost, d.bnx2 + 2 * d.nghost, d.bnx3 + 2 * d.nghost)
    nh = 4 if d.eos == 1 else 5
    return [(nb, nh, n3, n2, n1), (nb, n3, n2, n1 + 1), (nb, n3, n2 + 1, n1), (nb, n3 + 1, n2, n1)]


def morton(lx):
    key = 0
    for b in range(20):
        for a in range(3):
            key |= ((lx[a] >> b) & 1) << (3 * b + a)
    return key


def parse(path, d):
    raw = open(path, "rb").read()
    end = raw.index(b"<par_end>") + len(b"<par_end>") + 1
    text = raw[:end].decode()
    o = end
    nbtotal, root = struct.unpack_from("<ii", raw, o); o += 8
    rs = struct.unpack_from("<9d3i4x", raw, o); o += 88
    time, dt = struct.unpack_from("<dd", raw, o); o += 16
    ncycle, = struct.unpack_from("<i", raw, o); o += 4
    datasize, = struct.unpack_from("<Q", raw, o); o += 8
    ids = [struct.unpack_from("<qqqi4xd", raw, o + 40 * q) for q in range(nbtotal)]
    o += 40 * nbtotal
    sh = shapes(d)
    per = [int(np.prod(s[1:])) for s in sh]
    assert datasize == 8 * sum(per)
    data = []
    for q in range(nbtotal):
        blk, p = [], o + q * datasize
        for n in per:
            blk.append(np.frombuffer(raw, "<f8", n, p))
            p += 8 * n
        data.append(blk)
    assert len(raw) == o + nbtotal * datasize
    return text, nbtotal, root, rs, time, dt, ncycle, ids, data


@pytest.mark.parametrize("nx,bnx,eos", [((24, 8, 12), (4, 4, 4), 0), ((16, 16, 16), (8, 8, 8), 1)])
def test_layout_zorder_and_round_trip(nx, bnx, eos, tmp_path):
    agx = load_agx()
    d = desc_of(nx, bnx, eos)
    rs = np.random.default_rng(1)
    arrs = [rs.standard_normal(s) for s in shapes(d)]
    path = str(tmp_path / "run.00003.rst")
    agx.restart_write_host(path, d, 0.625, 1.5e-3, 17, *arrs)
    text, nbtotal, root, region, time, dt, ncycle, ids, data = parse(path, d)
    nb1, nb2, nb3 = nx[0] // bnx[0], nx[1] // bnx[1], nx[2] // bnx[2]
    assert text.startswith("#------------------------- PAR_DUMP") and text.endswith("<par_end>\n")
    assert "<mesh>" in text and f"nx1 = {nx[0]}" in text
    assert nbtotal == nb1 * nb2 * nb3
    assert root == int(np.ceil(np.log2(max(nb1, nb2, nb3))))
    assert region == (-1.0, 0.0, 0.0, 1.0, 0.5, 0.75, 1.0, 1.0, 1.0) + tuple(nx)
    assert (time, dt, ncycle) == (0.625, 1.5e-3, 17)
    locs = [i[:3] for i in ids]
    assert all(i[3] == root and i[4] == 1.0 for i in ids)
    want = sorted(((a, b, c) for c in range(nb3) for b in range(nb2) for a in range(nb1)), key=morton)
    assert locs == want
    for (a, b, c), blk in zip(locs, data):
        g = (c * nb2 + b) * nb1 + a
        for x, arr in zip(blk, arrs):
            assert np.array_equal(x, arr[g].ravel())
    t, h, n, *back = agx.restart_read_host(path, d, shapes(d))
    assert (t, h, n) == (0.625, 1.5e-3, 17)
    for x, y in zip(back, arrs):
        assert np.array_equal(x, y)


def test_caller_parameters_and_refusals(tmp_path):
    agx = load_agx()
    d = desc_of((16, 8, 8), (8, 4, 4))
    arrs = [np.ones(s) for s in shapes(d)]
    path = str(tmp_path / "a.rst")
    params = "<job>\nproblem_id = collapse\n<mesh>\nnx1 = 16\n"
    agx.restart_write_host(path, d, 1.0, 0.1, 3, *arrs, params=params)
    assert open(path, "rb").read().startswith((params + "<par_end>\n").encode())
    with pytest.raises(agx.AgxError, match="par_end"):
        agx.restart_write_host(path, d, 1.0, 0.1, 3, *arrs, params=params + "<par_end>\n")
    other = desc_of((16, 8, 16), (8, 4, 4))
    with pytest.raises(agx.AgxError, match="another mesh"):
        agx.restart_read_host(path, other, shapes(other))
    deeper = desc_of((16, 8, 8), (8, 4, 4), ng=3)
    with pytest.raises(agx.AgxError, match="bytes"):
        agx.restart_read_host(path, deeper, shapes(deeper))
    raw = open(path, "rb").read()
    open(path, "wb").write(raw[:-8])
    with pytest.raises(agx.AgxError, match="truncated"):
        agx.restart_read_host(path, d, shapes(d))
