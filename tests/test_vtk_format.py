"""VTK legacy output in the reference's format (outputs/vtk.cpp:54-220, quantities of
outputs/outputs.cpp:355-660) through the library's host writer (no GPU needed), and the
reference's MeshBlock numbering (agx_block_gids).

Each file is parsed here independently: the text lines, the big-endian float32
coordinate arrays and cell data.  Checks: the file names (gid = position in the Z order
of the block tree), the header, the face coordinates of every block (the oracle's
coordinates, rounded to float32), every quantity of "prim" and "cons" in
LoadOutputData's order and values (the oracle's arrays rounded to float32, vectors
interleaved per cell), the isothermal layout (no pressure or energy), ghost zones, single
components, and the refusal of a token the library cannot serve."""
import os

import numpy as np
import pytest

from oracle.orc import Oracle, make_desc
from tests.conftest import load_agx
from tests.states import load_state, turbulent_state


def morton_order(nb1, nb2, nb3):
    """lattice indices sorted by their Morton key (x1 the least significant bit of each
    level), which is the order a depth-first walk of the octree visits the leaves"""
    bits = max(nb1, nb2, nb3).bit_length()
    keys = []
    for k in range(nb3):
        for j in range(nb2):
            for i in range(nb1):
                key = 0
                for b in range(bits):
                    key |= ((i >> b) & 1) << (3 * b) | ((j >> b) & 1) << (3 * b + 1) | ((k >> b) & 1) << (3 * b + 2)
                keys.append((key, (k * nb2 + j) * nb1 + i))
    return [g for _, g in sorted(keys)]


@pytest.mark.parametrize("nb", [(1, 1, 1), (2, 2, 2), (4, 2, 2), (3, 2, 5), (2, 3, 1)])
def test_block_gids_are_the_z_order(nb):
    agx = load_agx()
    desc = make_desc(nx=tuple(4 * n for n in nb), bnx=(4, 4, 4))
    assert list(agx.block_gids(desc)) == morton_order(*nb)


def parse(path):
    raw = open(path, "rb").read()
    pos = 0

    def line():
        nonlocal pos
        e = raw.index(b"\n", pos)
        s = raw[pos:e].decode()
        pos = e + 1
        return s

    def floats(n):
        nonlocal pos
        a = np.frombuffer(raw, dtype=">f4", count=n, offset=pos).astype(np.float32)
        pos += 4 * n
        return a
    out = {"version": line(), "header": line(), "format": line(), "dataset": line()}
    dims = [int(x) for x in line().split()[1:]]
    out["dims"] = dims
    coords = []
    for ax, n in zip("XYZ", dims):
        if ax != "X":
            assert line() == ""
        h = line().split()
        assert h == [f"{ax}_COORDINATES", str(n), "float"]
        coords.append(floats(n))
    out["coords"] = coords
    assert line() == ""
    cd = line().split()
    assert cd[0] == "CELL_DATA"
    ncells = int(cd[1])
    data = []
    while pos < len(raw):
        if data:   # each quantity's binary block ends where "\n<kind> <name> float" begins
            assert line() == ""
        kind, name, typ = line().split()
        assert typ == "float"
        nv = 3 if kind == "VECTORS" else 1
        if nv == 1:
            assert line() == "LOOKUP_TABLE default"
        data.append((kind, name, floats(nv * ncells).reshape(-1, nv)))
    out["data"] = data
    return out


def oracle_state(sg=False, iso=False):
    kw = dict(self_gravity=1, four_pi_G=1.0, mg_mode=0, mg_threshold=-1.0, mg_niter=3) if sg else {}
    # logically cubic MeshBlocks, as multigrid needs
    desc = make_desc(nx=(16, 8, 8), bnx=(4, 4, 4), box=((0, 2), (0, 1), (0, 1)), eos=1 if iso else 0, iso_cs=0.8, **kw)
    o = Oracle(desc)
    load_state(o, turbulent_state(desc, seed=3, mach=1.0, iso=iso))
    o.initialize()
    o.cycle(1e9)
    return desc, o


def expected(o, name, nh, sl):
    u, w, bcc, phi = o.get("u")[:, :nh], o.get("w")[:, :nh], o.get("bcc"), o.get("phi")
    table = {"dens": (u, [0]), "rho": (w, [0]), "Etot": (u, [4]), "press": (w, [4]), "mom": (u, [1, 2, 3]),
             "mom1": (u, [1]), "mom2": (u, [2]), "mom3": (u, [3]), "vel": (w, [1, 2, 3]), "vel1": (w, [1]),
             "vel2": (w, [2]), "vel3": (w, [3]), "phi": (phi, [0]), "Bcc": (bcc, [0, 1, 2]), "Bcc1": (bcc, [0]),
             "Bcc2": (bcc, [1]), "Bcc3": (bcc, [2])}
    arr, comps = table[name]
    return lambda b: np.stack([arr[b, c][sl].reshape(-1) for c in comps], axis=1).astype(np.float32)


@pytest.mark.parametrize("variable,sg,iso,ghost,names", [
    ("prim", False, False, False, ["rho", "press", "vel", "Bcc"]),
    ("cons", True, False, False, ["dens", "Etot", "mom", "phi", "Bcc"]),
    ("prim", True, False, True, ["rho", "press", "vel", "phi", "Bcc"]),
    ("prim", False, True, False, ["rho", "vel", "Bcc"]),
    ("cons", False, True, False, ["dens", "mom", "Bcc"]),
    ("v2,d,bcc3,m1,E", False, False, False, ["rho", "Etot", "mom1", "vel2", "Bcc3"]),
])
def test_files_hold_the_reference_layout(tmp_path, variable, sg, iso, ghost, names):
    agx = load_agx()
    desc, o = oracle_state(sg=sg, iso=iso)
    nh = 4 if iso else 5
    t, _, ncyc = o.time()
    base = str(tmp_path / "run")
    agx.vtk_write_host(base, desc, t, ncyc, o.get("u")[:, :nh], o.get("w")[:, :nh], o.get("bcc"),
                       o.get("phi") if sg else None, file_id="out2", file_number=7, variable=variable, ghost=ghost)
    ng = desc.nghost
    sl = (slice(None), slice(None), slice(None)) if ghost else (slice(ng, -ng),) * 3
    gids = agx.block_gids(desc)
    assert sorted(os.listdir(tmp_path)) == sorted(f"run.block{g}.out2.00007.vtk" for g in range(len(gids)))
    for gid, b in enumerate(gids):
        f = parse(str(tmp_path / f"run.block{gid}.out2.00007.vtk"))
        assert f["version"] == "# vtk DataFile Version 2.0"
        assert f["header"] == f"# Athena++ data at time={t:e}  cycle={ncyc}  variables={variable} "
        assert f["format"] == "BINARY" and f["dataset"] == "DATASET RECTILINEAR_GRID"
        n = [desc.bnx1, desc.bnx2, desc.bnx3]
        if ghost:
            n = [x + 2 * ng for x in n]
        assert f["dims"] == [x + 1 for x in n]
        for d in range(3):
            xf = o.coords(int(b), d)[0]
            want = xf if ghost else xf[ng:len(xf) - ng]
            np.testing.assert_array_equal(f["coords"][d], want.astype(np.float32))
        assert [nm for _, nm, _ in f["data"]] == names
        for kind, nm, vals in f["data"]:
            assert kind == ("VECTORS" if nm in ("vel", "mom", "Bcc") else "SCALARS")
            np.testing.assert_array_equal(vals, expected(o, nm, nh, sl)(int(b)))


def test_refusals(tmp_path):
    agx = load_agx()
    desc, o = oracle_state()
    args = (desc, 0.0, 0, o.get("u"), o.get("w"), o.get("bcc"))
    with pytest.raises(agx.AgxError, match="b1"):
        agx.vtk_write_host(str(tmp_path / "x"), *args, variable="prim,b1")
    with pytest.raises(agx.AgxError, match="five digits"):
        agx.vtk_write_host(str(tmp_path / "x"), *args, file_number=100000)
    assert os.listdir(tmp_path) == []
