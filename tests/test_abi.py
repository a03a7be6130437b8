"""The C-ABI library loads, exports every symbol include/athena_gamma_amd.h declares,
and rejects invalid descriptors on the host (no GPU needed for these checks)."""
import ctypes
import os
import re

import pytest

from tests.conftest import ROOT, load_agx

HEADER = os.path.join(ROOT, "include", "athena_gamma_amd.h")


def declared_functions():
    src = open(HEADER).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(agx_[a-z0-9_]+)\s*\(", src)))


def test_header_matches_binding_list(agx):
    assert declared_functions() == sorted(agx.EXPORTS)


def test_library_exports_all(agx):
    L = agx.lib()
    for name in declared_functions():
        assert hasattr(L, name), name


def test_desc_layout(agx):
    from oracle.orc import Desc as ODesc
    assert ctypes.sizeof(agx.Desc) == ctypes.sizeof(ODesc)
    assert [f[0] for f in agx.Desc._fields_] == [f[0] for f in ODesc._fields_]
    # 14 int32 (56 B), 13 doubles, 3 int32 (+4 B padding), 3 doubles, 1 int32 (+4 B
    # padding), naturally aligned
    assert agx.Desc.mp_origin.offset == 56 + 13 * 8 + 12 + 4
    assert agx.Desc.mp_nodipole.offset == 56 + 13 * 8 + 16 + 3 * 8
    assert ctypes.sizeof(agx.Desc) == 56 + 13 * 8 + 16 + 3 * 8 + 8


@pytest.mark.parametrize("bad,code", [
    (dict(nx=(30, 16, 16), bnx=(16, 16, 16)), -3),      # mesh not a multiple of the block
    (dict(nghost=1), -5),
    (dict(xorder=4), -6),
    (dict(rsolver=9), -8),
    (dict(gamma=1.0), -11),
    (dict(eos=2), -9),                                  # unknown EOS
    (dict(eos=1, rsolver=2), -9),                       # Roe has no isothermal form here
    (dict(eos=1, iso_cs=0.0), -9),
    (dict(precision=3), -10),
    (dict(self_gravity=1, mg_bc=1, mp_auto_origin=1, mp_nodipole=1), -14),   # mg_gravity.cpp:106-113
])
def test_invalid_desc_rejected(agx, bad, code):
    kw = dict(nx=(32, 16, 16), bnx=(16, 16, 16))
    kw.update(bad)
    d = agx.make_desc(**kw)
    h = ctypes.c_void_p()
    rc = agx.lib().agx_mesh_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == code
    assert not h.value
    assert len(agx.lib().agx_last_error()) > 0


def test_null_mesh_calls_fail(agx):
    L = agx.lib()
    assert L.agx_calculate_fluxes(None, 2) < 0
    assert L.agx_cycle(None, 1.0, 1, None) < 0
    assert L.agx_mg_solve(None, 0) < 0


def test_scatter_gather_roundtrip(agx):
    import numpy as np
    d = agx.make_desc(nx=(8, 12, 4), bnx=(4, 4, 2))
    G = np.random.default_rng(0).normal(size=(3, 4, 12, 8))
    assert np.array_equal(agx.gather_cells(d, agx.scatter_cells(d, G)), G)


@pytest.mark.parametrize("kind,it,nbf,nbc,code", [
    ("restrict_cc", dict(fb=0, cb=0, si=0, ei=5, sj=2, ej=5, sk=2, ek=5), 1, 1, -40),    # fine i = -2
    ("restrict_cc", dict(fb=1, cb=0, si=2, ei=5, sj=2, ej=5, sk=2, ek=5), 1, 1, -40),    # block out of range
    ("prolong_cc", dict(fb=0, cb=0, si=0, ei=3, sj=2, ej=5, sk=2, ek=5), 1, 1, -40),     # stencil at -1
    ("prolong_fc1", dict(fb=0, cb=0, si=2, ei=9, sj=2, ej=5, sk=2, ek=5), 1, 1, -40),    # x1 face 9 > nc
    ("flux_from_finer", dict(fb=0, cb=0, fid=6), 1, 1, -40),
    ("flux_from_finer", dict(fb=0, cb=0, fid=1, fi1=2), 1, 1, -40),
    ("emf_to_coarser", dict(fb=0, fid=18), 1, 1, -40),
    ("emf_set_finer", dict(fb=0, cb=0, fid=7, fi2=1), 1, 1, -40),       # an edge has one half index
    ("emf_set_same", dict(fb=3, cb=0, fid=0), 2, 1, -40),               # buffer slot out of range
    ("emf_average", dict(cb=0, fid=8, fi1=1), 1, 1, -40),               # an edge needs 2..4 blocks
    ("emf_clear", dict(cb=1, fid=0), 1, 1, -40),
])
def test_refinement_lists_reject_bad_items(agx, kind, it, nbf, nbc, code):
    """agx_ref_list_create checks every item on the host before any device work."""
    g = agx.RefGeom(8, 8, 8, 2, 2, 0)
    full = dict(fb=0, cb=0, si=0, ei=0, sj=0, ej=0, sk=0, ek=0, fid=0, fi1=0, fi2=0, aux=0)
    full.update(it)
    items = (agx.RefItem * 1)(agx.RefItem(**full))
    h = ctypes.c_void_p()
    rc = agx.lib().agx_ref_list_create(ctypes.byref(g), agx.REF_KINDS[kind], items, 1, nbf, nbc, ctypes.byref(h))
    assert rc == code, agx.lib().agx_last_error()
    assert not h.value
    assert b"item 0" in agx.lib().agx_last_error()


def test_refinement_emf_count(agx):
    g = agx.RefGeom(8, 6, 4, 2, 2, 0)
    # x1 face: e2 (nx3/2 + 1) x nx2/2 + e3 nx3/2 x (nx2/2 + 1), flux_correction_fc.cpp:334-356
    assert agx.ref_emf_count(g, 0) == 3 * 3 + 2 * 4
    assert agx.ref_emf_count(g, 3) == 3 * 4 + 2 * 5
    assert agx.ref_emf_count(g, 5) == 4 * 4 + 3 * 5


def test_refinement_emf_lists_reject_duplicates(agx):
    """two items writing the same face of one block cannot share a launch"""
    g = agx.RefGeom(8, 8, 8, 2, 2, 0)
    it = dict(fb=0, cb=1, si=0, ei=0, sj=0, ej=0, sk=0, ek=0, fid=2, fi1=0, fi2=0, aux=0)
    items = (agx.RefItem * 2)(agx.RefItem(**it), agx.RefItem(**it))
    h = ctypes.c_void_p()
    for kind in ("emf_set_same", "emf_clear", "emf_average"):
        rc = agx.lib().agx_ref_list_create(ctypes.byref(g), agx.REF_KINDS[kind], items, 2, 4, 4, ctypes.byref(h))
        assert rc == -40 and not h.value, kind
        assert b"same block" in agx.lib().agx_last_error()


def test_refinement_emf_counts_edges(agx):
    g = agx.RefGeom(8, 6, 4, 2, 2, 0)
    # edges to a coarser block: one restricted line, flux_correction_fc.cpp:481-546
    assert [agx.ref_emf_count(g, 6 + e) for e in (0, 4, 8)] == [2, 3, 4]
    # same level: x1 face e2 (nx3+1) x nx2 + e3 nx3 x (nx2+1), edges one line
    assert agx.ref_emf_count_same_level(g, 0) == 5 * 6 + 4 * 7
    assert [agx.ref_emf_count_same_level(g, 6 + e) for e in (0, 4, 8)] == [4, 6, 8]
