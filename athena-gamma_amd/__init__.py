"""Python binding (ctypes) of the gfx950 MHD + multigrid library's C ABI.

The library is include/athena_gamma_amd.h built into lib/libathena_gamma_amd.so by
build.py.  This module is plumbing for tests and benchmarks: every call goes through
the C ABI, and nothing here computes physics.  Loading fails loudly when the library
has not been built; there is no fallback.

The directory name is not a Python identifier; load it with importlib, e.g.

    spec = importlib.util.spec_from_file_location(
        "athena_gamma_amd", "athena-gamma_amd/__init__.py")
"""
import ctypes
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libathena_gamma_amd.so")

ARR = dict(u=0, u1=1, w=2, bcc=3, b1f=4, b2f=5, b3f=6, r1f=7, r2f=8, r3f=9,
           fx1=10, fx2=11, fx3=12, e3x1=13, e2x1=14, e1x2=15, e3x2=16, e1x3=17,
           e2x3=18, wx1=19, wx2=20, wx3=21, e1=22, e2=23, e3=24, phi=25)
KCLASS = dict(flux_x1=0, flux_x2=1, flux_x3=2, corner_e=3, integrate=4, bvals=5, c2p=6,
              newdt=7, mg_smooth=8, mg_restrict=9, mg_prolong=10, mg_other=11)

# every symbol include/athena_gamma_amd.h declares
EXPORTS = [
    "agx_mesh_create", "agx_mesh_destroy", "agx_last_error", "agx_array_dims",
    "agx_mesh_bytes", "agx_upload", "agx_download", "agx_set_time", "agx_get_time",
    "agx_synchronize", "agx_calculate_fluxes", "agx_compute_corner_e", "agx_integrate",
    "agx_bvals_exchange", "agx_cons2prim", "agx_new_block_dt", "agx_new_time_step",
    "agx_initialize", "agx_cycle", "agx_mg_solve", "agx_mg_set_mode", "agx_mg_last",
    "agx_mg_vcycles", "agx_autotune", "agx_set_tile", "agx_set_graph", "agx_timing_enable", "agx_timing_reset",
    "agx_timing_get",
    "agx_ref_list_create", "agx_ref_list_destroy", "agx_ref_emf_count", "agx_ref_restrict_cc",
    "agx_ref_restrict_fc", "agx_ref_prolong_cc", "agx_ref_prolong_fc_shared", "agx_ref_prolong_fc_internal",
    "agx_ref_flux_from_finer", "agx_ref_emf_to_coarser", "agx_ref_sync",
    "agx_device_alloc", "agx_device_free", "agx_device_copy", "agx_ref_prim_to_cons", "agx_ref_prolong_ranges",
    "agx_upload_blocks", "agx_download_blocks", "agx_set_mg_fuse",
    "agx_ref_emf_count_same_level", "agx_ref_emf_same_level", "agx_ref_emf_set", "agx_ref_emf_clear",
    "agx_ref_emf_average", "agx_ref_count_fine_edges", "agx_ref_copy_cc", "agx_ref_pack_fc", "agx_ref_unpack_fc",
    "agx_ref_cons_to_prim", "agx_stream_create", "agx_stream_destroy", "agx_graph_begin", "agx_graph_end",
    "agx_graph_launch", "agx_graph_destroy", "agx_ref_restrict_all", "agx_ref_prolong_all",
    "agx_ref_internal_to_cons", "agx_mg_set_source_mask_sphere", "agx_set_mg_coarse",
    "agx_history", "agx_history_sum", "agx_history_sum_noalloc",
    "agx_restart_write", "agx_restart_read", "agx_restart_write_host", "agx_restart_read_host",
    "agx_vtk_write", "agx_vtk_write_host", "agx_block_gids",
    "agx_mg_setup", "agx_mg_smooth", "agx_mg_fill_ghost", "agx_mg_fas", "agx_mg_restrict", "agx_mg_prolong",
    "agx_mg_fmg_prolong", "agx_mg_solve_coarsest", "agx_mg_defect_norm", "agx_mg_retrieve", "agx_mg_coffset",
    "agx_mg_levels", "agx_array_ptr", "agx_mesh_stream", "agx_mesh_set_levels",
]


class Desc(ctypes.Structure):
    """agx_mesh_desc (include/athena_gamma_amd.h)."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "nx1", "nx2", "nx3", "bnx1", "bnx2", "bnx3", "nghost", "xorder", "rsolver",
        "eos", "self_gravity", "mg_mode", "mg_niter", "precision")] + \
        [(n, ctypes.c_double) for n in (
            "x1min", "x1max", "x2min", "x2max", "x3min", "x3max", "gamma", "iso_cs",
            "cfl", "dfloor", "pfloor", "four_pi_G", "mg_threshold")] + \
        [(n, ctypes.c_int32) for n in ("mg_bc", "mporder", "mp_auto_origin")] + \
        [("mp_origin", ctypes.c_double * 3), ("mp_nodipole", ctypes.c_int32), ("char_proj", ctypes.c_int32)]


def make_desc(nx=(64, 32, 32), bnx=None, box=((0.0, 3.0), (0.0, 1.5), (0.0, 1.5)),
              nghost=2, xorder=2, rsolver=0, gamma=1.666666666666667, cfl=0.3,
              self_gravity=0, four_pi_G=1.0, mg_mode=0, mg_threshold=-1.0, mg_niter=10,
              precision=0, eos=0, iso_cs=1.0, mg_bc=0, mporder=4, mp_auto_origin=1,
              mp_origin=(0.0, 0.0, 0.0), mp_nodipole=0, char_proj=0):
    bnx = bnx or nx
    d = Desc()
    d.nx1, d.nx2, d.nx3 = nx
    d.bnx1, d.bnx2, d.bnx3 = bnx
    d.nghost, d.xorder, d.rsolver, d.eos = nghost, xorder, rsolver, eos
    d.self_gravity, d.mg_mode, d.mg_niter, d.precision = self_gravity, mg_mode, mg_niter, precision
    (d.x1min, d.x1max), (d.x2min, d.x2max), (d.x3min, d.x3max) = box
    d.gamma, d.iso_cs, d.cfl = gamma, iso_cs, cfl
    # EquationOfState defaults sqrt(1024*FLT_MIN), eos/adiabatic_mhd.cpp:26-31
    d.dfloor = d.pfloor = float(np.sqrt(1024 * np.float64(np.finfo(np.float32).tiny)))
    d.four_pi_G, d.mg_threshold = four_pi_G, mg_threshold
    d.mg_bc, d.mporder, d.mp_auto_origin = mg_bc, mporder, mp_auto_origin
    d.mp_origin[:] = mp_origin
    d.mp_nodipole = mp_nodipole
    d.char_proj = char_proj
    return d


class RefItem(ctypes.Structure):
    """agx_ref_item: one work item of a refinement operator."""
    _fields_ = [(n, ctypes.c_int32) for n in ("fb", "cb", "si", "ei", "sj", "ej", "sk", "ek",
                                              "fid", "fi1", "fi2", "aux")]


class RefGeom(ctypes.Structure):
    """agx_ref_geom: block interior size and the ghost widths of blocks and coarse buffers."""
    _fields_ = [(n, ctypes.c_int32) for n in ("nx1", "nx2", "nx3", "ng", "cng", "layout")]


class RefCoords(ctypes.Structure):
    """agx_ref_coords: device pointers to per-block x?v and dx?f tables."""
    _fields_ = [(n, ctypes.c_void_p) for n in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")]


REF_KINDS = dict(restrict_cc=0, restrict_fc1=1, restrict_fc2=2, restrict_fc3=3, prolong_cc=4,
                 prolong_fc1=5, prolong_fc2=6, prolong_fc3=7, prolong_fc_int=8, flux_from_finer=9,
                 emf_to_coarser=10, prim_to_cons=11, emf_same_level=12, emf_set_same=13, emf_set_finer=14,
                 emf_clear=15, emf_average=16, copy_cc_bb=17, copy_cc_cb=18, copy_cc_bc=19,
                 pack_fc_b=20, pack_fc_c=21, unpack_fc_b=22, unpack_fc_c=23, cons_to_prim=24)

_LIB = None


def lib():
    """Load the built library; raises if it is missing (no fallback path exists)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libathena_gamma_amd.so not built: run athena-gamma_amd/build.py")
    L = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    I = ctypes.c_int
    L.agx_mesh_create.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(P)]
    L.agx_mesh_destroy.argtypes = [P]
    L.agx_last_error.restype = ctypes.c_char_p
    L.agx_array_dims.argtypes = [P, I, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]
    L.agx_mesh_bytes.restype = ctypes.c_int64
    L.agx_mesh_bytes.argtypes = [P]
    L.agx_array_ptr.argtypes = [P, I, ctypes.POINTER(ctypes.c_void_p)]
    L.agx_mesh_stream.argtypes = [P, ctypes.POINTER(ctypes.c_void_p)]
    L.agx_mesh_set_levels.argtypes = [P, ctypes.POINTER(ctypes.c_int32), D, ctypes.POINTER(ctypes.c_int32),
                                      ctypes.POINTER(ctypes.c_int64)]
    L.agx_upload.argtypes = [P, I, D]
    L.agx_download.argtypes = [P, I, D]
    L.agx_set_time.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int32]
    I32P = ctypes.POINTER(ctypes.c_int32)
    L.agx_upload_blocks.argtypes = [P, I, ctypes.c_int32, I32P, I32P, ctypes.POINTER(D)]
    L.agx_download_blocks.argtypes = [P, I, ctypes.c_int32, I32P, I32P, ctypes.POINTER(D)]
    L.agx_get_time.argtypes = [P, D, D, ctypes.POINTER(ctypes.c_int32)]
    L.agx_synchronize.argtypes = [P]
    L.agx_calculate_fluxes.argtypes = [P, I]
    L.agx_compute_corner_e.argtypes = [P]
    L.agx_integrate.argtypes = [P, I]
    L.agx_bvals_exchange.argtypes = [P, I]
    L.agx_cons2prim.argtypes = [P]
    L.agx_new_block_dt.argtypes = [P, D]
    L.agx_new_time_step.argtypes = [P, ctypes.c_double]
    L.agx_initialize.argtypes = [P, ctypes.c_double]
    L.agx_cycle.argtypes = [P, ctypes.c_double, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.agx_mg_solve.argtypes = [P, I]
    L.agx_mg_set_mode.argtypes = [P, I, ctypes.c_double, ctypes.c_int32]
    L.agx_mg_last.argtypes = [P, D, ctypes.POINTER(ctypes.c_int32)]
    L.agx_mg_vcycles.argtypes = [P, ctypes.c_int32]
    L.agx_autotune.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
    L.agx_set_tile.argtypes = [P, ctypes.c_int32]
    L.agx_set_mg_fuse.argtypes = [P, ctypes.c_int32]
    L.agx_set_mg_coarse.argtypes = [P, ctypes.c_int32]
    L.agx_mg_set_source_mask_sphere.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_double]
    L.agx_set_graph.argtypes = [P, ctypes.c_int32]
    L.agx_timing_enable.argtypes = [P, I]
    L.agx_timing_reset.argtypes = [P]
    L.agx_timing_get.argtypes = [P, I, D, ctypes.POINTER(ctypes.c_int64)]
    G, C = ctypes.POINTER(RefGeom), RefCoords
    L.agx_ref_list_create.argtypes = [G, ctypes.c_int32, ctypes.POINTER(RefItem), ctypes.c_int32,
                                      ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(P)]
    L.agx_ref_list_destroy.argtypes = [P]
    L.agx_ref_emf_count.restype = ctypes.c_int64
    L.agx_ref_emf_count.argtypes = [G, ctypes.c_int32]
    L.agx_ref_restrict_cc.argtypes = [P, ctypes.c_int32, P, C, P, P]
    L.agx_ref_restrict_fc.argtypes = [P, P, C, P, P]
    L.agx_ref_prolong_cc.argtypes = [P, ctypes.c_int32, P, C, P, C, P]
    L.agx_ref_prolong_fc_shared.argtypes = [P, P, C, P, C, P]
    L.agx_ref_prolong_fc_internal.argtypes = [P, P, P, P, C, P]
    L.agx_ref_flux_from_finer.argtypes = [P, ctypes.c_int32, P, P, P, C, P, P, P, P]
    L.agx_ref_emf_to_coarser.argtypes = [P, P, P, P, C, P, ctypes.c_int64, P]
    L.agx_ref_sync.argtypes = [P]
    L.agx_ref_prim_to_cons.argtypes = [P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_int32, P]
    IP = ctypes.POINTER(ctypes.c_int32)
    L.agx_ref_prolong_ranges.argtypes = [G, IP, IP, IP, IP]
    L.agx_ref_emf_count_same_level.restype = ctypes.c_int64
    L.agx_ref_emf_count_same_level.argtypes = [G, ctypes.c_int32]
    L.agx_ref_emf_same_level.argtypes = [P, P, P, P, P, ctypes.c_int64, P]
    L.agx_ref_emf_set.argtypes = [P, P, P, P, P, ctypes.c_int64, P]
    L.agx_ref_emf_clear.argtypes = [P, P, P, P, P]
    L.agx_ref_emf_average.argtypes = [P, P, P, P, P]
    L.agx_ref_count_fine_edges.argtypes = [IP, IP, IP]
    L.agx_ref_copy_cc.argtypes = [P, ctypes.c_int32, P, P, P]
    L.agx_ref_restrict_all.argtypes = [P, P, P, P, ctypes.c_int32, P, P, P, P, C, P, P, P, P, P]
    L.agx_ref_prolong_all.argtypes = [P, P, P, P, ctypes.c_int32, P, P, P, P, C, P, P, P, P, C, P]
    L.agx_ref_internal_to_cons.argtypes = [P, P, P, P, C, P, P, P, ctypes.c_double, ctypes.c_int32, P]
    L.agx_stream_create.argtypes = [ctypes.POINTER(P)]
    L.agx_stream_destroy.argtypes = [P]
    L.agx_graph_begin.argtypes = [P]
    L.agx_graph_end.argtypes = [P, ctypes.POINTER(P)]
    L.agx_graph_launch.argtypes = [P, P]
    L.agx_graph_destroy.argtypes = [P]
    L.agx_ref_pack_fc.argtypes = [P, P, P, P, P, ctypes.c_int64, P]
    L.agx_ref_cons_to_prim.argtypes = [P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int32, P]
    L.agx_ref_unpack_fc.argtypes = [P, P, P, P, P, ctypes.c_int64, P]
    L.agx_history.argtypes = [P, D, ctypes.c_int32]
    for fn in ("agx_history_sum", "agx_history_sum_noalloc"):
        getattr(L, fn).argtypes = [P, P, P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                   ctypes.c_int32, ctypes.c_int32, P, P, P]
    for fn in ("agx_mg_setup", "agx_mg_solve_coarsest", "agx_mg_retrieve"):
        getattr(L, fn).argtypes = [P]
    for fn in ("agx_mg_fill_ghost", "agx_mg_fas", "agx_mg_restrict", "agx_mg_prolong", "agx_mg_fmg_prolong"):
        getattr(L, fn).argtypes = [P, ctypes.c_int32]
    L.agx_mg_smooth.argtypes = [P, ctypes.c_int32, ctypes.c_int32]
    L.agx_mg_defect_norm.argtypes = [P, D]
    L.agx_mg_coffset.argtypes = [P, ctypes.c_int32, ctypes.POINTER(ctypes.c_int32)]
    L.agx_mg_levels.argtypes = [P, ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32), ctypes.c_int32]
    C = ctypes.c_char_p
    L.agx_restart_write.argtypes = [P, C, C]
    L.agx_restart_read.argtypes = [P, C]
    L.agx_restart_write_host.argtypes = [C, ctypes.POINTER(Desc), C, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int32, D, D, D, D]
    L.agx_restart_read_host.argtypes = [C, ctypes.POINTER(Desc), D, D, ctypes.POINTER(ctypes.c_int32), D, D, D, D]
    L.agx_vtk_write.argtypes = [P, C, C, ctypes.c_int32, C, ctypes.c_int32]
    L.agx_vtk_write_host.argtypes = [C, C, ctypes.c_int32, C, ctypes.c_int32, ctypes.POINTER(Desc), ctypes.c_double,
                                     ctypes.c_int32, D, D, D, D]
    L.agx_block_gids.argtypes = [ctypes.POINTER(Desc), ctypes.POINTER(ctypes.c_int64)]
    L.agx_device_alloc.argtypes = [ctypes.c_int64, ctypes.POINTER(P)]
    L.agx_device_free.argtypes = [P]
    L.agx_device_copy.argtypes = [P, P, ctypes.c_int64, ctypes.c_int32]
    _LIB = L
    return L


class AgxError(RuntimeError):
    pass


def _check(rc, what):
    if rc != 0:
        msg = lib().agx_last_error().decode(errors="replace")
        raise AgxError(f"{what} failed ({rc}): {msg}")


class Mesh:
    """One device-resident Mesh (all MeshBlocks); thin wrapper over agx_mesh."""

    def __init__(self, desc):
        self.L = lib()
        if not isinstance(desc, Desc):   # same layout from another binding: copy fields
            d = Desc()
            for name, _ in Desc._fields_:
                setattr(d, name, getattr(desc, name))
            desc = d
        self.desc = desc
        h = ctypes.c_void_p()
        _check(self.L.agx_mesh_create(ctypes.byref(desc), ctypes.byref(h)), "agx_mesh_create")
        self.h = h
        nb = ctypes.c_int64()
        dims = (ctypes.c_int64 * 4)()
        _check(self.L.agx_array_dims(self.h, 0, ctypes.byref(nb), dims), "agx_array_dims")
        self.nb = nb.value

    def close(self):
        if getattr(self, "h", None):
            self.L.agx_mesh_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def shape(self, name):
        nb = ctypes.c_int64()
        dims = (ctypes.c_int64 * 4)()
        _check(self.L.agx_array_dims(self.h, ARR[name], ctypes.byref(nb), dims), "agx_array_dims")
        d = list(dims)
        if d[0] == 1 and name not in ("phi",):
            return (nb.value, d[1], d[2], d[3])
        return (nb.value, d[0], d[1], d[2], d[3])

    def get(self, name):
        a = np.zeros(self.shape(name))
        _check(self.L.agx_download(self.h, ARR[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "agx_download")
        return a

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).reshape(self.shape(name))
        _check(self.L.agx_upload(self.h, ARR[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
               "agx_upload")

    def _blocks(self, name, blocks, arrays, up):
        n = len(blocks)
        org = (ctypes.c_int32 * (3 * n))(*[v for o, _ in blocks for v in o])
        siz = (ctypes.c_int32 * (3 * n))(*[v for _, z in blocks for v in z])
        ptrs = (ctypes.POINTER(ctypes.c_double) * n)(*[a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
                                                        for a in arrays])
        f = self.L.agx_upload_blocks if up else self.L.agx_download_blocks
        _check(f(self.h, ARR[name], n, org, siz, ptrs), "agx_upload_blocks" if up else "agx_download_blocks")

    def set_blocks(self, name, blocks, arrays):
        """Upload MeshBlocks of uneven sizes: blocks = [((i0, j0, k0), (nx1, nx2, nx3)), ...],
        arrays in the reference layout with ghosts (agx_upload_blocks)."""
        arrays = [np.ascontiguousarray(a, dtype=np.float64) for a in arrays]
        self._blocks(name, blocks, arrays, True)

    def get_blocks(self, name, blocks, shapes):
        """Download MeshBlocks of uneven sizes, ghosts included (agx_download_blocks)."""
        arrays = [np.zeros(s) for s in shapes]
        self._blocks(name, blocks, arrays, False)
        return arrays

    def time(self):
        t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
        _check(self.L.agx_get_time(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(n)),
               "agx_get_time")
        return t.value, dt.value, n.value

    def set_time(self, t, dt, ncycle=0):
        _check(self.L.agx_set_time(self.h, t, dt, ncycle), "agx_set_time")

    def sync(self):
        _check(self.L.agx_synchronize(self.h), "agx_synchronize")

    def calculate_fluxes(self, order):
        _check(self.L.agx_calculate_fluxes(self.h, order), "agx_calculate_fluxes")

    def corner_e(self):
        _check(self.L.agx_compute_corner_e(self.h), "agx_compute_corner_e")

    def integrate(self, stage):
        _check(self.L.agx_integrate(self.h, stage), "agx_integrate")

    def bvals(self, which=3):
        _check(self.L.agx_bvals_exchange(self.h, which), "agx_bvals_exchange")

    def cons2prim(self):
        _check(self.L.agx_cons2prim(self.h), "agx_cons2prim")

    def new_block_dt(self):
        v = ctypes.c_double()
        _check(self.L.agx_new_block_dt(self.h, ctypes.byref(v)), "agx_new_block_dt")
        return v.value

    def new_time_step(self, tlim):
        _check(self.L.agx_new_time_step(self.h, tlim), "agx_new_time_step")

    def initialize(self, tlim=1e300):
        _check(self.L.agx_initialize(self.h, tlim), "agx_initialize")

    def cycle(self, tlim, n=1):
        done = ctypes.c_int32()
        _check(self.L.agx_cycle(self.h, tlim, n, ctypes.byref(done)), "agx_cycle")
        return done.value

    def run(self, tlim, nlim=1 << 30):
        return self.cycle(tlim, nlim)

    def mg_solve(self, stage=0):
        _check(self.L.agx_mg_solve(self.h, stage), "agx_mg_solve")

    def mg_set_mode(self, mode, threshold, niter):
        _check(self.L.agx_mg_set_mode(self.h, mode, threshold, niter), "agx_mg_set_mode")

    def mg_last(self):
        v, n = ctypes.c_double(), ctypes.c_int32()
        _check(self.L.agx_mg_last(self.h, ctypes.byref(v), ctypes.byref(n)), "agx_mg_last")
        return v.value, n.value

    def mg_vcycles(self, n):
        _check(self.L.agx_mg_vcycles(self.h, n), "agx_mg_vcycles")

    def autotune(self):
        """Pick the fastest stage-kernel tile shape on the current state; returns its index."""
        c = ctypes.c_int32()
        _check(self.L.agx_autotune(self.h, ctypes.byref(c)), "agx_autotune")
        return c.value

    def set_graph(self, enable=True):
        """Replay agx_cycle as one hipGraph launch per cycle (no self-gravity)."""
        _check(self.L.agx_set_graph(self.h, 1 if enable else 0), "agx_set_graph")

    def set_mg_coarse(self, on):
        _check(lib().agx_set_mg_coarse(self.h, int(on)), "agx_set_mg_coarse")

    def mg_set_source_mask_sphere(self, xc, yc, zc, radius):
        _check(lib().agx_mg_set_source_mask_sphere(self.h, xc, yc, zc, radius), "agx_mg_set_source_mask_sphere")

    def set_mg_fuse(self, mode):
        _check(self.L.agx_set_mg_fuse(self.h, mode), "agx_set_mg_fuse")

    def set_tile(self, tile):
        _check(self.L.agx_set_tile(self.h, tile), "agx_set_tile")

    def timing(self, enable=True):
        _check(self.L.agx_timing_enable(self.h, 1 if enable else 0), "agx_timing_enable")

    def timing_reset(self):
        _check(self.L.agx_timing_reset(self.h), "agx_timing_reset")

    def timing_get(self, kclass):
        ms, n = ctypes.c_double(), ctypes.c_int64()
        _check(self.L.agx_timing_get(self.h, KCLASS[kclass], ctypes.byref(ms), ctypes.byref(n)),
               "agx_timing_get")
        return ms.value, n.value

    def device_bytes(self):
        return self.L.agx_mesh_bytes(self.h)

    def ptr(self, name):
        """device address of a registered array (agx_array_ptr)"""
        v = ctypes.c_void_p()
        _check(self.L.agx_array_ptr(self.h, ARR[name], ctypes.byref(v)), "agx_array_ptr")
        return v.value

    def stream(self):
        v = ctypes.c_void_p()
        _check(self.L.agx_mesh_stream(self.h, ctypes.byref(v)), "agx_mesh_stream")
        return v.value

    def set_levels(self, nrb, box, leaves):
        """agx_mesh_set_levels: block b becomes the MeshBlock at LogicalLocation leaves[b] =
        (level, lx1, lx2, lx3) of a root grid of nrb blocks over box = ((x1min, x1max), ...)"""
        n = len(leaves)
        lev = (ctypes.c_int32 * n)(*[int(x[0]) for x in leaves])
        lx = (ctypes.c_int64 * (3 * n))(*[int(v) for x in leaves for v in x[1:]])
        nr = (ctypes.c_int32 * 3)(*[int(x) for x in nrb])
        bx = (ctypes.c_double * 6)(*[float(v) for b in box for v in b])
        _check(self.L.agx_mesh_set_levels(self.h, nr, bx, lev, lx), "agx_mesh_set_levels")

    def mg_step(self, name, *args):
        """One per-level multigrid step: agx_mg_<name>(mesh, *args)."""
        _check(getattr(self.L, "agx_mg_" + name)(self.h, *args), "agx_mg_" + name)

    def mg_levels(self):
        n = ctypes.c_int32()
        dims = (ctypes.c_int32 * 96)()
        _check(self.L.agx_mg_levels(self.h, ctypes.byref(n), dims, 32), "agx_mg_levels")
        return [tuple(dims[3 * l:3 * l + 3]) for l in range(n.value)]

    def mg_defect_norm(self):
        v = ctypes.c_double()
        _check(self.L.agx_mg_defect_norm(self.h, ctypes.byref(v)), "agx_mg_defect_norm")
        return v.value

    def mg_coffset(self, flip=False):
        v = ctypes.c_int32()
        _check(self.L.agx_mg_coffset(self.h, 1 if flip else 0, ctypes.byref(v)), "agx_mg_coffset")
        return v.value

    def restart_write(self, path, params=None):
        """Write a restart file in the reference's format (agx_restart_write)."""
        _check(self.L.agx_restart_write(self.h, os.fsencode(path), None if params is None else params.encode()),
               "agx_restart_write")

    def vtk_write(self, basename, file_id="out1", file_number=0, variable="prim", ghost=False):
        """VTK legacy files of the current state, one per MeshBlock (agx_vtk_write)."""
        _check(self.L.agx_vtk_write(self.h, os.fsencode(basename), file_id.encode(), file_number, variable.encode(),
                                    1 if ghost else 0), "agx_vtk_write")

    def restart_read(self, path):
        """Load u, b, time, dt and ncycle from a restart file (agx_restart_read); call
        initialize() next, as the reference's Mesh::Initialize follows its restart
        constructor."""
        _check(self.L.agx_restart_read(self.h, os.fsencode(path)), "agx_restart_read")

    def history(self):
        """The built-in history sums (agx_history): mass, 3 momenta, 3 kinetic energies,
        [total energy], [gravitational energy], 3 magnetic energies."""
        out = np.zeros(16)
        _check(self.L.agx_history(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 16), "agx_history")
        n = (4 if self.desc.eos == 1 else 5) + 3 + (1 if self.desc.self_gravity else 0) + 3
        return out[:n].copy()


# ---------------------------------------------------------------- host utilities
def restart_write_host(path, desc, time, dt, ncycle, u, b1f, b2f, b3f, params=None):
    """agx_restart_write_host: a restart file from host arrays (no device work)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, b1f, b2f, b3f)]
    D = ctypes.POINTER(ctypes.c_double)
    _check(lib().agx_restart_write_host(os.fsencode(path), ctypes.byref(_desc(desc)),
                                        None if params is None else params.encode(), time, dt, ncycle,
                                        *[a.ctypes.data_as(D) for a in arrs]), "agx_restart_write_host")


def restart_read_host(path, desc, shapes):
    """agx_restart_read_host into new host arrays of the given shapes (u, b1f, b2f, b3f);
    returns (time, dt, ncycle, u, b1f, b2f, b3f)."""
    arrs = [np.zeros(s) for s in shapes]
    t, dt, n = ctypes.c_double(), ctypes.c_double(), ctypes.c_int32()
    D = ctypes.POINTER(ctypes.c_double)
    _check(lib().agx_restart_read_host(os.fsencode(path), ctypes.byref(_desc(desc)), ctypes.byref(t), ctypes.byref(dt),
                                       ctypes.byref(n), *[a.ctypes.data_as(D) for a in arrs]), "agx_restart_read_host")
    return (t.value, dt.value, n.value, *arrs)


def vtk_write_host(basename, desc, time, ncycle, u, w, bcc, phi=None, file_id="out1", file_number=0,
                   variable="prim", ghost=False):
    """agx_vtk_write_host: VTK legacy files from host arrays (no device work)."""
    D = ctypes.POINTER(ctypes.c_double)
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (u, w, bcc)]
    ph = None if phi is None else np.ascontiguousarray(phi, dtype=np.float64)
    _check(lib().agx_vtk_write_host(os.fsencode(basename), file_id.encode(), file_number, variable.encode(),
                                    1 if ghost else 0, ctypes.byref(_desc(desc)), time, ncycle,
                                    *[a.ctypes.data_as(D) for a in arrs],
                                    None if ph is None else ph.ctypes.data_as(D)), "agx_vtk_write_host")


def block_gids(desc):
    """agx_block_gids: the lattice index of each of the reference's MeshBlock gids."""
    nb = (desc.nx1 // desc.bnx1) * (desc.nx2 // desc.bnx2) * (desc.nx3 // desc.bnx3)
    out = np.zeros(nb, dtype=np.int64)
    _check(lib().agx_block_gids(ctypes.byref(_desc(desc)), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))),
           "agx_block_gids")
    return out


def _desc(desc):
    if isinstance(desc, Desc):
        return desc
    d = Desc()
    for name, _ in Desc._fields_:
        setattr(d, name, getattr(desc, name))
    return d

def block_lattice(desc):
    return desc.nx1 // desc.bnx1, desc.nx2 // desc.bnx2, desc.nx3 // desc.bnx3


def scatter_cells(desc, G):
    """Global interior array [nv, N3, N2, N1] -> block array [nb, nv, n3, n2, n1] (ghosts 0)."""
    ng = desc.nghost
    nb1, nb2, nb3 = block_lattice(desc)
    nv = G.shape[0]
    out = np.zeros((nb1 * nb2 * nb3, nv, desc.bnx3 + 2 * ng, desc.bnx2 + 2 * ng, desc.bnx1 + 2 * ng))
    for g in range(out.shape[0]):
        l1, l2, l3 = g % nb1, (g // nb1) % nb2, g // (nb1 * nb2)
        out[g, :, ng:ng + desc.bnx3, ng:ng + desc.bnx2, ng:ng + desc.bnx1] = \
            G[:, l3 * desc.bnx3:(l3 + 1) * desc.bnx3, l2 * desc.bnx2:(l2 + 1) * desc.bnx2,
              l1 * desc.bnx1:(l1 + 1) * desc.bnx1]
    return out


def gather_cells(desc, a):
    """Block array [nb, nv, n3, n2, n1] -> global interior [nv, N3, N2, N1]."""
    ng = desc.nghost
    nb1, nb2, nb3 = block_lattice(desc)
    nv = a.shape[1]
    G = np.zeros((nv, desc.nx3, desc.nx2, desc.nx1))
    for g in range(a.shape[0]):
        l1, l2, l3 = g % nb1, (g // nb1) % nb2, g // (nb1 * nb2)
        G[:, l3 * desc.bnx3:(l3 + 1) * desc.bnx3, l2 * desc.bnx2:(l2 + 1) * desc.bnx2,
          l1 * desc.bnx1:(l1 + 1) * desc.bnx1] = a[g, :, ng:ng + desc.bnx3, ng:ng + desc.bnx2,
                                                   ng:ng + desc.bnx1]
    return G


class RefList:
    """A device work list of one refinement operator (agx_ref_list_create).  The
    operator calls take device pointers (integers) to arrays in the layouts of
    include/athena_gamma_amd.h and run on the default stream; sync() waits."""

    def __init__(self, geom, kind, items, nb_fine, nb_coarse, stream=None):
        self.geom = geom
        self.stream = stream   # hipStream_t from agx_stream_create, or None (default stream)
        self.kind = REF_KINDS[kind] if isinstance(kind, str) else int(kind)
        arr = (RefItem * max(len(items), 1))(*items)
        self.h = ctypes.c_void_p()
        _check(lib().agx_ref_list_create(ctypes.byref(geom), self.kind, arr, len(items), nb_fine, nb_coarse,
                                         ctypes.byref(self.h)), "agx_ref_list_create")

    def close(self):
        if self.h:
            lib().agx_ref_list_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def restrict_cc(self, nvar, fine, fc, coarse):
        _check(lib().agx_ref_restrict_cc(self.h, nvar, fine, fc, coarse, self.stream), "agx_ref_restrict_cc")

    def restrict_fc(self, fine, fc, coarse):
        _check(lib().agx_ref_restrict_fc(self.h, fine, fc, coarse, self.stream), "agx_ref_restrict_fc")

    def prolong_cc(self, nvar, coarse, cc, fine, fc):
        _check(lib().agx_ref_prolong_cc(self.h, nvar, coarse, cc, fine, fc, self.stream), "agx_ref_prolong_cc")

    def prolong_fc_shared(self, coarse, cc, fine, fc):
        _check(lib().agx_ref_prolong_fc_shared(self.h, coarse, cc, fine, fc, self.stream), "agx_ref_prolong_fc_shared")

    def prolong_fc_internal(self, b1, b2, b3, fc):
        _check(lib().agx_ref_prolong_fc_internal(self.h, b1, b2, b3, fc, self.stream), "agx_ref_prolong_fc_internal")

    def flux_from_finer(self, nvar, f1, f2, f3, fc, c1, c2, c3):
        _check(lib().agx_ref_flux_from_finer(self.h, nvar, f1, f2, f3, fc, c1, c2, c3, self.stream),
               "agx_ref_flux_from_finer")

    def prim_to_cons(self, w, b1, b2, b3, bcc, u, gamma, iso=False):
        _check(lib().agx_ref_prim_to_cons(self.h, w, b1, b2, b3, bcc, u, gamma, int(iso), self.stream), "agx_ref_prim_to_cons")

    def emf_to_coarser(self, e1, e2, e3, fc, buf, stride):
        _check(lib().agx_ref_emf_to_coarser(self.h, e1, e2, e3, fc, buf, stride, self.stream), "agx_ref_emf_to_coarser")

    def copy_cc(self, nvar, src, dst):
        _check(lib().agx_ref_copy_cc(self.h, nvar, src, dst, self.stream), "agx_ref_copy_cc")

    def cons_to_prim(self, u, b1, b2, b3, w, bcc, gamma, dfloor, pfloor, iso=False):
        _check(lib().agx_ref_cons_to_prim(self.h, u, b1, b2, b3, w, bcc, gamma, dfloor, pfloor, int(iso), self.stream),
               "agx_ref_cons_to_prim")

    def pack_fc(self, f1, f2, f3, buf, stride):
        _check(lib().agx_ref_pack_fc(self.h, f1, f2, f3, buf, stride, self.stream), "agx_ref_pack_fc")

    def unpack_fc(self, f1, f2, f3, buf, stride):
        _check(lib().agx_ref_unpack_fc(self.h, f1, f2, f3, buf, stride, self.stream), "agx_ref_unpack_fc")

    def emf_same_level(self, e1, e2, e3, buf, stride):
        _check(lib().agx_ref_emf_same_level(self.h, e1, e2, e3, buf, stride, self.stream), "agx_ref_emf_same_level")

    def emf_set(self, e1, e2, e3, buf, stride):
        _check(lib().agx_ref_emf_set(self.h, e1, e2, e3, buf, stride, self.stream), "agx_ref_emf_set")

    def emf_clear(self, e1, e2, e3):
        _check(lib().agx_ref_emf_clear(self.h, e1, e2, e3, self.stream), "agx_ref_emf_clear")

    def emf_average(self, e1, e2, e3):
        _check(lib().agx_ref_emf_average(self.h, e1, e2, e3, self.stream), "agx_ref_emf_average")

    @staticmethod
    def sync(stream=None):
        _check(lib().agx_ref_sync(stream), "agx_ref_sync")


def ref_emf_count(geom, fid):
    return int(lib().agx_ref_emf_count(ctypes.byref(geom), fid))


def ref_emf_count_same_level(geom, fid):
    return int(lib().agx_ref_emf_count_same_level(ctypes.byref(geom), fid))


def ref_count_fine_edges(nblevel):
    """agx_ref_count_fine_edges: (edge_flag[12] as bools, nedge_fine[12])."""
    I27, I12 = ctypes.c_int32 * 27, ctypes.c_int32 * 12
    fl, nf = I12(), I12()
    _check(lib().agx_ref_count_fine_edges(I27(*nblevel), fl, nf), "agx_ref_count_fine_edges")
    return [bool(x) for x in fl], list(nf)


def ref_prolong_ranges(geom, ox, lx_odd, nblevel):
    """agx_ref_prolong_ranges: [si, ei, sj, ej, sk, ek, il, iu, jl, ju, kl, ku]."""
    I3, I27, I12 = ctypes.c_int32 * 3, ctypes.c_int32 * 27, ctypes.c_int32 * 12
    out = I12()
    _check(lib().agx_ref_prolong_ranges(ctypes.byref(geom), I3(*ox), I3(*lx_odd), I27(*nblevel), out),
           "agx_ref_prolong_ranges")
    return list(out)


class DeviceArray:
    """A device copy of a host array (agx_device_alloc / agx_device_copy).  `ptr` is
    the device address the refinement operators take; get() copies back."""

    def __init__(self, host):
        host = np.ascontiguousarray(host)
        self.shape, self.dtype, self.nbytes = host.shape, host.dtype, host.nbytes
        self.p = ctypes.c_void_p()
        _check(lib().agx_device_alloc(self.nbytes, ctypes.byref(self.p)), "agx_device_alloc")
        _check(lib().agx_device_copy(self.p, host.ctypes.data_as(ctypes.c_void_p), self.nbytes, 0),
               "agx_device_copy")

    @property
    def ptr(self):
        return self.p.value

    def set(self, host):
        """copy a host array of the same shape and dtype in (synchronous)"""
        host = np.ascontiguousarray(host, dtype=self.dtype)
        if host.shape != self.shape:
            raise AgxError(f"DeviceArray.set: shape {host.shape} != {self.shape}")
        _check(lib().agx_device_copy(self.p, host.ctypes.data_as(ctypes.c_void_p), self.nbytes, 0),
               "agx_device_copy")

    def get(self):
        out = np.empty(self.shape, self.dtype)
        _check(lib().agx_device_copy(out.ctypes.data_as(ctypes.c_void_p), self.p, self.nbytes, 1),
               "agx_device_copy")
        return out

    def close(self):
        if self.p:
            lib().agx_device_free(self.p)
            self.p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def device_coords(tab):
    """RefCoords over device copies of per-block tables x1v, x2v, x3v, dx1f, dx2f, dx3f;
    returns (coords, arrays) - keep the arrays alive while the coords are in use."""
    arrs = [DeviceArray(tab[k]) for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")]
    return RefCoords(*[a.ptr for a in arrs]), arrs


def _load_smr():
    import importlib.util
    name = "agx_smr"
    if name in sys.modules:
        return sys.modules[name]
    spec = importlib.util.spec_from_file_location(name, os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                                     "smr.py"))
    mod = importlib.util.module_from_spec(spec)
    sys.modules[name] = mod
    spec.loader.exec_module(mod)
    return mod


class SMRBoundary:
    """The boundary step of a static-refinement MHD mesh on the device, with every work
    list built once from the block tree (smr.py):
      exchange_cells(u, cu)           CellCenteredBoundaryVariable across levels
      exchange_faces(b, cb)           FaceCenteredBoundaryVariable across levels
      prolongate(u, b, w, bcc, cu, cb, cw, cbcc)
                                      BoundaryValues::ProlongateBoundaries (MHD), per
                                      slot of coarser neighbours
      correct_emf(e1, e2, e3)         SendFluxCorrection / ReceiveFluxCorrection
    Arguments are device pointers (integers) to the reference's dense layouts, block-
    major: u, w (nb, nhydro, n, n, n), bcc (nb, 3, n, n, n), b = (x1f, x2f, x3f)
    (nb, n, n, n+1) ..., coarse buffers cu, cw, cbcc, cb with nc = nx/2 + 2 ng, edges
    e1 (nb, n+1, n+1, n) ...  Cubic blocks of nx^3 cells, ng ghosts, the adiabatic EOS.
    Launches are asynchronous on the driver's own stream (agx_stream_create), on
    `stream` (for instance a mesh's, agx_mesh_stream) or, with own_stream=False, on the
    default stream; sync() waits.  layout: AGX_REF_LAYOUT_DENSE (0) for the dense arrays
    above, AGX_REF_LAYOUT_MESH (1) for the block arrays of an agx_mesh (agx_array_ptr).
    tables: the blocks' and coarse buffers' coordinate tables (fine, coarse), by default
    smr.block_coords over [0, box).  capture(...) records one boundary step (and
    optionally the EMF correction) with fixed arrays into a hipGraph; replay() launches it
    (about 100 kernels in one launch)."""

    def __init__(self, tree, nx, ng, gamma, dfloor, pfloor, box=(1.0, 1.0, 1.0), own_stream=True, stream=None,
                 layout=0, tables=None, iso=False):
        smr = _load_smr()
        self.stream, self.graph, self.own = None, None, False
        if stream is not None:
            self.stream = stream
        elif own_stream:
            st = ctypes.c_void_p()
            _check(lib().agx_stream_create(ctypes.byref(st)), "agx_stream_create")
            self.stream, self.own = st.value, True
        self.tree, self.nx, self.ng = tree, nx, ng
        self.gamma, self.dfloor, self.pfloor = gamma, dfloor, pfloor
        self.iso, self.nh = bool(iso), 4 if iso else 5   # hydro variables (NHYDRO)
        nb = self.nb = len(tree.leaves)
        self.geom = RefGeom(nx, nx, nx, ng, ng, layout)
        self._keep = []
        if tables is None:
            tables = (smr.block_coords(tree, nx, ng, box=box), smr.block_coords(tree, nx, ng, coarse=True, box=box))
        self.fc, a1 = device_coords(tables[0])
        self.cc, a2 = device_coords(tables[1])
        self._keep += a1 + a2
        self._lists = []
        L = self._list
        ccp, fcp = smr.cc_plan(tree, nx, ng), smr.fc_plan(tree, nx, ng)
        slots, _ = smr.mhd_prolong_plan(tree, nx, ng, ccp, fcp)
        self.c_same, self.c_send = L("copy_cc_bb", ccp.copy_same), L("restrict_cc", ccp.restrict_send)
        self.c_from_finer, self.c_to_finer = L("copy_cc_cb", ccp.copy_from_finer), L("copy_cc_bc", ccp.copy_to_finer)
        self.f_send = [L(f"restrict_fc{c + 1}", fcp.restrict_send[c]) for c in range(3)]
        self.nslot_b = len(fcp.pack_b)
        nbuf = max(1, len(fcp.pack_b) + len(fcp.pack_c))
        self.f_stride = max([1] + [_box_size(d) for d in fcp.pack_b + fcp.pack_c])
        self.f_pack_b, self.f_pack_c = L("pack_fc_b", fcp.pack_b), L("pack_fc_c", fcp.pack_c)
        self.f_unpack_b = L("unpack_fc_b", fcp.unpack_b, nbuf)
        self.f_unpack_c = L("unpack_fc_c", fcp.unpack_c, nbuf)
        self.f_buf = DeviceArray(np.zeros((nbuf, self.f_stride)))
        self.slots = []
        for st in slots:
            self.slots.append(dict(
                rcc=L("restrict_cc", st["restrict_cc"]),
                rfc=[L(f"restrict_fc{c + 1}", st["restrict_fc"][c]) for c in range(3)],
                c2p=L("cons_to_prim", st["c2p"]), pcc=L("prolong_cc", st["prolong_cc"]),
                pfc=[L(f"prolong_fc{c + 1}", st["prolong_fc"][c]) for c in range(3)],
                pint=L("prolong_fc_int", st["prolong_int"]), p2c=L("prim_to_cons", st["p2c"])))
        q = smr.emf_plan(tree)
        g = self.geom
        self.e_ss = max([1] + [ref_emf_count_same_level(g, d["fid"]) for d in q.same_level])
        self.e_sc = max([1] + [ref_emf_count(g, d["fid"]) for d in q.to_coarser])
        self.e_buf_s = DeviceArray(np.zeros((max(1, len(q.same_level)), self.e_ss)))
        self.e_buf_c = DeviceArray(np.zeros((max(1, len(q.to_coarser)), self.e_sc)))
        self.e_same, self.e_coarse = L("emf_same_level", q.same_level), L("emf_to_coarser", q.to_coarser)
        self.e_set_same = L("emf_set_same", q.set_same, max(1, len(q.same_level)))
        self.e_clear = L("emf_clear", q.clear)
        self.e_set_finer = L("emf_set_finer", q.set_finer, max(1, len(q.to_coarser)))
        self.e_avg = L("emf_average", q.average)
        self._keep += [self.f_buf, self.e_buf_s, self.e_buf_c]

    def _list(self, kind, items, nbf=None):
        lst = RefList(self.geom, kind, [RefItem(**d) for d in items], self.nb if nbf is None else nbf, self.nb,
                      stream=self.stream)
        self._lists.append(lst)
        return lst

    def exchange_cells(self, u, cu, nvar=None):
        nvar = self.nh if nvar is None else nvar
        self.c_same.copy_cc(nvar, u, u)
        self.c_send.restrict_cc(nvar, u, self.fc, cu)
        self.c_from_finer.copy_cc(nvar, cu, u)
        self.c_to_finer.copy_cc(nvar, u, cu)

    def exchange_faces(self, b, cb):
        for c in range(3):
            self.f_send[c].restrict_fc(b[c], self.fc, cb[c])
        p = self.f_buf.ptr
        self.f_pack_b.pack_fc(*b, p, self.f_stride)
        self.f_pack_c.pack_fc(*cb, p + self.nslot_b * self.f_stride * 8, self.f_stride)
        self.f_unpack_b.unpack_fc(*b, p, self.f_stride)
        self.f_unpack_c.unpack_fc(*cb, p, self.f_stride)

    def prolongate(self, u, b, w, bcc, cu, cb, cw, cbcc, fused=True):
        if fused:   # four launches per slot
            L = lib()
            for s in self.slots:
                _check(L.agx_ref_restrict_all(s["rcc"].h, *[x.h for x in s["rfc"]], self.nh, u, *b, self.fc, cu,
                                              *cb, self.stream), "agx_ref_restrict_all")
                s["c2p"].cons_to_prim(cu, *cb, cw, cbcc, self.gamma, self.dfloor, self.pfloor, self.iso)
                _check(L.agx_ref_prolong_all(s["pcc"].h, *[x.h for x in s["pfc"]], self.nh, cw, *cb, self.cc, w,
                                             *b, self.fc, self.stream), "agx_ref_prolong_all")
                _check(L.agx_ref_internal_to_cons(s["pint"].h, *b, self.fc, w, bcc, u, self.gamma, int(self.iso),
                                                  self.stream), "agx_ref_internal_to_cons")
            return
        for s in self.slots:
            s["rcc"].restrict_cc(self.nh, u, self.fc, cu)
            for c in range(3):
                s["rfc"][c].restrict_fc(b[c], self.fc, cb[c])
            s["c2p"].cons_to_prim(cu, *cb, cw, cbcc, self.gamma, self.dfloor, self.pfloor, self.iso)
            s["pcc"].prolong_cc(self.nh, cw, self.cc, w, self.fc)
            for c in range(3):
                s["pfc"][c].prolong_fc_shared(cb[c], self.cc, b[c], self.fc)
            s["pint"].prolong_fc_internal(*b, self.fc)
            s["p2c"].prim_to_cons(w, *b, bcc, u, self.gamma, self.iso)

    def correct_emf(self, e1, e2, e3):
        self.e_same.emf_same_level(e1, e2, e3, self.e_buf_s.ptr, self.e_ss)
        self.e_coarse.emf_to_coarser(e1, e2, e3, self.fc, self.e_buf_c.ptr, self.e_sc)
        self.e_set_same.emf_set(e1, e2, e3, self.e_buf_s.ptr, self.e_ss)
        self.e_clear.emf_clear(e1, e2, e3)
        self.e_set_finer.emf_set(e1, e2, e3, self.e_buf_c.ptr, self.e_sc)
        self.e_avg.emf_average(e1, e2, e3)

    def boundary_step(self, u, b, w, bcc, cu, cb, cw, cbcc, fused=True):
        """ghost zones of u and b everywhere: both exchanges, then the prolongation"""
        self.exchange_cells(u, cu)
        self.exchange_faces(b, cb)
        self.prolongate(u, b, w, bcc, cu, cb, cw, cbcc, fused)

    def capture(self, u, b, w, bcc, cu, cb, cw, cbcc, emf=None):
        """record boundary_step (then correct_emf(*emf) if given) into a graph"""
        if not self.stream:
            raise AgxError("capture needs the driver's own stream")
        self.sync()
        _check(lib().agx_graph_begin(self.stream), "agx_graph_begin")
        try:
            self.boundary_step(u, b, w, bcc, cu, cb, cw, cbcc)
            if emf is not None:
                self.correct_emf(*emf)
        finally:
            ex = ctypes.c_void_p()
            rc = lib().agx_graph_end(self.stream, ctypes.byref(ex))
        _check(rc, "agx_graph_end")
        if self.graph:
            lib().agx_graph_destroy(self.graph)
        self.graph = ex.value

    def replay(self):
        _check(lib().agx_graph_launch(self.graph, self.stream), "agx_graph_launch")

    def sync(self):
        RefList.sync(self.stream)

    def close(self):
        if self.graph:
            lib().agx_graph_destroy(self.graph)
            self.graph = None
        for x in self._lists:
            x.close()
        for a in self._keep:
            a.close()
        self._lists, self._keep = [], []
        if self.stream and self.own:
            lib().agx_stream_destroy(self.stream)
        self.stream = None


class SMRMesh:
    """MHD on a static-refinement mesh on the device: the leaves of `tree` (smr.SMRTree)
    as the MeshBlocks of one agx_mesh (agx_mesh_set_levels), cubic blocks of nx^3 cells,
    ng ghosts (even), FP64, adiabatic EOS or isothermal (eos = 1, NHYDRO = 4, sound speed
    iso_cs), periodic or block boundaries only.  Each VL2 stage runs the task order of
    TimeIntegratorTaskList with refinement (time_integrator.cpp:1460-1558):
      CalculateHydroFlux              agx_calculate_fluxes (donor cell in stage 1)
      Send/ReceiveAndCorrectHydroFlux flux_from_finer (flux_correction_cc.cpp:77-283)
      CalculateEMF                    agx_compute_corner_e
      Send/ReceiveAndCorrectEMF       SMRBoundary.correct_emf (flux_correction_fc.cpp)
      IntegrateHydro, IntegrateField  agx_integrate
      Send/Receive/SetBoundaries, Prolongation
                                      SMRBoundary.boundary_step on the mesh's own arrays
      Primitives                      agx_cons2prim (interior and ghosts)
    then time += dt and Mesh::NewTimeStep (main.cpp:459-489, mesh.cpp:1078-1119).  Every
    launch runs on the mesh's stream.  Coordinates follow the reference's Coordinates on
    a uniform mesh (smr.mesh_coords)."""

    def __init__(self, tree, nx, ng=2, box=((0.0, 1.0),) * 3, gamma=5.0 / 3.0, cfl=0.3, xorder=2, rsolver=0, eos=0,
                 iso_cs=1.0):
        smr = _load_smr()
        self.tree, self.nx, self.ng, self.xorder = tree, nx, ng, xorder
        nb = self.nb = len(tree.leaves)
        self.nh = 4 if eos == 1 else 5
        desc = make_desc(nx=(nb * nx, nx, nx), bnx=(nx, nx, nx), box=((0.0, float(nb)), (0.0, 1.0), (0.0, 1.0)),
                         nghost=ng, xorder=xorder, rsolver=rsolver, gamma=gamma, cfl=cfl, eos=eos, iso_cs=iso_cs)
        self.desc = desc
        self.m = Mesh(desc)
        self._keep = []
        try:
            self.m.set_levels(tree.nrb, box, tree.leaves)
            tabs = (smr.mesh_coords(tree, nx, ng, box), smr.mesh_coords(tree, nx, ng, box, coarse=True))
            self.tables = tabs
            self.bnd = SMRBoundary(tree, nx, ng, gamma, desc.dfloor, desc.pfloor, stream=self.m.stream(), layout=1,
                                   tables=tabs, iso=eos == 1)
            plan = smr.flux_plan(tree)
            self.flux = RefList(self.bnd.geom, "flux_from_finer", [RefItem(**d) for d in plan], nb, nb,
                                stream=self.m.stream()) if plan else None
            nc = nx // 2 + 2 * ng
            self.cu = DeviceArray(np.zeros((nb, self.nh, nc, nc, nc)))
            self.cw = DeviceArray(np.zeros((nb, self.nh, nc, nc, nc)))
            self.cbcc = DeviceArray(np.zeros((nb, 3, nc, nc, nc)))
            self.cb = [DeviceArray(np.zeros(sh)) for sh in ((nb, nc, nc, nc + 1), (nb, nc, nc + 1, nc),
                                                              (nb, nc + 1, nc, nc))]
            self._keep += [self.cu, self.cw, self.cbcc] + self.cb
        except Exception:
            self.close()
            raise

    def set(self, name, arr):
        self.m.set(name, arr)

    def get(self, name):
        return self.m.get(name)

    def time(self):
        return self.m.time()

    def sync(self):
        self.m.sync()

    def boundary(self):
        """ghost zones of u and b across levels, then ProlongateBoundaries"""
        p = self.m.ptr
        self.bnd.boundary_step(p("u"), [p("b1f"), p("b2f"), p("b3f")], p("w"), p("bcc"), self.cu.ptr,
                               [x.ptr for x in self.cb], self.cw.ptr, self.cbcc.ptr)

    def initialize(self, tlim=1e300):
        """Mesh::Initialize after the problem generator: ghost zones, primitives, first dt"""
        self.boundary()
        self.m.cons2prim()
        t, _, n = self.m.time()
        self.m.set_time(t, 1.7976931348623157e308, n)   # a fresh Mesh starts from dt = DBL_MAX
        self.m.new_time_step(tlim)

    def stage(self, s):
        m, p = self.m, self.m.ptr
        m.calculate_fluxes(1 if s == 1 else self.xorder)
        if self.flux is not None:
            f = [p("fx1"), p("fx2"), p("fx3")]
            self.flux.flux_from_finer(self.nh, *f, self.bnd.fc, *f)
        m.corner_e()
        self.bnd.correct_emf(p("e1"), p("e2"), p("e3"))
        m.integrate(s)
        self.boundary()
        m.cons2prim()

    def cycle(self, tlim, n=1):
        """up to n cycles while time < tlim; returns the number run"""
        done = 0
        while done < n:
            t, dt, nc = self.m.time()
            if t >= tlim:
                break
            self.stage(1)
            self.stage(2)
            self.m.set_time(t + dt, dt, nc + 1)
            self.m.new_time_step(tlim)
            done += 1
        return done

    def close(self):
        for x in (getattr(self, "flux", None), getattr(self, "bnd", None)):
            if x is not None:
                x.close()
        for a in getattr(self, "_keep", []):
            a.close()
        self._keep, self.flux, self.bnd = [], None, None
        if getattr(self, "m", None) is not None:
            self.m.close()
            self.m = None


def _box_size(d):
    return (d["ei"] - d["si"] + 1) * (d["ej"] - d["sj"] + 1) * (d["ek"] - d["sk"] + 1)
