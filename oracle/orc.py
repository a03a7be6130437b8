"""ctypes binding of the CPU oracle (oracle/liborc.so).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liborc.so")

# array ids shared with the product C-ABI (AGX_ARR_* in include/athena_gamma_amd.h)
ARR = dict(u=0, u1=1, w=2, bcc=3, b1f=4, b2f=5, b3f=6, r1f=7, r2f=8, r3f=9,
           fx1=10, fx2=11, fx3=12, e3x1=13, e2x1=14, e1x2=15, e3x2=16, e1x3=17,
           e2x3=18, wx1=19, wx2=20, wx3=21, e1=22, e2=23, e3=24, phi=25)


class Desc(ctypes.Structure):
    """Mesh description; identical layout to agx_mesh_desc (product) and orc_desc."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "nx1", "nx2", "nx3", "bnx1", "bnx2", "bnx3", "nghost", "xorder", "rsolver",
        "eos", "self_gravity", "mg_mode", "mg_niter", "precision")] + \
        [(n, ctypes.c_double) for n in (
            "x1min", "x1max", "x2min", "x2max", "x3min", "x3max", "gamma", "iso_cs",
            "cfl", "dfloor", "pfloor", "four_pi_G", "mg_threshold")] + \
        [(n, ctypes.c_int32) for n in ("mg_bc", "mporder", "mp_auto_origin")] + \
        [("mp_origin", ctypes.c_double * 3), ("mp_nodipole", ctypes.c_int32), ("char_proj", ctypes.c_int32)]


FLT_MIN = np.finfo(np.float32).tiny


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
    # EquationOfState defaults, eos/adiabatic_mhd.cpp:26-31: sqrt(1024*FLT_MIN)
    d.dfloor = d.pfloor = float(np.sqrt(1024 * np.float64(FLT_MIN)))
    d.four_pi_G, d.mg_threshold = four_pi_G, mg_threshold
    d.mg_bc, d.mporder, d.mp_auto_origin = mg_bc, mporder, mp_auto_origin
    d.mp_origin[:] = mp_origin
    d.mp_nodipole = mp_nodipole
    d.char_proj = char_proj
    return d


def _lib():
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    D = ctypes.POINTER(ctypes.c_double)
    lib.orc_create.restype = P
    lib.orc_create.argtypes = [ctypes.POINTER(Desc)]
    lib.orc_destroy.argtypes = [P]
    lib.orc_size.restype = ctypes.c_int64
    lib.orc_size.argtypes = [P, ctypes.c_int]
    lib.orc_array.restype = D
    lib.orc_array.argtypes = [P, ctypes.c_int]
    lib.orc_set_time.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int]
    lib.orc_get_time.argtypes = [P, D]
    lib.orc_coords.argtypes = [P, ctypes.c_int, ctypes.c_int, D, D, D]
    lib.orc_set_levels.argtypes = [P, ctypes.POINTER(ctypes.c_int), D, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int64)]
    for f in ("orc_bvals_cc", "orc_bvals_fc", "orc_calculate_fluxes", "orc_integrate",
              "orc_mg_solve"):
        getattr(lib, f).argtypes = [P, ctypes.c_int]
    for f in ("orc_cons2prim", "orc_corner_e", "orc_initialize", "orc_initialize_restarted"):
        getattr(lib, f).argtypes = [P]
    lib.orc_new_block_dt.restype = ctypes.c_double
    lib.orc_new_block_dt.argtypes = [P]
    lib.orc_new_time_step.argtypes = [P, ctypes.c_double]
    lib.orc_cycle.argtypes = [P, ctypes.c_double]
    lib.orc_mg_last_norm.restype = ctypes.c_double
    lib.orc_mg_last_norm.argtypes = [P]
    lib.orc_mg_last_iters.restype = ctypes.c_int
    lib.orc_mg_last_iters.argtypes = [P]
    lib.orc_mg_multipole.argtypes = [P, ctypes.POINTER(ctypes.c_double)]
    lib.orc_mp_phis.restype = ctypes.c_double
    lib.orc_mp_phis.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double]
    lib.orc_mg_set_mode.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_int]
    lib.orc_mg_set_source_mask_sphere.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                                  ctypes.c_double]
    lib.orc_pgen_linear_wave.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double]
    lib.orc_linear_wave_errors.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double, D]
    lib.orc_max_divb.restype = ctypes.c_double
    lib.orc_max_divb.argtypes = [P]
    lib.orc_conserved_totals.argtypes = [P, D]
    lib.orc_history.restype = ctypes.c_int
    lib.orc_history.argtypes = [P, D]
    lib.orc_riemann.argtypes = [ctypes.c_int, ctypes.c_double, D, D, ctypes.c_double, D]
    lib.orc_riemann_iso.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, D, D,
                                    ctypes.c_double, D]
    lib.orc_riemann_lhlld.argtypes = [ctypes.c_double, D, D, ctypes.c_double, ctypes.c_double, ctypes.c_double, D]
    lib.orc_plm_pencil.argtypes = [ctypes.c_int, D, D, D, D, D]
    lib.orc_ppm_pencil.argtypes = [ctypes.c_int, D, D, D]
    lib.orc_set_noise.argtypes = [P, ctypes.c_double, ctypes.c_uint64]
    lib.orc_char_matrices.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, D,
                                      ctypes.c_double, D, D]
    I = ctypes.c_int
    lib.orc_recon_cell.argtypes = [P, I, I, I, I, I, I, D, D]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _LIB = _lib()
    return _LIB


class Oracle:
    """One oracle mesh (all MeshBlocks of a uniform periodic mesh)."""

    def __init__(self, desc):
        self.desc = desc
        self.L = lib()
        self.h = self.L.orc_create(ctypes.byref(desc))
        if not self.h:
            raise RuntimeError("orc_create failed")
        self.nb = self.L.orc_size(self.h, 7)
        ng = desc.nghost
        self.n = (desc.bnx1 + 2 * ng, desc.bnx2 + 2 * ng, desc.bnx3 + 2 * ng)

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def shape(self, name):
        n1, n2, n3 = self.n
        nb = self.nb
        cell = {"u": 5, "u1": 5, "w": 5, "bcc": 3, "phi": 1}
        if name in cell:
            return (nb, cell[name], n3, n2, n1)
        face = {"b1f": (n3, n2, n1 + 1), "r1f": (n3, n2, n1 + 1), "wx1": (n3, n2, n1 + 1),
                "e3x1": (n3, n2, n1 + 1), "e2x1": (n3, n2, n1 + 1),
                "b2f": (n3, n2 + 1, n1), "r2f": (n3, n2 + 1, n1), "wx2": (n3, n2 + 1, n1),
                "e1x2": (n3, n2 + 1, n1), "e3x2": (n3, n2 + 1, n1),
                "b3f": (n3 + 1, n2, n1), "r3f": (n3 + 1, n2, n1), "wx3": (n3 + 1, n2, n1),
                "e1x3": (n3 + 1, n2, n1), "e2x3": (n3 + 1, n2, n1),
                "e1": (n3 + 1, n2 + 1, n1), "e2": (n3 + 1, n2, n1 + 1), "e3": (n3, n2 + 1, n1 + 1)}
        if name in face:
            return (nb,) + face[name]
        fl = {"fx1": (5, n3, n2, n1 + 1), "fx2": (5, n3, n2 + 1, n1), "fx3": (5, n3 + 1, n2, n1)}
        return (nb,) + fl[name]

    def view(self, name):
        """numpy view (no copy) of an oracle array; re-fetch after register swaps."""
        shp = self.shape(name)
        ptr = self.L.orc_array(self.h, ARR[name])
        return np.ctypeslib.as_array(ptr, shape=shp)

    def get(self, name):
        return self.view(name).copy()

    def set_noise(self, eps, seed=0):
        """test-only: perturb w and bcc after every C2P of a cycle by a relative eps
        (uniform, seeded), or round them to FP32 (eps < 0); 0 turns it off"""
        self.L.orc_set_noise(self.h, eps, seed)

    def recon_cell(self, dir, order, g, k, j, i):
        """(wm, wp): the reconstructed states of one cell at its lower and upper faces along
        dir, 7 slots in the mesh frame (IBY, IBZ = the field transverse to dir)."""
        wm, wp = np.zeros(7), np.zeros(7)
        self.L.orc_recon_cell(self.h, dir, order, g, k, j, i, _dp(wm), _dp(wp))
        return wm, wp

    def set(self, name, arr):
        self.view(name)[...] = arr

    def coords(self, gid, d):
        n = self.n[d]
        xf = np.zeros(n + 1)
        xv = np.zeros(n)
        dx = np.zeros(n)
        D = ctypes.POINTER(ctypes.c_double)
        self.L.orc_coords(self.h, gid, d, xf.ctypes.data_as(D), xv.ctypes.data_as(D),
                          dx.ctypes.data_as(D))
        return xf, xv, dx

    def set_levels(self, nrb, box, leaves):
        """block g becomes the MeshBlock at LogicalLocation leaves[g] = (level, lx1, lx2,
        lx3) of a root grid of nrb blocks over box = ((x1min, x1max), ...)"""
        n = len(leaves)
        lev = (ctypes.c_int * n)(*[int(x[0]) for x in leaves])
        lx = (ctypes.c_int64 * (3 * n))(*[int(v) for x in leaves for v in x[1:]])
        nr = (ctypes.c_int * 3)(*[int(x) for x in nrb])
        bx = (ctypes.c_double * 6)(*[float(v) for b in box for v in b])
        self.L.orc_set_levels(self.h, nr, bx, lev, lx)

    def time(self):
        out = np.zeros(3)
        self.L.orc_get_time(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out[0], out[1], int(out[2])

    def set_time(self, t, dt, ncycle=0):
        self.L.orc_set_time(self.h, t, dt, ncycle)

    # reference steps
    def initialize(self, restarted=False):
        """Mesh::Initialize; restarted: after a restart file's time, dt and ncycle were set
        (set_time), so the first time step is capped by twice that dt"""
        (self.L.orc_initialize_restarted if restarted else self.L.orc_initialize)(self.h)

    def cycle(self, tlim):
        self.L.orc_cycle(self.h, tlim)

    def run(self, tlim, nlim=-1):
        n = 0
        while self.time()[0] < tlim and (nlim < 0 or n < nlim):
            self.cycle(tlim)
            n += 1
        return n

    def pgen_linear_wave(self, wave_flag=0, amp=1e-6, vflow=0.0):
        self.L.orc_pgen_linear_wave(self.h, wave_flag, amp, vflow)

    def linear_wave_errors(self, wave_flag=0, amp=1e-6, vflow=0.0):
        out = np.zeros(18)
        self.L.orc_linear_wave_errors(self.h, wave_flag, amp, vflow,
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def max_divb(self):
        return self.L.orc_max_divb(self.h)

    def totals(self):
        out = np.zeros(8)
        self.L.orc_conserved_totals(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def history(self):
        """history.cpp built-in sums (orc_history)."""
        out = np.zeros(16)
        n = self.L.orc_history(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out[:n].copy()

    def mg_solve(self, stage=0):
        self.L.orc_mg_solve(self.h, stage)

    def mg_set_mode(self, mode, threshold, niter):
        self.L.orc_mg_set_mode(self.h, mode, threshold, niter)

    def mg_set_source_mask_sphere(self, xc, yc, zc, radius):
        self.L.orc_mg_set_source_mask_sphere(self.h, xc, yc, zc, radius)

    def mg_last(self):
        return self.L.orc_mg_last_norm(self.h), self.L.orc_mg_last_iters(self.h)

    def mg_multipole(self):
        """(origin[3], coefficients[25]) of the last multipole solve."""
        out = np.zeros(28)
        self.L.orc_mg_multipole(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out[:3], out[3:]

    def mp_phis(self, x, y, z):
        return self.L.orc_mp_phis(self.h, x, y, z)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def riemann(rs, gamma, wl, wr, bx):
    """One interface flux [7] from the oracle's HLLD (rs=0) or HLLE (rs=1)."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    out = np.zeros(7)
    lib().orc_riemann(rs, gamma, _dp(wl), _dp(wr), bx, _dp(out))
    return out


def riemann_lhlld(gamma, wl, wr, bx, dvn=0.0, dvt=0.0):
    """One interface flux [7] from the oracle's LHLLD with the shock detector's normal and
    transverse velocity differences dvn, dvt."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    out = np.zeros(7)
    lib().orc_riemann_lhlld(gamma, _dp(wl), _dp(wr), bx, dvn, dvt, _dp(out))
    return out


def riemann_iso(rs, cs, wl, wr, bx, dfloor=None):
    """One interface flux [7] (energy slot 0) from the oracle's isothermal HLLD (rs=0) or
    HLLE (rs=1)."""
    wl = np.ascontiguousarray(wl, dtype=np.float64)
    wr = np.ascontiguousarray(wr, dtype=np.float64)
    out = np.zeros(7)
    dfl = float(np.sqrt(1024 * np.float64(FLT_MIN))) if dfloor is None else dfloor
    lib().orc_riemann_iso(rs, cs, dfl, _dp(wl), _dp(wr), bx, _dp(out))
    return out


def plm_pencil(q, dxp=None, dxm=None):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.size
    dxp = np.full(n, 0.5) if dxp is None else np.ascontiguousarray(dxp, dtype=np.float64)
    dxm = np.full(n, 0.5) if dxm is None else np.ascontiguousarray(dxm, dtype=np.float64)
    wl = np.full(n, np.nan)
    wr = np.full(n, np.nan)
    lib().orc_plm_pencil(n, _dp(q), _dp(dxp), _dp(dxm), _dp(wl), _dp(wr))
    return wl, wr


def char_matrices(w7, b1, dir=0, gamma=5.0 / 3.0, iso=False, iso_cs=1.0):
    """(L, R) of the eigenvector projections at one state: L @ v gives the characteristic
    fields of the primitive vector v, R @ ch the primitive vector (7 x 7; isothermal: the
    first 6 fields)."""
    w7 = np.ascontiguousarray(w7, dtype=np.float64)
    L = np.zeros((7, 7))
    R = np.zeros((7, 7))
    lib().orc_char_matrices(int(iso), gamma, iso_cs, dir, _dp(w7), b1, _dp(L), _dp(R))
    return L, R


def ppm_pencil(q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q.size
    wl = np.full(n, np.nan)
    wr = np.full(n, np.nan)
    lib().orc_ppm_pencil(n, _dp(q), _dp(wl), _dp(wr))
    return wl, wr


# ---------------------------------------------------------------- refinement operators
class RefItem(ctypes.Structure):
    """orc_ref_item (same layout as agx_ref_item)."""
    _fields_ = [(n, ctypes.c_int32) for n in ("fb", "cb", "si", "ei", "sj", "ej", "sk", "ek",
                                              "fid", "fi1", "fi2", "aux")]


class RefGeom(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("nx1", "nx2", "nx3", "ng", "cng", "pad")]


class RefCoords(ctypes.Structure):
    _fields_ = [(n, ctypes.c_void_p) for n in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")]


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.c_void_p)


def ref_coords(tabs):
    """tabs: dict of contiguous float64 arrays x1v, x2v, x3v, dx1f, dx2f, dx3f (nb, n)."""
    return RefCoords(*[_ptr(tabs[k]) for k in ("x1v", "x2v", "x3v", "dx1f", "dx2f", "dx3f")])


def _items(items):
    return (RefItem * max(len(items), 1))(*items), len(items)


def _ref_lib():
    L = lib()
    if not getattr(L, "_ref_bound", False):
        G, C, IT, V = ctypes.POINTER(RefGeom), RefCoords, ctypes.POINTER(RefItem), ctypes.c_void_p
        L.orc_ref_restrict_cc.argtypes = [G, ctypes.c_int, V, C, V, IT, ctypes.c_int]
        L.orc_ref_restrict_fc.argtypes = [G, ctypes.c_int, V, C, V, IT, ctypes.c_int]
        L.orc_ref_prolong_cc.argtypes = [G, ctypes.c_int, V, C, V, C, IT, ctypes.c_int]
        L.orc_ref_prolong_fc_shared.argtypes = [G, ctypes.c_int, V, C, V, C, IT, ctypes.c_int]
        L.orc_ref_prolong_fc_internal.argtypes = [G, V, V, V, C, IT, ctypes.c_int]
        L.orc_ref_flux_from_finer.argtypes = [G, ctypes.c_int, V, V, V, C, V, V, V, IT, ctypes.c_int]
        L.orc_ref_emf_to_coarser.argtypes = [G, V, V, V, C, V, ctypes.c_int64, IT, ctypes.c_int]
        L.orc_ref_prim_to_cons.argtypes = [G, V, V, V, V, V, V, ctypes.c_double, ctypes.c_int, IT, ctypes.c_int]
        IP = ctypes.POINTER(ctypes.c_int)
        L.orc_ref_prolong_ranges.argtypes = [G, IP, IP, IP, IP]
        L.orc_ref_cons_to_prim.argtypes = [G, V, V, V, V, V, V, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int, IT, ctypes.c_int]
        L.orc_ref_pack_fc.argtypes = [G, V, V, V, ctypes.c_int, V, ctypes.c_int64, IT, ctypes.c_int]
        L.orc_ref_unpack_fc.argtypes = [G, V, V, V, ctypes.c_int, V, ctypes.c_int64, IT, ctypes.c_int]
        L.orc_ref_copy_cc.argtypes = [G, ctypes.c_int, V, ctypes.c_int, V, ctypes.c_int, IT, ctypes.c_int]
        L.orc_ref_emf_same_level.argtypes = [G, V, V, V, V, ctypes.c_int64, IT, ctypes.c_int]
        L.orc_ref_emf_set.argtypes = [G, V, V, V, V, ctypes.c_int64, ctypes.c_int, IT, ctypes.c_int]
        L.orc_ref_emf_clear.argtypes = [G, V, V, V, IT, ctypes.c_int]
        L.orc_ref_emf_average.argtypes = [G, V, V, V, IT, ctypes.c_int]
        L.orc_ref_count_fine_edges.argtypes = [IP, IP, IP]
        L._ref_bound = True
    return L


def ref_restrict_cc(g, nvar, fine, ftab, coarse, items):
    a, n = _items(items)
    _ref_lib().orc_ref_restrict_cc(ctypes.byref(g), nvar, _ptr(fine), ref_coords(ftab), _ptr(coarse), a, n)


def ref_restrict_fc(g, d, fine, ftab, coarse, items):
    a, n = _items(items)
    _ref_lib().orc_ref_restrict_fc(ctypes.byref(g), d, _ptr(fine), ref_coords(ftab), _ptr(coarse), a, n)


def ref_prolong_cc(g, nvar, coarse, ctab, fine, ftab, items):
    a, n = _items(items)
    _ref_lib().orc_ref_prolong_cc(ctypes.byref(g), nvar, _ptr(coarse), ref_coords(ctab), _ptr(fine),
                                  ref_coords(ftab), a, n)


def ref_prolong_fc_shared(g, d, coarse, ctab, fine, ftab, items):
    a, n = _items(items)
    _ref_lib().orc_ref_prolong_fc_shared(ctypes.byref(g), d, _ptr(coarse), ref_coords(ctab), _ptr(fine),
                                         ref_coords(ftab), a, n)


def ref_prolong_fc_internal(g, b1, b2, b3, ftab, items):
    a, n = _items(items)
    _ref_lib().orc_ref_prolong_fc_internal(ctypes.byref(g), _ptr(b1), _ptr(b2), _ptr(b3), ref_coords(ftab), a, n)


def ref_flux_from_finer(g, nvar, f1, f2, f3, ftab, c1, c2, c3, items):
    a, n = _items(items)
    _ref_lib().orc_ref_flux_from_finer(ctypes.byref(g), nvar, _ptr(f1), _ptr(f2), _ptr(f3), ref_coords(ftab),
                                       _ptr(c1), _ptr(c2), _ptr(c3), a, n)


def ref_emf_to_coarser(g, e1, e2, e3, ftab, buf, stride, items):
    a, n = _items(items)
    _ref_lib().orc_ref_emf_to_coarser(ctypes.byref(g), _ptr(e1), _ptr(e2), _ptr(e3), ref_coords(ftab), _ptr(buf),
                                      stride, a, n)


def ref_prim_to_cons(g, w, b1, b2, b3, bcc, u, gamma, iso, items):
    a, n = _items(items)
    _ref_lib().orc_ref_prim_to_cons(ctypes.byref(g), _ptr(w), _ptr(b1), _ptr(b2), _ptr(b3), _ptr(bcc), _ptr(u),
                                    gamma, int(iso), a, n)


def ref_prolong_ranges(g, ox, lx_odd, nblevel):
    """nblevel: 27 levels relative to the block, index ((ox3+1)*3 + ox2+1)*3 + ox1+1."""
    I3, I27, I12 = ctypes.c_int * 3, ctypes.c_int * 27, ctypes.c_int * 12
    out = I12()
    rc = _ref_lib().orc_ref_prolong_ranges(ctypes.byref(g), I3(*ox), I3(*lx_odd), I27(*nblevel), out)
    assert rc == 0
    return list(out)


def ref_emf_same_level(g, e1, e2, e3, buf, stride, items):
    a, n = _items(items)
    _ref_lib().orc_ref_emf_same_level(ctypes.byref(g), _ptr(e1), _ptr(e2), _ptr(e3), _ptr(buf), stride, a, n)


def ref_emf_set(g, e1, e2, e3, buf, stride, finer, items):
    a, n = _items(items)
    _ref_lib().orc_ref_emf_set(ctypes.byref(g), _ptr(e1), _ptr(e2), _ptr(e3), _ptr(buf), stride, int(finer), a, n)


def ref_emf_clear(g, e1, e2, e3, items):
    a, n = _items(items)
    _ref_lib().orc_ref_emf_clear(ctypes.byref(g), _ptr(e1), _ptr(e2), _ptr(e3), a, n)


def ref_emf_average(g, e1, e2, e3, items):
    a, n = _items(items)
    _ref_lib().orc_ref_emf_average(ctypes.byref(g), _ptr(e1), _ptr(e2), _ptr(e3), a, n)


def ref_count_fine_edges(nblevel):
    """CountFineEdges: (edge_flag[12], nedge_fine[12]) from 27 relative levels (self = 0)."""
    I27, I12 = ctypes.c_int * 27, ctypes.c_int * 12
    fl, nf = I12(), I12()
    _ref_lib().orc_ref_count_fine_edges(I27(*nblevel), fl, nf)
    return [bool(x) for x in fl], list(nf)


def ref_copy_cc(g, nvar, src, src_coarse, dst, dst_coarse, items):
    a, n = _items(items)
    _ref_lib().orc_ref_copy_cc(ctypes.byref(g), nvar, _ptr(src), int(src_coarse), _ptr(dst), int(dst_coarse), a, n)


def ref_pack_fc(g, f, coarse, buf, stride, items):
    a, n = _items(items)
    _ref_lib().orc_ref_pack_fc(ctypes.byref(g), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), int(coarse), _ptr(buf), stride, a, n)


def ref_unpack_fc(g, f, coarse, buf, stride, items):
    a, n = _items(items)
    _ref_lib().orc_ref_unpack_fc(ctypes.byref(g), _ptr(f[0]), _ptr(f[1]), _ptr(f[2]), int(coarse), _ptr(buf), stride,
                                 a, n)


def ref_cons_to_prim(g, u, b, w, bcc, gamma, dfloor, pfloor, iso, items):
    a, n = _items(items)
    _ref_lib().orc_ref_cons_to_prim(ctypes.byref(g), _ptr(u), _ptr(b[0]), _ptr(b[1]), _ptr(b[2]), _ptr(w), _ptr(bcc),
                                    gamma, dfloor, pfloor, int(iso), a, n)
