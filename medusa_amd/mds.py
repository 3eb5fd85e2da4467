"""ctypes binding of the C-ABI in include/mds_hydro.h (libmds.so).

This is plumbing for tests and the benchmark: the hot path runs entirely in
the HIP kernels behind the C-ABI.  There is no CPU fallback: if libmds.so is
missing, or no GPU is visible, the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libmds.so")

NG = 2
FP64, FP32_BULK = 0, 1
EIO = -5                       # MDS_EIO: a table file missing or short
ERANGE = -6                    # MDS_ERANGE: a size beyond an entry point's fixed maximum
MAX_NOALLOC_FACES = 32768      # MDS_NOALLOC_MAX_FACES
BC_OUTFLOW, BC_REFLECT, BC_EXTRAP, BC_PERIODIC, BC_PROB = 0, 1, 2, 3, 4
BC_DISK = 8
MONOPOLE_NEWTON, MONOPOLE_GR = 1, 2
GEOM_CARTESIAN, GEOM_SPHERICAL = 0, 1
RCOORD_UNIFORM, RCOORD_LOG, RCOORD_SINH = 0, 1, 2
RC_COUNT = 8
TILE_ROWS2D = 0
TILE_TY3D = 1
TILE_KC3D = 2
TILE_FUSE1D = 3
TILE_GRID1D = 4
TILE_HALO3D = 5
TILE_UPD3S = 6
ZG_COUNT = 24
ZG_VOLUME, ZG_SCALE0 = 0, 12            # 1D spherical record fields (MDS_ZG_*)
# 3D spherical per-(i,j) record (MDS_ZS_*)
(ZS_VOLUME, ZS_AREA0, ZS_AREA1, ZS_AREA2, ZS_GCOV0) = range(5)
ZS_GCOVF0, ZS_GCOVF1, ZS_GCON, ZS_SCALE0 = 7, 10, 13, 16
ZS_SCALEF0, ZS_SCALEF1 = 19, 20
(ZS_C000, ZS_C011, ZS_C101, ZS_C110, ZS_C202, ZS_C220, ZS_C212, ZS_C221, ZS_C022, ZS_C122,
 ZS_C111, ZS_STH, ZS_CTH, ZS_COUNT) = range(21, 35)
BC_SPHERICAL_ORIGIN = 5

_dp = ctypes.POINTER(ctypes.c_double)


class Grid(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int),
                ("n", ctypes.c_int * 3),
                ("geom", ctypes.c_int),
                ("rcoord", ctypes.c_int),
                ("startx", ctypes.c_double * 3),
                ("dx", ctypes.c_double * 3),
                ("alpha", _dp * 3),
                ("beta", _dp * 3),
                ("Gamma", _dp * 3),
                ("rcoef", _dp * 3),
                ("volume", ctypes.c_double),
                ("area", ctypes.c_double * 3),
                ("zone", _dp),
                ("rtrans", ctypes.c_double),
                ("dendritic", ctypes.c_int),
                ("dj", ctypes.POINTER(ctypes.c_int)),
                ("dk", ctypes.POINTER(ctypes.c_int)),
                ("maxjlev", ctypes.c_int),
                ("maxklev", ctypes.c_int),
                ("lalpha", _dp * 2),
                ("lbeta", _dp * 2),
                ("lGamma", _dp * 2),
                ("lrcoef", _dp * 2)]


class EosTableDesc(ctypes.Structure):
    _fields_ = [("nr", ctypes.c_int), ("nt", ctypes.c_int), ("ny", ctypes.c_int),
                ("lr0", ctypes.c_double), ("lr1", ctypes.c_double),
                ("lt0", ctypes.c_double), ("lt1", ctypes.c_double),
                ("ye0", ctypes.c_double), ("ye1", ctypes.c_double),
                ("data", ctypes.POINTER(ctypes.c_float)),
                ("e_scale", ctypes.c_double), ("e_shift", ctypes.c_double), ("p_scale", ctypes.c_double)]


class Problem(ctypes.Structure):
    _fields_ = [("grid", ctypes.POINTER(Grid)),
                ("bc_lo", ctypes.c_int * 3),
                ("bc_hi", ctypes.c_int * 3),
                ("gam", ctypes.c_double),
                ("g0", ctypes.c_double * 3),
                ("nvel", ctypes.c_int),
                ("eos", ctypes.POINTER(EosTableDesc)),
                ("monopole", ctypes.c_int),
                ("multipole_lmax", ctypes.c_int)]


class RunParams(ctypes.Structure):
    _fields_ = [("cfl", ctypes.c_double),
                ("dt_init_safety", ctypes.c_double),
                ("dt_grow_safety", ctypes.c_double),
                ("dtmin", ctypes.c_double),
                ("dtmax", ctypes.c_double),
                ("initial_dt", ctypes.c_double),
                ("tmax", ctypes.c_double),
                ("hydro_interp_order", ctypes.c_int)]


class Health(ctypes.Structure):
    """mds_health: device health of one problem (include/mds_hydro.h)."""
    _fields_ = [("t", ctypes.c_double), ("dt", ctypes.c_double), ("istep", ctypes.c_int),
                ("dt_below_min", ctypes.c_int), ("bad_state", ctypes.c_longlong),
                ("eos_off_table", ctypes.c_longlong)]


_LIB = None

EXPORTED = ["mds_grid_build", "mds_grid_build_batch", "mds_grid_free", "mds_default_run_params", "mds_create", "mds_destroy",
            "mds_last_error", "mds_upload_prims", "mds_download_prims", "mds_update_eos", "mds_run",
            "mds_run_graph", "mds_sync", "mds_get_clock", "mds_get_health", "mds_get_gr_state", "mds_conserved_totals", "mds_counters",
            "mds_stream", "mds_recon_points", "mds_riemann_points", "mds_riemann_points_noalloc", "mds_run_timed", "mds_copy_probe",
            "mds_valu_probe", "mds_set_tile", "mds_get_tile", "mds_autotune", "mds_set_prob_bounds", "mds_set_prob_bounds_t",
            "mds_splitmix64", "mds_uniform", "mds_init_sod", "mds_init_kh", "mds_init_rt",
            "mds_init_sph_riemann", "mds_ensemble_params",
            "mds_eos_table_default", "mds_eos_model_params", "mds_eos_model_eval", "mds_eos_table_build",
            "mds_eos_table_free", "mds_eos_points", "mds_init_polytrope", "mds_eos_burrows_default",
            "mds_eos_table_load_burrows", "mds_eos_rtx"]


def lib():
    """Load libmds.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m medusa_amd.build` "
                           "(there is no CPU fallback for the device path)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    L.mds_grid_build.argtypes = [ctypes.POINTER(Grid)]
    L.mds_grid_build_batch.argtypes = [ctypes.POINTER(Grid), ctypes.c_int]
    L.mds_grid_free.argtypes = [ctypes.POINTER(Grid)]
    L.mds_grid_free.restype = None
    L.mds_default_run_params.argtypes = [ctypes.POINTER(RunParams)]
    L.mds_default_run_params.restype = None
    L.mds_create.argtypes = [ctypes.POINTER(Problem), ctypes.c_int, ctypes.POINTER(RunParams), ctypes.c_int,
                             ctypes.POINTER(vp)]
    L.mds_destroy.argtypes = [vp]
    L.mds_last_error.argtypes = [vp]
    L.mds_last_error.restype = ctypes.c_char_p
    L.mds_upload_prims.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
    L.mds_download_prims.argtypes = [vp, ctypes.c_int, _dp, ctypes.c_int]
    L.mds_update_eos.argtypes = [vp]
    L.mds_run.argtypes = [vp, ctypes.c_int]
    L.mds_run_graph.argtypes = [vp, ctypes.c_int]
    L.mds_sync.argtypes = [vp]
    L.mds_get_clock.argtypes = [vp, ctypes.c_int, _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    L.mds_get_health.argtypes = [vp, ctypes.c_int, ctypes.POINTER(Health)]
    L.mds_get_gr_state.argtypes = [vp, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_int)]
    L.mds_conserved_totals.argtypes = [vp, ctypes.c_int, _dp]
    L.mds_counters.argtypes = [vp, ctypes.POINTER(ctypes.c_longlong)]
    L.mds_stream.argtypes = [vp]
    L.mds_stream.restype = vp
    L.mds_run_timed.argtypes = [vp, ctypes.c_int, _dp]
    L.mds_copy_probe.argtypes = [ctypes.c_longlong, ctypes.c_int, _dp]
    L.mds_valu_probe.argtypes = [ctypes.c_int, ctypes.c_int, _dp]
    L.mds_set_tile.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.mds_get_tile.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.mds_autotune.argtypes = [vp, ctypes.c_int, _dp]
    L.mds_set_prob_bounds.argtypes = [vp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.mds_set_prob_bounds_t.argtypes = [vp, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    L.mds_recon_points.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.mds_riemann_points.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, ctypes.c_int,
                                     _dp, _dp, ctypes.POINTER(ctypes.c_int)]
    L.mds_riemann_points_noalloc.argtypes = L.mds_riemann_points.argtypes
    L.mds_eos_table_default.argtypes = [ctypes.POINTER(EosTableDesc)]
    L.mds_eos_table_default.restype = None
    L.mds_eos_model_params.argtypes = [ctypes.c_ulonglong, _dp]
    L.mds_eos_model_params.restype = None
    L.mds_eos_model_eval.argtypes = [_dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    L.mds_eos_model_eval.restype = None
    L.mds_eos_table_build.argtypes = [ctypes.POINTER(EosTableDesc), ctypes.c_ulonglong]
    L.mds_eos_burrows_default.argtypes = [ctypes.POINTER(EosTableDesc)]
    L.mds_eos_burrows_default.restype = None
    L.mds_eos_table_load_burrows.argtypes = [ctypes.c_char_p, ctypes.POINTER(EosTableDesc)]
    L.mds_eos_table_free.argtypes = [ctypes.POINTER(EosTableDesc)]
    L.mds_eos_table_free.restype = None
    L.mds_eos_points.argtypes = [ctypes.c_int, ctypes.POINTER(EosTableDesc), ctypes.c_int, _dp, _dp, _dp,
                                 _dp, _dp, _dp, _dp]
    L.mds_eos_rtx.argtypes = [ctypes.POINTER(EosTableDesc), ctypes.c_double, ctypes.c_double, ctypes.c_double, _dp]
    _LIB = L
    return L


def _ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _grid_desc(ndim, n, startx, dx, geom, rcoord, g=None, rtrans=0.0, dendritic=False):
    g = g if g is not None else Grid()
    g.rtrans = float(rtrans)
    g.dendritic = int(bool(dendritic))
    g.ndim = ndim
    n = list(n) + [1] * (3 - len(n))
    for d in range(3):
        g.n[d] = int(n[d])
        g.startx[d] = float(startx[d]) if d < len(startx) else 0.0
        g.dx[d] = float(dx[d]) if d < len(dx) else 1.0
    g.geom = geom
    g.rcoord = rcoord
    return g


class GridBuilder:
    """Owns an mds_grid built by mds_grid_build (the reference geometry)."""

    def __init__(self, ndim, n, startx, dx, geom=GEOM_CARTESIAN, rcoord=RCOORD_UNIFORM, _built=None,
                 rtrans=0.0, dendritic=False):
        if _built is not None:
            self.g = _built
        else:
            self.g = _grid_desc(ndim, n, startx, dx, geom, rcoord, rtrans=rtrans, dendritic=dendritic)
            rc = lib().mds_grid_build(ctypes.byref(self.g))
            if rc:
                raise RuntimeError(f"mds_grid_build failed ({rc})")
        self.ndim = self.g.ndim
        self.n = [self.g.n[d] for d in range(3)]

    @staticmethod
    def batch(descs):
        """Build many grids with host threads: descs = [(ndim, n, startx, dx, geom, rcoord), ...]."""
        arr = (Grid * len(descs))()
        for k, d in enumerate(descs):
            _grid_desc(*d, g=arr[k])
        rc = lib().mds_grid_build_batch(arr, len(descs))
        if rc:
            raise RuntimeError(f"mds_grid_build_batch failed ({rc})")
        out = []
        for k in range(len(descs)):
            g = Grid()
            ctypes.pointer(g)[0] = arr[k]
            out.append(GridBuilder(0, [], [], [], _built=g))
        return out

    def __del__(self):
        if getattr(self, "g", None) is not None and _LIB is not None:
            _LIB.mds_grid_free(ctypes.byref(self.g))
            self.g = None

    def moments(self, d):
        nz = self.n[d] + 2 * NG
        return [np.ctypeslib.as_array(getattr(self.g, k)[d], shape=(nz,)).copy() for k in ("alpha", "beta", "Gamma")]

    def rcoef(self, d):
        nz = self.n[d] + 2 * NG
        return np.ctypeslib.as_array(self.g.rcoef[d], shape=(nz * RC_COUNT,)).reshape(nz, RC_COUNT).copy()

    def zone(self):
        """1D spherical: [n1+2NG][MDS_ZG_COUNT]; 3D spherical: [n2+2NG][n1+2NG][MDS_ZS_COUNT]."""
        if not self.g.zone:
            return None
        nz = self.n[0] + 2 * NG
        if self.ndim >= 2 and self.g.geom == GEOM_SPHERICAL:
            nj = self.n[1] + 2 * NG
            return np.ctypeslib.as_array(self.g.zone, shape=(nj * nz * ZS_COUNT,)).reshape(nj, nz, ZS_COUNT).copy()
        return np.ctypeslib.as_array(self.g.zone, shape=(nz * ZG_COUNT,)).reshape(nz, ZG_COUNT).copy()

    def levels(self):
        """Dendritic grids: (nj[n1], nk[n1, n2] with 0 beyond nj[i]); None on uniform grids."""
        if not self.g.dj:
            return None
        n1, n2, n3 = self.n
        dj = np.ctypeslib.as_array(self.g.dj, shape=(n1 + 2 * NG,))[NG:NG + n1]
        dk = np.ctypeslib.as_array(self.g.dk, shape=((n1 + 2 * NG) * (n2 + 2 * NG),)).reshape(n1 + 2 * NG, n2 + 2 * NG)
        nj = (n2 // dj).astype(np.int32)
        nk = np.zeros((n1, n2), dtype=np.int32)
        for i in range(n1):
            nk[i, :nj[i]] = n3 // dk[i + NG, NG:NG + nj[i]]
        return nj, nk

    def valid_mask(self):
        """Boolean [n3, n2, n1]: which (i, j, k) positions are zones (all on uniform grids)."""
        n1, n2, n3 = self.n
        lv = self.levels()
        if lv is None:
            return np.ones((n3, n2, n1), bool)
        nj, nk = lv
        m = np.zeros((n3, n2, n1), bool)
        for i in range(n1):
            for j in range(nj[i]):
                m[:nk[i, j], j, i] = True
        return m

    def level_moments(self, d, lev):
        """theta (d=1) or phi (d=2) [alpha, beta, Gamma] of refinement level lev (length n + 2 NG)."""
        t = d - 1
        maxl = self.g.maxjlev if t == 0 else self.g.maxklev
        if not self.g.lalpha[t] or lev > maxl:
            return None
        nn = self.n[d] + 2 * NG
        return [np.ctypeslib.as_array(getattr(self.g, f)[t], shape=((maxl + 1) * nn,))[lev * nn:(lev + 1) * nn].copy()
                for f in ("lalpha", "lbeta", "lGamma")]


def run_params(cfl=0.5, tmax=1.0e300, **kw):
    rp = RunParams()
    lib().mds_default_run_params(ctypes.byref(rp))
    rp.cfl = cfl
    rp.tmax = tmax
    for k, v in kw.items():
        setattr(rp, k, v)
    return rp


# mds_prob_bounds_fn: (i, j, k, p, user)
PROB_BOUNDS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                  ctypes.c_void_p)
PROB_BOUNDS_T_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


class Context:
    """A device context holding one or more problems (mds_ctx)."""

    def __init__(self, problems, precision=FP64, rp=None):
        """problems: list of dicts {grid: GridBuilder, bc_lo, bc_hi, gam, g0, nvel}."""
        L = lib()
        self.grids = [p["grid"] for p in problems]
        arr = (Problem * len(problems))()
        for k, p in enumerate(problems):
            arr[k].grid = ctypes.pointer(p["grid"].g)
            for d in range(3):
                arr[k].bc_lo[d] = p.get("bc_lo", [0, 0, 0])[d] if d < len(p.get("bc_lo", [])) else 0
                arr[k].bc_hi[d] = p.get("bc_hi", [0, 0, 0])[d] if d < len(p.get("bc_hi", [])) else 0
                arr[k].g0[d] = p.get("g0", [0.0, 0.0, 0.0])[d]
            arr[k].gam = p["gam"]
            arr[k].nvel = p.get("nvel", 0)
            if p.get("eos") is not None:
                arr[k].eos = ctypes.pointer(p["eos"].t)
            # True / 1: Newtonian, 2 (MONOPOLE_GR): the GR effective monopole
            arr[k].monopole = int(p.get("monopole", 0))
            arr[k].multipole_lmax = int(p.get("multipole_lmax", 0))
        self._tables = [p.get("eos") for p in problems]     # keep the tables alive
        self.nstate = 6 if problems[0].get("eos") is not None else 5
        self.rp = rp if rp is not None else run_params()
        h = ctypes.c_void_p()
        rc = L.mds_create(arr, len(problems), ctypes.byref(self.rp), precision, ctypes.byref(h))
        if rc:
            raise RuntimeError(f"mds_create failed ({rc})")
        self.h = h
        self.precision = precision
        self.nprob = len(problems)
        self.ndim = problems[0]["grid"].g.ndim

    def __del__(self):
        if getattr(self, "h", None) and _LIB is not None:
            _LIB.mds_destroy(self.h)
            self.h = None

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {lib().mds_last_error(self.h).decode()}")

    def set_prob_bounds(self, fn, prob=0):
        """PROB ghost values: fn(i, j, k) -> (rho, u, v1, v2, v3[, Ye]) for every
        ghost zone of the problem's PROB sides (the reference's prob_bounds)."""
        def _cb(i, j, k, p, user):
            vals = fn(i, j, k)
            for v in range(min(len(vals), 6)):
                p[v] = float(vals[v])
        cb = PROB_BOUNDS_FN(_cb)
        self._check(lib().mds_set_prob_bounds(self.h, prob, ctypes.cast(cb, ctypes.c_void_p), None),
                    "set_prob_bounds")

    def set_prob_bounds_t(self, fn, prob=0):
        """Time-dependent PROB ghost values: fn(i, j, k, t) -> (rho, u, v1, v2,
        v3[, Ye]) at the time the reference's fill sees (t in the first RK
        stage, t + dt in the second); the context then steps synchronously."""
        def _cb(i, j, k, t, p, user):
            vals = fn(i, j, k, t)
            for v in range(min(len(vals), 6)):
                p[v] = float(vals[v])
        cb = PROB_BOUNDS_T_FN(_cb)
        if not hasattr(self, "_tdep_keep"):
            self._tdep_keep = {}
        self._tdep_keep[prob] = cb                     # the library calls it every step
        self._check(lib().mds_set_prob_bounds_t(self.h, prob, ctypes.cast(cb, ctypes.c_void_p), None),
                    "set_prob_bounds_t")

    def _shape(self, prob):
        g = self.grids[prob]
        return (self.nstate, g.n[2], g.n[1], g.n[0])

    def upload(self, prims, prob=0):
        a = np.ascontiguousarray(np.asarray(prims, dtype=np.float64)[:self.nstate]).reshape(self._shape(prob))
        self._check(lib().mds_upload_prims(self.h, prob, _ptr(a), self.nstate), "upload")

    def download(self, prob=0):
        a = np.zeros(self._shape(prob), dtype=np.float64)
        self._check(lib().mds_download_prims(self.h, prob, _ptr(a), self.nstate), "download")
        return a

    def update_eos(self):
        self._check(lib().mds_update_eos(self.h), "update_eos")

    def run(self, nsteps, graph=False):
        fn = lib().mds_run_graph if graph else lib().mds_run
        self._check(fn(self.h, nsteps), "run")

    def run_timed(self, nsteps):
        """Eager run with HIP events between kernels: (total_ms, stage1_kernels_ms, stage2_kernels_ms)."""
        out = np.zeros(3)
        self._check(lib().mds_run_timed(self.h, nsteps, _ptr(out)), "run_timed")
        return tuple(out)

    def set_tile(self, param, value):
        self._check(lib().mds_set_tile(self.h, param, value), "set_tile")

    def get_tile(self, param):
        v = ctypes.c_int()
        self._check(lib().mds_get_tile(self.h, param, ctypes.byref(v)), "get_tile")
        return v.value

    def autotune(self, nsteps=3):
        """Pick the fastest launch shape on the current state (state is restored)."""
        ms = np.zeros(1)
        self._check(lib().mds_autotune(self.h, nsteps, _ptr(ms)), "autotune")
        return float(ms[0])

    def sync(self):
        self._check(lib().mds_sync(self.h), "sync")

    def clock(self, prob=0):
        t, dt, ist = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        self._check(lib().mds_get_clock(self.h, prob, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(ist)),
                    "get_clock")
        return t.value, dt.value, ist.value

    def health(self, prob=0):
        """Device health counters of a problem: dict of bad_state, eos_off_table, dt_below_min (and the clock)."""
        h = Health()
        self._check(lib().mds_get_health(self.h, prob, ctypes.byref(h)), "get_health")
        return {f: getattr(h, f) for f, _ in Health._fields_}

    def gr_state(self, prob=0):
        """GR effective monopole state of a problem (mds_get_gr_state): dict of
        gtov, mr, lapse_edge (n1 + 1 faces), lapse (n1 zones), passes, nonconverged."""
        n1 = self.grids[prob].n[0]
        out = {k: np.zeros(n1 + 1) for k in ("gtov", "mr", "lapse_edge")}
        out["lapse"] = np.zeros(n1)
        ps, nc = ctypes.c_int(), ctypes.c_int()
        self._check(lib().mds_get_gr_state(self.h, prob, _ptr(out["gtov"]), _ptr(out["mr"]), _ptr(out["lapse_edge"]),
                                           _ptr(out["lapse"]), ctypes.byref(ps), ctypes.byref(nc)), "get_gr_state")
        out["passes"], out["nonconverged"] = ps.value, nc.value
        return out

    def totals(self, prob=0):
        out = np.zeros(2)
        self._check(lib().mds_conserved_totals(self.h, prob, _ptr(out)), "totals")
        return out

    def counters(self):
        out = (ctypes.c_longlong * 2)()
        self._check(lib().mds_counters(self.h, out), "counters")
        return [out[0], out[1]]

    def run_to(self, tmax_steps_cap=10 ** 7, batch=64):
        """Run until every problem is frozen (t >= tmax), checking the clock every `batch` steps."""
        done = 0
        while done < tmax_steps_cap:
            self.run(batch)
            done += batch
            if all(self.clock(p)[0] >= self.rp.tmax for p in range(self.nprob)):
                break
        return done


def copy_probe_ms(nbytes, reps=20):
    """Mean duration (ms) of a device copy kernel over nbytes."""
    out = np.zeros(1)
    if lib().mds_copy_probe(int(nbytes), reps, _ptr(out)):
        raise RuntimeError("mds_copy_probe failed")
    return float(out[0])


VALU_PROBE_KINDS = {"fma32": 0, "fma64": 1, "rcp32": 2, "rcp64": 3, "sqrt32": 4, "sqrt64": 5}


def valu_probe_rate(fp64=False, reps=5, kind=None):
    """Wavefront instructions per second of the device: FMA (FP32 or FP64), or
    the chain named by kind (a VALU_PROBE_KINDS key)."""
    out = np.zeros(1)
    k = int(bool(fp64)) if kind is None else VALU_PROBE_KINDS[kind]
    if lib().mds_valu_probe(k, reps, _ptr(out)):
        raise RuntimeError("mds_valu_probe failed")
    return float(out[0])


def recon_points(precision, a, mom=None, x=None, rcoef=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    al = np.zeros(n)
    ar = np.zeros(n)
    m = _ptr(np.ascontiguousarray(mom, dtype=np.float64)) if mom is not None else None
    xx = _ptr(np.ascontiguousarray(x, dtype=np.float64)) if x is not None else None
    rc = _ptr(np.ascontiguousarray(rcoef, dtype=np.float64)) if rcoef is not None else None
    r = lib().mds_recon_points(precision, n, _ptr(a), m, xx, rc, _ptr(al), _ptr(ar))
    if r:
        raise RuntimeError(f"mds_recon_points failed ({r})")
    return al, ar


def riemann_points(precision, solver, pl, pr, fg, direction, noalloc=False):
    """Pointwise device Riemann fluxes (mds_riemann_points, or with noalloc
    its allocation-free variant, at most MAX_NOALLOC_FACES faces)."""
    pl = np.ascontiguousarray(pl, dtype=np.float64)
    pr = np.ascontiguousarray(pr, dtype=np.float64)
    fg = np.ascontiguousarray(fg, dtype=np.float64)
    n = pl.shape[0]
    flux = np.zeros((n, 5))
    sig = np.zeros(n)
    sh = np.zeros(n, dtype=np.int32)
    fn = lib().mds_riemann_points_noalloc if noalloc else lib().mds_riemann_points
    r = fn(precision, solver, n, _ptr(pl), _ptr(pr), _ptr(fg), direction, _ptr(flux),
           _ptr(sig), sh.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    if r:
        raise RuntimeError(f"{fn.__name__} failed ({r})")
    return flux, sig, sh


# ---------------------------------------------------------------------------
# host problem initialisers (include/mds_host.h)
# ---------------------------------------------------------------------------
def _host():
    L = lib()
    if not hasattr(L, "_host_ready"):
        L.mds_init_sod.argtypes = [ctypes.POINTER(Grid), ctypes.c_double, _dp]
        L.mds_init_kh.argtypes = [ctypes.POINTER(Grid), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_ulonglong, _dp]
        L.mds_init_rt.argtypes = [ctypes.POINTER(Grid)] + [ctypes.c_double] * 7 + [ctypes.c_ulonglong, _dp]
        L.mds_init_sph_riemann.argtypes = [ctypes.POINTER(Grid)] + [ctypes.c_double] * 6 + [_dp]
        L.mds_ensemble_params.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_ulonglong, _dp]
        L.mds_init_polytrope.argtypes = ([ctypes.POINTER(Grid)] + [ctypes.c_double] * 5
                                         + [ctypes.POINTER(EosTableDesc), ctypes.c_ulonglong, ctypes.c_double,
                                            ctypes.c_ulonglong, _dp])
        L.mds_uniform.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong]
        L.mds_uniform.restype = ctypes.c_double
        L._host_ready = True
    return L


def _state(grid):
    return np.zeros((5, grid.n[2], grid.n[1], grid.n[0]))


def init_sod(grid, gam=1.4):
    p = _state(grid)
    if _host().mds_init_sod(ctypes.byref(grid.g), gam, _ptr(p)):
        raise RuntimeError("mds_init_sod failed")
    return p


def init_kh(grid, gam=5.0 / 3.0, p0=2.5, amp=0.01, noise=1.0e-3, seed=0x5EED0002):
    p = _state(grid)
    if _host().mds_init_kh(ctypes.byref(grid.g), gam, p0, amp, noise, seed, _ptr(p)):
        raise RuntimeError("mds_init_kh failed")
    return p


def init_rt(grid, gam=1.4, p0=2.5, gz=-0.1, rho_lo=1.0, rho_hi=2.0, amp=0.01, noise=1.0e-4, seed=0x5EED0003):
    p = _state(grid)
    if _host().mds_init_rt(ctypes.byref(grid.g), gam, p0, gz, rho_lo, rho_hi, amp, noise, seed, _ptr(p)):
        raise RuntimeError("mds_init_rt failed")
    return p


def init_sph_riemann(grid, gam, rho_in, p_in, r0, frho, fp):
    p = _state(grid)
    if _host().mds_init_sph_riemann(ctypes.byref(grid.g), gam, rho_in, p_in, r0, frho, fp, _ptr(p)):
        raise RuntimeError("mds_init_sph_riemann failed")
    return p


def init_polytrope(grid, n_index=3.0, rho_c=1.0e10, K=6.0e14, rho_atm=1.0e4, gam=4.0 / 3.0, table=None,
                   sigma=0.01, seed=0x5EED0006):
    """Polytropic star: prims [6][n3][n2][n1] (rho, u, v1, v2, v3, Ye)."""
    p = np.zeros((6, grid.n[2], grid.n[1], grid.n[0]))
    tp = ctypes.byref(table.t) if table is not None else None
    es = table.seed if table is not None else 0
    if _host().mds_init_polytrope(ctypes.byref(grid.g), n_index, rho_c, K, rho_atm, gam, tp, es, sigma, seed,
                                  _ptr(p)):
        raise RuntimeError("mds_init_polytrope failed")
    return p


def ensemble_params(nprob, nmin=256, nmax=8192, seed=0x5EED0004):
    out = np.zeros((nprob, 7))
    if _host().mds_ensemble_params(nprob, nmin, nmax, seed, _ptr(out)):
        raise RuntimeError("mds_ensemble_params failed")
    return out


# --------------------------------------------------------------------------
# tabulated EOS
# --------------------------------------------------------------------------
EOS_SEED = 0x5EED0005


class EosTable:
    """Synthetic tabulated EOS (mds_eos_table): host FP32 records
    [MDS_EOS_NFIELD][ny][nr][nt] (e, ln P, Gamma_1)."""

    NFIELD = 3

    def __init__(self, seed=EOS_SEED, burrows_path=None, **dims):
        """The synthetic table (seeded), or with burrows_path a table file in the
        reference's Burrows binary layout (dimensions from mds_eos_burrows_default
        unless given in dims)."""
        self.t = EosTableDesc()
        if burrows_path is None:
            lib().mds_eos_table_default(ctypes.byref(self.t))
        else:
            lib().mds_eos_burrows_default(ctypes.byref(self.t))
        for k, v in dims.items():
            setattr(self.t, k, v)
        self.seed = seed
        if burrows_path is None:
            rc = lib().mds_eos_table_build(ctypes.byref(self.t), seed)
        else:
            rc = lib().mds_eos_table_load_burrows(os.fsencode(burrows_path), ctypes.byref(self.t))
        if rc:
            raise RuntimeError(f"EOS table set-up failed ({rc})")
        n = self.NFIELD * self.t.ny * self.t.nr * self.t.nt
        self.data = np.ctypeslib.as_array(self.t.data, shape=(n,)).reshape(self.NFIELD, self.t.ny, self.t.nr,
                                                                           self.t.nt)

    def __del__(self):
        if getattr(self, "t", None) is not None and _LIB is not None and self.t.data:
            _LIB.mds_eos_table_free(ctypes.byref(self.t))

    def nodes(self):
        """(log10 rho, log10 T, Ye) node coordinates."""
        t = self.t
        return (t.lr0 + np.arange(t.nr) * ((t.lr1 - t.lr0) / (t.nr - 1)),
                t.lt0 + np.arange(t.nt) * ((t.lt1 - t.lt0) / (t.nt - 1)),
                t.ye0 + np.arange(t.ny) * ((t.ye1 - t.ye0) / (t.ny - 1)))


def eos_model_params(seed=EOS_SEED):
    out = np.zeros(3)
    lib().mds_eos_model_params(seed, _ptr(out))
    return out


def eos_model_eval(prm, rho, temp, ye):
    """(P, e, Gamma_1) of the analytic model at one point (extended precision)."""
    out = np.zeros(3)
    prm = np.ascontiguousarray(prm, dtype=np.float64)
    lib().mds_eos_model_eval(_ptr(prm), rho, temp, ye, _ptr(out))
    return out


def eos_rtx(table, rho, temp, ye):
    """Host table at a given temperature (mds_eos_rtx): (u, P, Gamma_1, cs)."""
    out = np.zeros(4)
    if lib().mds_eos_rtx(ctypes.byref(table.t), float(rho), float(temp), float(ye), _ptr(out)):
        raise RuntimeError("mds_eos_rtx failed")
    return out


def eos_points(precision, table, rho, u, ye):
    """Device EOS fill of arrays of states: (P, cs, T, Gamma_1)."""
    rho = np.ascontiguousarray(rho, dtype=np.float64)
    u = np.ascontiguousarray(u, dtype=np.float64)
    ye = np.ascontiguousarray(ye, dtype=np.float64)
    n = rho.size
    outs = [np.zeros(n) for _ in range(4)]
    rc = lib().mds_eos_points(precision, ctypes.byref(table.t), n, _ptr(rho), _ptr(u), _ptr(ye),
                              *[_ptr(o) for o in outs])
    if rc:
        raise RuntimeError(f"mds_eos_points failed ({rc})")
    return tuple(outs)
