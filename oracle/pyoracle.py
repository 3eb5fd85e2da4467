"""ctypes binding of the CPU oracle (test infrastructure only).

Imported only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
leg, always as the checker.  Builds oracle/liboracle.so on first use if it is
missing (gcc, strict IEEE) so that the CPU test suite needs no manual step.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

RHO, UU, U1, U2, U3, YE = 0, 1, 2, 3, 4, 5
BC_OUTFLOW, BC_REFLECT, BC_EXTRAP, BC_PERIODIC = 0, 1, 2, 3
GEOM_CARTESIAN, GEOM_SPHERICAL = 0, 1
RCOORD_UNIFORM, RCOORD_LOG, RCOORD_SINH = 0, 1, 2
BC_PROB = 4
BC_DISK = 8
BC_SPHERICAL_ORIGIN = 5


class ZoneGeom(ctypes.Structure):
    _fields_ = [("area", ctypes.c_double * 3),
                ("conn", ctypes.c_double * 27),
                ("gcov", ctypes.c_double * 12),
                ("gcon", ctypes.c_double * 3),
                ("scale", ctypes.c_double * 12),
                ("volume", ctypes.c_double)]


class Config(ctypes.Structure):
    _fields_ = [("ndim", ctypes.c_int),
                ("n", ctypes.c_int * 3),
                ("geom", ctypes.c_int),
                ("rcoord", ctypes.c_int),
                ("startx", ctypes.c_double * 3),
                ("dx", ctypes.c_double * 3),
                ("bc_lo", ctypes.c_int * 3),
                ("bc_hi", ctypes.c_int * 3),
                ("nvars", ctypes.c_int),
                ("interp_order", ctypes.c_int),
                ("gam", ctypes.c_double),
                ("cfl", ctypes.c_double),
                ("dt_init_safety", ctypes.c_double),
                ("dt_grow_safety", ctypes.c_double),
                ("dtmin", ctypes.c_double),
                ("dtmax", ctypes.c_double),
                ("initial_dt", ctypes.c_double),
                ("g0", ctypes.c_double * 3),
                ("use_fixed_grav", ctypes.c_int),
                ("use_monopole", ctypes.c_int),
                ("eos_tab", ctypes.c_void_p),
                ("rtrans", ctypes.c_double),
                ("dendritic", ctypes.c_int),
                ("prob_bounds", ctypes.c_void_p),
                ("prob_user", ctypes.c_void_p),
                ("prob_bounds_t", ctypes.c_void_p),
                ("multipole_lmax", ctypes.c_int)]


# prob_bounds(i, j, k, p, user): the ghost values of a PROB boundary zone
PROB_BOUNDS_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
# prob_bounds_t(i, j, k, t, p, user): the same at the time the fill sees
PROB_BOUNDS_T_FN = ctypes.CFUNCTYPE(None, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def _load(fast=False):
    name = "liboracle_fast.so" if fast else "liboracle.so"
    path = os.path.join(HERE, name)
    srcs = [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith((".c", ".h"))]
    if (not os.path.exists(path)) or any(os.path.getmtime(s) > os.path.getmtime(path) for s in srcs):
        subprocess.check_call(["make", "-s", "-C", HERE, name])
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_create.restype = ctypes.c_void_p
    lib.orc_create.argtypes = [ctypes.POINTER(Config)]
    lib.orc_destroy.argtypes = [ctypes.c_void_p]
    lib.orc_set_prims.argtypes = [ctypes.c_void_p, dp]
    lib.orc_get_prims.argtypes = [ctypes.c_void_p, dp]
    lib.orc_run.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_int]
    lib.orc_run.restype = ctypes.c_int
    lib.orc_step.argtypes = [ctypes.c_void_p]
    lib.orc_estimate_dt.argtypes = [ctypes.c_void_p]
    lib.orc_estimate_dt.restype = ctypes.c_double
    lib.orc_update_eos_all.argtypes = [ctypes.c_void_p]
    for fn in ("orc_time", "orc_last_dt"):
        getattr(lib, fn).argtypes = [ctypes.c_void_p]
        getattr(lib, fn).restype = ctypes.c_double
    lib.orc_istep.argtypes = [ctypes.c_void_p]
    lib.orc_istep.restype = ctypes.c_int
    lib.orc_dendritic_levels.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]
    lib.orc_dendritic_levels.restype = ctypes.c_int
    lib.orc_level_moment.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_level_moment.restype = ctypes.POINTER(ctypes.c_double)
    for fn in ("orc_shock_face_count", "orc_hlle_face_count"):
        getattr(lib, fn).argtypes = [ctypes.c_void_p]
        getattr(lib, fn).restype = ctypes.c_long
    lib.orc_fill_ghosts.argtypes = [ctypes.c_void_p]
    lib.orc_zone_prims.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_zone_prims.restype = ctypes.c_int
    lib.orc_zone_count_all.argtypes = [ctypes.c_void_p]
    lib.orc_zone_count_all.restype = ctypes.c_int
    lib.orc_geometry.argtypes = [ctypes.c_void_p]
    lib.orc_geometry.restype = ctypes.POINTER(ZoneGeom)
    lib.orc_moment.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.orc_moment.restype = dp
    lib.orc_conserved_totals.argtypes = [ctypes.c_void_p, dp]
    lib.orc_para_recon.argtypes = [dp, dp, dp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
    lib.orc_lin_recon.argtypes = [dp, dp, dp, dp, dp, dp, ctypes.c_int, ctypes.c_int]
    for fn in ("orc_riemann_hllc", "orc_riemann_hlle"):
        f = getattr(lib, fn)
        f.argtypes = [ctypes.c_int, ctypes.c_int, dp, dp, ctypes.POINTER(ZoneGeom), ctypes.c_int, dp, dp, dp]
        f.restype = ctypes.c_int
    lib.orc_p_to_u.argtypes = [dp, dp, dp, ctypes.c_int]
    lib.orc_u_to_p.argtypes = [dp, dp, dp, ctypes.c_int]
    lib.orc_geom_src.argtypes = [dp, ctypes.c_double, dp, ctypes.POINTER(ZoneGeom), ctypes.c_int, ctypes.c_int]
    lib.orc_dt_source_control.argtypes = [dp, dp, dp]
    lib.orc_dt_source_control.restype = ctypes.c_double
    lib.orc_monopole_edges.argtypes = [ctypes.c_void_p, dp]
    lib.orc_monopole_edges.restype = None
    lib.orc_gr_state.argtypes = [ctypes.c_void_p, dp, dp, dp, dp, ctypes.POINTER(ctypes.c_int)]
    lib.orc_gr_state.restype = ctypes.c_int
    lib.orc_gr_edges.argtypes = [ctypes.c_void_p, dp]
    lib.orc_gr_edges.restype = ctypes.c_int
    lib.orc_ylms.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int]
    lib.orc_ylms.restype = None
    lib.orc_ylm_ints.argtypes = [dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_int, ctypes.c_int]
    lib.orc_ylm_ints.restype = None
    lib.orc_multipole_phi.argtypes = [ctypes.c_void_p, dp]
    lib.orc_multipole_phi.restype = ctypes.c_int
    lib.orc_multipole_sources.argtypes = [ctypes.c_void_p, dp]
    lib.orc_multipole_sources.restype = ctypes.c_int
    lib.orc_eos_walk.argtypes = [ctypes.POINTER(EosTab), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, dp]
    lib.orc_eos_walk.restype = ctypes.c_int
    lib.orc_eos_root.argtypes = [ctypes.POINTER(EosTab), ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                 ctypes.c_double, ctypes.c_double, dp, ctypes.POINTER(ctypes.c_int)]
    lib.orc_eos_root.restype = ctypes.c_int
    lib.orc_prepare_entries.argtypes = [ctypes.c_void_p]
    lib.orc_dend_radial_entry.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_dend_radial_entry.restype = ctypes.c_int
    lib.orc_dend_theta_entry.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp]
    lib.orc_dend_theta_entry.restype = ctypes.c_int
    lib.orc_polar_face_states.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, dp]
    lib.orc_polar_face_states.restype = ctypes.c_int
    return lib


_LIBS = {}


class EosTab(ctypes.Structure):
    """orc_eos_tab: dimensions, node ranges and the FP32 records."""
    _fields_ = [("nr", ctypes.c_int), ("nt", ctypes.c_int), ("ny", ctypes.c_int),
                ("lr0", ctypes.c_double), ("lr1", ctypes.c_double),
                ("lt0", ctypes.c_double), ("lt1", ctypes.c_double),
                ("ye0", ctypes.c_double), ("ye1", ctypes.c_double),
                ("data", ctypes.POINTER(ctypes.c_float)),
                ("e_scale", ctypes.c_double), ("e_shift", ctypes.c_double), ("p_scale", ctypes.c_double)]


def eos_tab(dims, data):
    """Describe a float32 table [3][ny][nr][nt]; dims: dict of nr, nt, ny, the ranges and
    (optionally) the units e_scale, e_shift, p_scale (default cgs: 1, 0, 1)."""
    data = np.ascontiguousarray(data, dtype=np.float32)
    t = EosTab(**{k: dims[k] for k in ("nr", "nt", "ny", "lr0", "lr1", "lt0", "lt1", "ye0", "ye1")},
               e_scale=dims.get("e_scale", 1.0), e_shift=dims.get("e_shift", 0.0), p_scale=dims.get("p_scale", 1.0))
    t.data = data.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
    t._keep = data
    return t


def eos_walk(tab, rho, u, ye, temp_guess=0.0):
    """eos_stuff.c:507-650 restated: (P, cs, T, Gamma_1), status."""
    out = np.zeros(4)
    st = lib().orc_eos_walk(ctypes.byref(tab), rho, u, ye, temp_guess, ptr(out))
    return out, st


def eos_root(tab, rho, u, ye, lt_guess, tol=1e-10):
    """eos_stellar_collapse.c:294-369 restated on the same interpolant: (log10 T, status, newton iterations)."""
    lt = np.zeros(1)
    it = ctypes.c_int()
    st = lib().orc_eos_root(ctypes.byref(tab), rho, u, ye, lt_guess, tol, ptr(lt), ctypes.byref(it))
    return float(lt[0]), st, it.value


def lib(fast=False):
    if fast not in _LIBS:
        _LIBS[fast] = _load(fast)
    return _LIBS[fast]


def ptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_config(ndim, n, startx, dx, bc_lo, bc_hi, gam, cfl=0.5, nvars=5,
                geom=GEOM_CARTESIAN, rcoord=RCOORD_UNIFORM, interp_order=3,
                g0=(0.0, 0.0, 0.0), dt_init_safety=0.1, dt_grow_safety=1.1,
                dtmin=1.0e-200, dtmax=1.0e200, initial_dt=1.0e100, monopole=False, eos_tab=None,
                rtrans=0.0, dendritic=False, prob_bounds=None, prob_bounds_t=None, multipole_lmax=0):
    """eos_tab: an EosTab (tabulated EOS; nvars must be 6 to carry Ye).
    rtrans: scale of the sinh radial map; dendritic: 3D spherical dendritic grid.
    prob_bounds: f(i, j, k) -> sequence of nvars ghost values for PROB sides;
    prob_bounds_t: f(i, j, k, t) -> the same at the time the fill sees."""
    c = Config()
    c.ndim = ndim
    n = list(n) + [1] * (3 - len(n))
    for d in range(3):
        c.n[d] = int(n[d])
        c.startx[d] = float(startx[d]) if d < len(startx) else 0.0
        c.dx[d] = float(dx[d]) if d < len(dx) else 1.0
        c.bc_lo[d] = int(bc_lo[d]) if d < len(bc_lo) else 0
        c.bc_hi[d] = int(bc_hi[d]) if d < len(bc_hi) else 0
        c.g0[d] = float(g0[d])
    c.geom = geom
    c.rcoord = rcoord
    c.nvars = nvars
    c.interp_order = interp_order
    c.gam = gam
    c.cfl = cfl
    c.dt_init_safety = dt_init_safety
    c.dt_grow_safety = dt_grow_safety
    c.dtmin = dtmin
    c.dtmax = dtmax
    c.initial_dt = initial_dt
    c.use_fixed_grav = int(any(v != 0.0 for v in g0))
    # monopole: False/0 none, True/1 Newtonian, 2 GR effective potential
    c.use_monopole = 2 if (monopole == 2 and monopole is not True) else int(bool(monopole))
    c.rtrans = float(rtrans)
    c.dendritic = int(bool(dendritic))
    c.multipole_lmax = int(multipole_lmax)
    if eos_tab is not None:
        c.eos_tab = ctypes.cast(ctypes.pointer(eos_tab), ctypes.c_void_p)
        c._eos_keep = eos_tab
    if prob_bounds is not None:
        def _cb(i, j, k, p, user, f=prob_bounds, nv=nvars):
            vals = f(i, j, k)
            for v in range(nv):
                p[v] = float(vals[v])
        cb = PROB_BOUNDS_FN(_cb)
        c.prob_bounds = ctypes.cast(cb, ctypes.c_void_p)
        c._prob_keep = cb
    if prob_bounds_t is not None:
        def _cbt(i, j, k, t, p, user, f=prob_bounds_t, nv=nvars):
            vals = f(i, j, k, t)
            for v in range(nv):
                p[v] = float(vals[v])
        cbt = PROB_BOUNDS_T_FN(_cbt)
        c.prob_bounds_t = ctypes.cast(cbt, ctypes.c_void_p)
        c._prob_t_keep = cbt
    return c


class Sim:
    """One oracle simulation (a restatement of the reference's globals + main loop)."""

    def __init__(self, cfg, fast=False):
        self.lib = lib(fast)
        self.cfg = cfg
        self.h = self.lib.orc_create(ctypes.byref(cfg))
        if not self.h:
            raise ValueError("oracle rejected the configuration (dendritic levels or geometry)")
        self.n = [cfg.n[d] if d < cfg.ndim else 1 for d in range(3)]
        self.nvars = cfg.nvars

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    @property
    def shape(self):
        return (self.nvars, self.n[2], self.n[1], self.n[0])

    def set_prims(self, prims):
        a = np.ascontiguousarray(prims, dtype=np.float64).reshape(self.shape)
        self.lib.orc_set_prims(self.h, ptr(a))

    def get_prims(self):
        a = np.zeros(self.shape, dtype=np.float64)
        self.lib.orc_get_prims(self.h, ptr(a))
        return a

    def fill_ghosts(self):
        """Apply the stage-start ghost fill (bc.c:705-1421) to the current state."""
        self.lib.orc_fill_ghosts(self.h)

    def zone(self, i, j=0, k=0):
        """Primitives of zone (i, j, k), ghost zones included (indices from -2)."""
        out = np.zeros(self.nvars)
        if self.lib.orc_zone_prims(self.h, i, j, k, ptr(out)) != 0:
            raise IndexError((i, j, k))
        return out

    def monopole_edges(self):
        out = np.zeros(self.n[0] + 1)
        self.lib.orc_monopole_edges(self.h, ptr(out))
        return out

    def gr_edges(self):
        """GR effective monopole face gravity of the current state (updates the kept TOV factor)."""
        out = np.zeros(self.n[0] + 1)
        if self.lib.orc_gr_edges(self.h, ptr(out)) < 0:
            raise ValueError("no GR monopole in this configuration")
        return out

    def multipole_phi(self):
        """multipole potential (l >= 1) at the zone corners of the current state:
        array [n1 + 1][n2 + 1][n3 + 1] over level-local corner indices (unused entries 0)."""
        out = np.zeros((self.n[0] + 1) * (self.n[1] + 1) * (self.n[2] + 1))
        if self.lib.orc_multipole_phi(self.h, ptr(out)) < 0:
            raise ValueError("no multipole gravity in this configuration")
        return out.reshape(self.n[0] + 1, self.n[1] + 1, self.n[2] + 1)

    def multipole_sources(self):
        """multipole momentum sources of the current state without fluxes: [3][n3][n2][n1]"""
        out = np.zeros(3 * self.n[0] * self.n[1] * self.n[2])
        if self.lib.orc_multipole_sources(self.h, ptr(out)) < 0:
            raise ValueError("no multipole gravity in this configuration")
        return np.moveaxis(out.reshape(self.n[2], self.n[1], self.n[0], 3), -1, 0)

    def gr_state(self):
        """GR effective monopole after the last stage: dict of gtov, mr (faces),
        lapse (zones), lapse_edge (faces), iter, fail."""
        n1 = self.n[0]
        g, m, lc, le = np.zeros(n1 + 1), np.zeros(n1 + 1), np.zeros(n1), np.zeros(n1 + 1)
        it = ctypes.c_int(0)
        rc = self.lib.orc_gr_state(self.h, ptr(g), ptr(m), ptr(lc), ptr(le), ctypes.byref(it))
        if rc < 0:
            raise ValueError("no GR monopole in this configuration")
        return dict(gtov=g, mr=m, lapse=lc, lapse_edge=le, iter=it.value, fail=rc)

    def run(self, tmax=1.0e300, max_steps=2 ** 31 - 1):
        return self.lib.orc_run(self.h, tmax, max_steps)

    @property
    def t(self):
        return self.lib.orc_time(self.h)

    @property
    def dt(self):
        return self.lib.orc_last_dt(self.h)

    @property
    def istep(self):
        return self.lib.orc_istep(self.h)

    def shock_faces(self):
        return self.lib.orc_shock_face_count(self.h)

    def hlle_faces(self):
        return self.lib.orc_hlle_face_count(self.h)

    def totals(self):
        out = np.zeros(2)
        self.lib.orc_conserved_totals(self.h, ptr(out))
        return out

    def geometry(self):
        nall = self.lib.orc_zone_count_all(self.h)
        g = self.lib.orc_geometry(self.h)
        arr = np.ctypeslib.as_array(ctypes.cast(g, ctypes.POINTER(ctypes.c_double)),
                                    shape=(nall, ctypes.sizeof(ZoneGeom) // 8))
        return arr.copy()

    def levels(self):
        """Dendritic level counts: (nj[n1], nk[n1, n2] with 0 beyond nj[i], maxjlev, maxklev)."""
        n1, n2 = self.n[0], self.n[1]
        nj = np.zeros(n1, dtype=np.int32)
        nk = np.zeros((n1, n2), dtype=np.int32)
        r = self.lib.orc_dendritic_levels(self.h, nj.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                          nk.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        return nj, nk, r // 256, r % 256

    def level_moments(self, d, lev):
        """theta (d=1) or phi (d=2) moments [alpha, beta, Gamma] of refinement level lev."""
        ng = self.cfg.n[d] + 4
        out = []
        for q in range(3):
            p = self.lib.orc_level_moment(self.h, q, d - 1, lev)
            if not p:
                return None
            out.append(np.ctypeslib.as_array(p, shape=(ng,)).copy())
        return out

    def prepare_entries(self):
        """Stage-start ghost fill and EOS update on the current state (before the entry evaluators)."""
        self.lib.orc_prepare_entries(self.h)

    def radial_entry(self, i, j, k, ii):
        """Row ii of the radial pencil of the face below zone (i, j, k) (restrict0 / prolongate0):
        nvars primitives, then P and Gamma_1."""
        out = np.zeros(self.nvars + 2)
        if self.lib.orc_dend_radial_entry(self.h, i, j, k, ii, ptr(out)) != 0:
            raise IndexError((i, j, k, ii))
        return out

    def theta_entry(self, i, j, kk, jj):
        """Row jj of the theta pencil of the face below row j of shell i at phi index kk of the face level."""
        out = np.zeros(self.nvars + 2)
        if self.lib.orc_dend_theta_entry(self.h, i, j, kk, jj, ptr(out)) != 0:
            raise IndexError((i, j, kk, jj))
        return out

    def polar_face_states(self, i, j, k):
        """(left, right) states of the phi face below zone (i, j, k) of the last theta row."""
        left, right = np.zeros(self.nvars + 2), np.zeros(self.nvars + 2)
        if self.lib.orc_polar_face_states(self.h, i, j, k, ptr(left), ptr(right)) != 0:
            raise IndexError((i, j, k))
        return left, right

    def moments(self, d):
        ng = self.cfg.n[d] + 4 if d < self.cfg.ndim else 5
        out = []
        for q in range(3):
            p = self.lib.orc_moment(self.h, q, d)
            out.append(np.ctypeslib.as_array(p, shape=(ng,)).copy())
        return out


# geometry record field offsets (in doubles) inside ZoneGeom
GEOM_AREA = 0
GEOM_CONN = 3
GEOM_GCOV = 30
GEOM_GCON = 42
GEOM_SCALE = 45
GEOM_VOLUME = 57


def ylms(th, phi, lmax, mmax):
    """real spherical harmonics of the multipole gravity (gravity.c:518-577), index l (l + 1) + m"""
    out = np.zeros(15 * 15)
    lib().orc_ylms(ptr(out), th, phi, lmax, mmax)
    return out[:(lmax + 1) ** 2]


def ylm_ints(thmin, thmax, phimin, phimax, lmax, mmax):
    """integrals of Y[l,m] sin(th) over a zone (gravity.c:609-697)"""
    out = np.zeros(15 * 15)
    lib().orc_ylm_ints(ptr(out), thmin, thmax, phimin, phimax, lmax, mmax)
    return out[:(lmax + 1) ** 2]
