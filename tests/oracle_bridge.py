"""ctypes bridge to the CPU oracle (oracle/rh3t_oracle.c).

Test infrastructure: the oracle is the checker for the HIP path and the
reported CPU baseline; nothing in the product imports this module.  The
ghost tables are read from a built :class:`Field` through the product's C
ABI (GhostGrid bookkeeping is checked separately, bit for bit, against a
brute-force recount in test_host.py), and the oracle evaluates the
reference's l() on them with the reference's own data model: AoS states,
ghost values by sorted id, and the 17-slot access table.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "_build", "librh3t_oracle.so")

_DP = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)
_U32P = ctypes.POINTER(ctypes.c_uint32)
_U8P = ctypes.POINTER(ctypes.c_uint8)


class ro_physics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "gamma_e", "gamma_i", "gamma_r", "kappa_e", "kappa_i", "kappa_r", "cve", "cvi", "a_rad",
        "omega_ei", "omega_er", "lambda_", "weno_q", "eps")]


class ro_derived(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in ("u", "v", "e_e", "e_i", "e_r", "pe", "pi", "pr", "te", "ti", "tr")]


class ro_problem(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("dx", ctypes.c_double), ("dy", ctypes.c_double),
                ("fluid", _U8P), ("access", _U32P), ("n_ghost", ctypes.c_int32), ("g_kind", _I32P),
                ("g_src", _I32P), ("g_heavy", _I32P), ("g_normal", _DP), ("g_dist", _DP), ("g_bcstate", _DP),
                ("g_outflow", _DP), ("h_nearest", _I32P), ("h_d", _DP), ("h_stencil", _U32P), ("h_s", _DP),
                ("h_w", _DP), ("time", ctypes.c_double), ("g_p0x", _DP), ("g_bcstate2", _DP), ("g_front", _DP)]


class ro_error(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("stage", ctypes.c_int32), ("i", ctypes.c_int32), ("j", ctypes.c_int32),
                ("ghost_id", ctypes.c_int32), ("state", ctypes.c_double * 6)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_LIB):
            build()
        L = ctypes.CDLL(ORACLE_LIB)
        L.ro_derived_from_state.argtypes = [ctypes.POINTER(ro_physics), _DP, ctypes.POINTER(ro_derived)]
        L.ro_weno5.argtypes = [_DP, ctypes.c_double]
        L.ro_weno5.restype = ctypes.c_double
        L.ro_build_l.argtypes = [ctypes.POINTER(ro_physics), _DP, _DP, ctypes.POINTER(ro_derived),
                                 ctypes.POINTER(ro_derived), ctypes.c_int, _DP]
        L.ro_build_r.argtypes = [ctypes.POINTER(ro_physics), _DP, _DP, ctypes.POINTER(ro_derived),
                                 ctypes.POINTER(ro_derived), ctypes.c_int, _DP, _DP]
        L.ro_reconstruction_fast.argtypes = [ctypes.POINTER(ro_physics), _DP, ctypes.POINTER(ro_derived),
                                             ctypes.c_int, ctypes.c_int, _DP]
        L.ro_noncon.argtypes = [ctypes.POINTER(ro_derived), ctypes.c_double, ctypes.c_int, _DP]
        L.ro_diffusion.argtypes = [ctypes.POINTER(ro_physics), ctypes.POINTER(ro_derived), _DP]
        L.ro_source.argtypes = [ctypes.POINTER(ro_physics), _DP, _DP]
        L.ro_local_dt.argtypes = [ctypes.POINTER(ro_physics), _DP, ctypes.c_double, ctypes.c_double]
        L.ro_local_dt.restype = ctypes.c_double
        L.ro_update_ghosts.argtypes = [ctypes.POINTER(ro_problem), ctypes.POINTER(ro_physics), _DP, _DP,
                                       ctypes.c_int, ctypes.POINTER(ro_error)]
        L.ro_l.argtypes = [ctypes.POINTER(ro_problem), ctypes.POINTER(ro_physics), _DP, _DP, _DP, ctypes.c_int,
                           ctypes.POINTER(ro_error)]
        L.ro_rk3_step.argtypes = [ctypes.POINTER(ro_problem), ctypes.POINTER(ro_physics), _DP, _DP,
                                  ctypes.c_double, ctypes.c_int, ctypes.POINTER(ro_error)]
        L.ro_global_dt.argtypes = [ctypes.POINTER(ro_problem), ctypes.POINTER(ro_physics), _DP]
        L.ro_global_dt.restype = ctypes.c_double
        _declare_heavy(L)
        _lib = L
    return _lib


def _declare_heavy(L):
    L.ro_heavy_terms.argtypes = [ctypes.POINTER(ro_problem), ctypes.POINTER(ro_physics), _DP, ctypes.c_int, _DP,
                                 _DP, _DP, _DP, _DP]
    L.ro_heavy_terms.restype = ctypes.c_int


def load_variant(path):
    """A separately compiled build of rh3t_oracle.c (same exports), e.g. a
    planted-mistake copy for the mutation tests; the default library stays
    in place (use :class:`using` to route calls through the variant)."""
    global ORACLE_LIB, _lib
    saved = (ORACLE_LIB, _lib)
    ORACLE_LIB, _lib = path, None
    try:
        return lib()
    finally:
        ORACLE_LIB, _lib = saved


class using:
    """Context manager: route every oracle call of this module through the
    given loaded library (restores the default afterwards)."""

    def __init__(self, library):
        self.library = library

    def __enter__(self):
        global _lib
        self.saved = _lib
        _lib = self.library
        return self.library

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def physics(**kw):
    """constant.rs defaults (src/constant.rs:1-25), overridable."""
    p = ro_physics(1.4, 1.4, 1.4, 0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.0, 0.0, 0.5, 10.0, 1e-12)
    for k, v in kw.items():
        setattr(p, k, float(v))
    return p


def _dp(a):
    return a.ctypes.data_as(_DP)


class OracleProblem:
    """ro_problem built from a Field's ghost tables (arrays kept alive)."""

    def __init__(self, field):
        g = field.grid
        self.nx, self.ny = g.nx, g.ny
        self.fluid = np.ascontiguousarray(field.fluid.reshape(-1).astype(np.uint8))
        self.access = field.access_table()
        gh = field.ghosts()
        n = len(gh)
        self.n_ghost = n
        self.g_kind = np.ascontiguousarray(gh["bc_kind"].astype(np.int32))
        src = np.where(gh["src_i"] >= 0, gh["src_i"].astype(np.int64) * g.ny + gh["src_j"], -1)
        self.g_src = np.ascontiguousarray(src.astype(np.int32))
        self.g_heavy = np.ascontiguousarray(gh["heavy_id"].astype(np.int32))
        self.g_normal = np.ascontiguousarray(np.stack([gh["nx"], gh["ny"]], axis=1).reshape(-1))
        self.g_dist = np.ascontiguousarray(gh["distance"].astype(np.float64))
        bcs = [field.bc(int(b), int(s)) for b, s in zip(gh["boundary"], gh["side_id"])]
        self.g_bcstate = np.ascontiguousarray(np.array([list(b.state) for b in bcs] or [[0.0] * 6]).reshape(-1))
        self.g_outflow = np.ascontiguousarray(np.array([[b.p_inf, b.sigma, b.l_domain] for b in bcs]
                                                       or [[0.0] * 3]).reshape(-1))
        self.g_p0x = np.ascontiguousarray(gh["p0x"].astype(np.float64) if n else np.zeros(1))
        self.g_bcstate2 = np.ascontiguousarray(np.array([list(b.state2) for b in bcs] or [[0.0] * 6]).reshape(-1))
        self.g_front = np.ascontiguousarray(np.array([[b.speed, b.front0] for b in bcs] or [[0.0] * 2]).reshape(-1))
        nh = field.counts()[2]
        near, d5, st, S, W = [], [], [], [], []
        for h in range(nh):
            b = field.ghost_bc(h)
            near.append(b.nearest_i * g.ny + b.nearest_j)
            d5.append(list(b.d))
            st.append(list(b.stencil))
            S.append(np.frombuffer(b.s, dtype=np.float64).copy())
            W.append(np.frombuffer(b.w, dtype=np.float64).copy())
        self.h_nearest = np.ascontiguousarray(np.array(near or [0], dtype=np.int32))
        self.h_d = np.ascontiguousarray(np.array(d5 or [[0.0] * 5], dtype=np.float64).reshape(-1))
        self.h_stencil = np.ascontiguousarray(np.array(st or [[0] * 55], dtype=np.uint32).reshape(-1))
        self.h_s = np.ascontiguousarray(np.concatenate(S) if S else np.zeros(979))
        self.h_w = np.ascontiguousarray(np.concatenate(W) if W else np.zeros(225))
        self.c = ro_problem(g.nx, g.ny, g.dx, g.dy, self.fluid.ctypes.data_as(_U8P),
                            self.access.ctypes.data_as(_U32P), n, self.g_kind.ctypes.data_as(_I32P),
                            self.g_src.ctypes.data_as(_I32P), self.g_heavy.ctypes.data_as(_I32P),
                            _dp(self.g_normal), _dp(self.g_dist), _dp(self.g_bcstate), _dp(self.g_outflow),
                            self.h_nearest.ctypes.data_as(_I32P), _dp(self.h_d),
                            self.h_stencil.ctypes.data_as(_U32P), _dp(self.h_s), _dp(self.h_w), 0.0,
                            _dp(self.g_p0x), _dp(self.g_bcstate2), _dp(self.g_front))


class OracleError(RuntimeError):
    def __init__(self, status, err):
        self.status = status
        self.info = {"kind": err.kind, "stage": err.stage, "i": err.i, "j": err.j, "ghost_id": err.ghost_id,
                     "state": list(err.state)}
        super().__init__(f"oracle status {status}: {self.info}")


def l(prob, u, phys=None, nthreads=8, with_ghosts=False, time=0.0):
    """l() of state u at Field.time `time` (the time TimeDependent ghosts
    are evaluated at)."""
    phys = phys or physics()
    prob.c.time = float(time)
    u = np.ascontiguousarray(u, dtype=np.float64)
    rhs = np.zeros_like(u)
    gv = np.zeros((max(prob.n_ghost, 1), 6))
    err = ro_error()
    st = lib().ro_l(ctypes.byref(prob.c), ctypes.byref(phys), _dp(u), _dp(rhs), _dp(gv), nthreads,
                    ctypes.byref(err))
    if st:
        raise OracleError(st, err)
    return (rhs, gv[: prob.n_ghost]) if with_ghosts else rhs


def rk3_step(prob, u, dt, time=0.0, phys=None, nthreads=8):
    phys = phys or physics()
    u = np.array(u, dtype=np.float64, copy=True, order="C")
    t = ctypes.c_double(time)
    err = ro_error()
    st = lib().ro_rk3_step(ctypes.byref(prob.c), ctypes.byref(phys), _dp(u), ctypes.byref(t), float(dt), nthreads,
                           ctypes.byref(err))
    if st:
        raise OracleError(st, err)
    return u, t.value


def global_dt(prob, u, phys=None):
    phys = phys or physics()
    u = np.ascontiguousarray(u, dtype=np.float64)
    return lib().ro_global_dt(ctypes.byref(prob.c), ctypes.byref(phys), _dp(u))


def derived(s, phys=None):
    phys = phys or physics()
    d = ro_derived()
    s = np.ascontiguousarray(s, dtype=np.float64)
    lib().ro_derived_from_state(ctypes.byref(phys), _dp(s), ctypes.byref(d))
    return d


def weno5(st, eps=1e-12):
    a = np.ascontiguousarray(st, dtype=np.float64)
    return lib().ro_weno5(_dp(a), eps)


def build_lr(s3, s2, direction, phys=None):
    phys = phys or physics()
    s3 = np.ascontiguousarray(s3, dtype=np.float64)
    s2 = np.ascontiguousarray(s2, dtype=np.float64)
    d3, d2 = derived(s3, phys), derived(s2, phys)
    L = np.zeros(36)
    R = np.zeros(36)
    lam = np.zeros(6)
    lib().ro_build_l(ctypes.byref(phys), _dp(s3), _dp(s2), ctypes.byref(d3), ctypes.byref(d2), direction, _dp(L))
    lib().ro_build_r(ctypes.byref(phys), _dp(s3), _dp(s2), ctypes.byref(d3), ctypes.byref(d2), direction, _dp(lam),
                     _dp(R))
    return L.reshape(6, 6), R.reshape(6, 6), lam


def reconstruction(pts, direction, recon=1, phys=None):
    phys = phys or physics()
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(36)
    d = (ro_derived * 6)(*[derived(pts[6 * k: 6 * k + 6], phys) for k in range(6)])
    out = np.zeros(6)
    lib().ro_reconstruction_fast(ctypes.byref(phys), _dp(pts), d, direction, recon, _dp(out))
    return out


def noncon(states9, ds, direction, phys=None):
    phys = phys or physics()
    d = (ro_derived * 9)(*[derived(s, phys) for s in states9])
    out = np.zeros(6)
    lib().ro_noncon(d, float(ds), direction, _dp(out))
    return out


def local_dt(s, dx, dy, phys=None):
    phys = phys or physics()
    s = np.ascontiguousarray(s, dtype=np.float64)
    return lib().ro_local_dt(ctypes.byref(phys), _dp(s), dx, dy)


def source(s, phys):
    s = np.ascontiguousarray(s, dtype=np.float64)
    out = np.zeros(6)
    lib().ro_source(ctypes.byref(phys), _dp(s), _dp(out))
    return out


def heavy_terms(prob, u, g, phys=None):
    """Internals of heavy ghost g: L, R (6x6), lambda, V^(k) (5x6) and the
    normal derivatives u^(k) (5x6) of the conservative state at the boundary
    point in the (n, t) frame."""
    phys = phys or physics()
    u = np.ascontiguousarray(u, dtype=np.float64)
    L, R, lam, v, uk = np.zeros(36), np.zeros(36), np.zeros(6), np.zeros(30), np.zeros(30)
    st = lib().ro_heavy_terms(ctypes.byref(prob.c), ctypes.byref(phys), _dp(u), int(g), _dp(L), _dp(R), _dp(lam),
                              _dp(v), _dp(uk))
    if st:
        raise OracleError(st, ro_error())
    return {"L": L.reshape(6, 6), "R": R.reshape(6, 6), "lam": lam, "v": v.reshape(5, 6), "uk": uk.reshape(5, 6)}


def update_ghosts(prob, u, phys=None, nthreads=8, time=0.0):
    """GhostGrid::update_values_parallel: ghost values (n_ghost, 6)."""
    phys = phys or physics()
    prob.c.time = float(time)
    u = np.ascontiguousarray(u, dtype=np.float64)
    gv = np.zeros((max(prob.n_ghost, 1), 6))
    err = ro_error()
    st = lib().ro_update_ghosts(ctypes.byref(prob.c), ctypes.byref(phys), _dp(u), _dp(gv), nthreads,
                                ctypes.byref(err))
    if st:
        raise OracleError(st, err)
    return gv[: prob.n_ghost]


def relerr(a, b, fluid=None):
    """Max-norm relative error per field: max|a-b| / max|b| over fluid cells
    (absolute where the reference field is identically zero)."""
    a = np.asarray(a).reshape(-1, 6)
    b = np.asarray(b).reshape(-1, 6)
    if fluid is not None:
        m = np.asarray(fluid).reshape(-1)
        a, b = a[m], b[m]
    out = []
    for f in range(6):
        scale = np.max(np.abs(b[:, f]))
        d = np.max(np.abs(a[:, f] - b[:, f]))
        out.append(d / scale if scale > 0 else d)
    return np.array(out)


# ------------------------------------------------- forward-mode oracle

DUAL_LIB = os.path.join(ORACLE_DIR, "_build", "librh3t_dual.so")
_dual = None


def dual_lib():
    """oracle/dual.cpp: rh3t_oracle.c instantiated with dual numbers."""
    global _dual
    if _dual is None:
        if not os.path.exists(DUAL_LIB):
            build()
        L = ctypes.CDLL(DUAL_LIB)
        phys_p = ctypes.POINTER(ro_physics)
        L.rod_l_jvp.argtypes = [ctypes.POINTER(ro_problem), phys_p, phys_p, _DP, _DP, _DP, _DP, ctypes.c_int]
        L.rod_rk3_jvp.argtypes = [ctypes.POINTER(ro_problem), phys_p, phys_p, _DP, _DP, _DP, ctypes.c_int,
                                  ctypes.c_int]
        _dual = L
    return _dual


PARAMS = ("omega_ei", "omega_er", "kappa_e", "kappa_i", "kappa_r")


def _dphys(dparams):
    """ro_physics holding only the tangents of the five parameters."""
    d = ro_physics(*([0.0] * 14))
    if dparams is not None:
        for name, v in zip(PARAMS, dparams):
            setattr(d, name, float(v))
    return d


def l_jvp(prob, u, du, phys=None, dparams=None, nthreads=8):
    """(l(u), d l(u)[du, dparams]) from the dual-number oracle."""
    phys = phys or physics()
    u = np.ascontiguousarray(u, dtype=np.float64)
    du = np.ascontiguousarray(du, dtype=np.float64)
    rhs, drhs = np.zeros_like(u), np.zeros_like(u)
    dp = _dphys(dparams)
    st = dual_lib().rod_l_jvp(ctypes.byref(prob.c), ctypes.byref(phys), ctypes.byref(dp), _dp(u), _dp(du),
                              _dp(rhs), _dp(drhs), nthreads)
    if st:
        raise OracleError(st, ro_error())
    return rhs, drhs


def rk3_jvp(prob, u, du, dts, phys=None, dparams=None, nthreads=8):
    """State and tangent after len(dts) rk3_ssp steps with frozen dts."""
    phys = phys or physics()
    u = np.array(u, dtype=np.float64, copy=True, order="C")
    du = np.array(du, dtype=np.float64, copy=True, order="C")
    dts = np.ascontiguousarray(dts, dtype=np.float64)
    dp = _dphys(dparams)
    st = dual_lib().rod_rk3_jvp(ctypes.byref(prob.c), ctypes.byref(phys), ctypes.byref(dp), _dp(u), _dp(du),
                                _dp(dts), len(dts), nthreads)
    if st:
        raise OracleError(st, ro_error())
    return u, du
