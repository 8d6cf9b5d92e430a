"""Python mirror of the l() / rk3_ssp() path of colinhu123/3T_RH on MI355X.

The package loads ``librh3t.so`` (host Field/GhostGrid construction + gfx950
HIP kernels, built by ``build.py``) through ctypes and mirrors the
reference's data model:

=======================  ==================================================
reference (Rust)         here
=======================  ==================================================
``BCType``               :class:`BCType` factories (src/bc1.rs:16-28)
``GridInfo``             :class:`GridInfo` (src/field1.rs:5-73)
``Field`` + ``GhostGrid``:class:`Field` (src/field1.rs:75-160,
                         src/ghost.rs:29-118): geometry, fluid mask,
                         ghost tables; ``build_ghosts()``
``l(u, ghosts, s)``      :meth:`Solver.l` (src/main.rs:135-305)
``rk3_ssp``              :meth:`Solver.rk3_ssp` (src/main.rs:351-373)
``calc_global_dt``       :meth:`Solver.calc_global_dt` (src/main.rs:375-408)
time loop body           :meth:`Solver.advance` (src/main.rs:459-467)
=======================  ==================================================

Errors the reference turns into panics raise :class:`RH3TError` carrying the
first failure (kind, RK stage, grid index, state).  All state arrays are
float64 numpy arrays of shape (nx, ny, 6) in the reference's AoS order.
There is no CPU fallback: without the built library every entry point
raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "librh3t.so")

OK = 0
ERR_ARG, ERR_HIP, ERR_NONPHYSICAL, ERR_GHOST, ERR_GEOMETRY = 1, 2, 3, 4, 5
ERR_UNSUPPORTED, ERR_NO_DEVICE, ERR_SINGULAR, ERR_DT, ERR_SIZE = 6, 7, 8, 9, 10
STATUS_NAMES = {
    0: "ok", 1: "bad argument", 2: "HIP error", 3: "non-physical state", 4: "non-finite ghost state",
    5: "geometry/ghost build error", 6: "unsupported", 7: "no HIP device", 8: "singular ILW system",
    9: "non-positive dt", 10: "input beyond a fixed-capacity entry point",
}

BC_REFLECTIVE_WALL, BC_WALL, BC_PERIODIC, BC_CONSTANT = 0, 1, 2, 3
BC_TIME_DEPENDENT, BC_OUTFLOW, BC_ZEROTH_ORDER, BC_FARFIELD = 4, 5, 6, 7
ARITH_EXACT, ARITH_FAST = 0, 1
TD_MOVING_FRONT = 1  # TimeDependent profile: planar front normal to x


class RH3TError(RuntimeError):
    def __init__(self, status, message="", error=None):
        self.status = status
        self.error = error
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")


# ----------------------------------------------------------------- C types

class c_bc(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("profile", ctypes.c_int32), ("state", ctypes.c_double * 6),
                ("p_inf", ctypes.c_double), ("sigma", ctypes.c_double), ("l_domain", ctypes.c_double),
                ("state2", ctypes.c_double * 6), ("speed", ctypes.c_double), ("front0", ctypes.c_double)]


class c_physics(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "gamma_e", "gamma_i", "gamma_r", "kappa_e", "kappa_i", "kappa_r", "cve", "cvi", "a_rad",
        "omega_ei", "omega_er", "lambda_", "weno_q", "eps")]


class c_grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("dx", ctypes.c_double),
                ("dy", ctypes.c_double), ("x0", ctypes.c_double), ("y0", ctypes.c_double)]


class c_ghost(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in (
        "i", "j", "boundary", "side_id", "bc_kind", "heavy_id", "src_i", "src_j", "nearest_i",
        "nearest_j")] + [(n, ctypes.c_double) for n in ("p0x", "p0y", "nx", "ny", "distance")]


class c_ghost_bc(ctypes.Structure):
    _fields_ = [("nearest_i", ctypes.c_int32), ("nearest_j", ctypes.c_int32), ("normal", ctypes.c_double * 2),
                ("d", ctypes.c_double * 5), ("stencil", ctypes.c_uint32 * 55), ("s", ctypes.c_double * 979),
                ("w", ctypes.c_double * 225)]


class c_error(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("stage", ctypes.c_int32), ("i", ctypes.c_int32),
                ("j", ctypes.c_int32), ("ghost_id", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("state", ctypes.c_double * 6), ("message", ctypes.c_char * 256)]


_P = ctypes.c_void_p
_DP = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes); the complete C ABI of include/rh3t.h
SIGNATURES = {
    "rh3t_physics_default": (None, [ctypes.POINTER(c_physics)]),
    "rh3t_problem_create": (ctypes.c_int, [ctypes.POINTER(c_grid), _DP, ctypes.c_int32, ctypes.POINTER(c_bc),
                                           _DP, ctypes.c_int32, ctypes.POINTER(c_bc), ctypes.POINTER(_P)]),
    "rh3t_problem_destroy": (None, [_P]),
    "rh3t_problem_add_arc": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                            _DP, _DP, _DP]),
    "rh3t_problem_build_ghosts": (ctypes.c_int, [_P]),
    "rh3t_problem_init_cylinder": (ctypes.c_int, [ctypes.c_int32, ctypes.c_double, ctypes.POINTER(_P)]),
    "rh3t_problem_cylinder_state": (ctypes.c_int, [_P, ctypes.c_double, _DP]),
    "rh3t_problem_clone": (ctypes.c_int, [_P, ctypes.POINTER(_P)]),
    "rh3t_problem_init_double_mach": (ctypes.c_int, [ctypes.c_double, ctypes.POINTER(_P)]),
    "rh3t_problem_init_double_mach_top": (ctypes.c_int, [ctypes.c_double, ctypes.POINTER(c_bc), ctypes.POINTER(_P)]),
    "rh3t_problem_double_mach_state": (ctypes.c_int, [_P, _DP]),
    "rh3t_problem_set_bc_state": (ctypes.c_int, [_P, ctypes.c_int32, _DP]),
    "rh3t_problem_grid": (ctypes.c_int, [_P, ctypes.POINTER(c_grid)]),
    "rh3t_problem_counts": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                           ctypes.POINTER(ctypes.c_int64)]),
    "rh3t_problem_fluid_mask": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint8)]),
    "rh3t_problem_ghosts": (ctypes.c_int, [_P, ctypes.POINTER(c_ghost)]),
    "rh3t_problem_ghost_bc": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(c_ghost_bc)]),
    "rh3t_problem_bc": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.POINTER(c_bc)]),
    "rh3t_problem_access_table": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_uint32)]),
    "rh3t_device_count": (ctypes.c_int, [ctypes.POINTER(ctypes.c_int32)]),
    "rh3t_ctx_create": (ctypes.c_int, [_P, ctypes.POINTER(c_physics), ctypes.c_int32, ctypes.POINTER(_P)]),
    "rh3t_ctx_destroy": (None, [_P]),
    "rh3t_set_state": (ctypes.c_int, [_P, _DP, ctypes.c_double]),
    "rh3t_get_state": (ctypes.c_int, [_P, _DP, _DP]),
    "rh3t_l": (ctypes.c_int, [_P, _DP, _DP]),
    "rh3t_rk3_step": (ctypes.c_int, [_P, ctypes.c_double]),
    "rh3t_calc_global_dt": (ctypes.c_int, [_P, _DP]),
    "rh3t_calc_global_dt_noalloc": (ctypes.c_int, [_P, _P, ctypes.POINTER(c_grid), ctypes.POINTER(c_physics), _P,
                                                   _P]),
    "rh3t_advance": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double, ctypes.c_double, _DP]),
    "rh3t_ctx_set_arith": (ctypes.c_int, [_P, ctypes.c_int32]),
    "rh3t_ctx_set_launch": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "rh3t_ctx_set_launch_x": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "rh3t_batch_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int32, ctypes.POINTER(c_physics),
                                         ctypes.c_int32, ctypes.POINTER(_P)]),
    "rh3t_batch_destroy": (None, [_P]),
    "rh3t_batch_set_state": (ctypes.c_int, [_P, ctypes.c_int32, _DP, ctypes.c_double]),
    "rh3t_batch_get_state": (ctypes.c_int, [_P, ctypes.c_int32, _DP, _DP]),
    "rh3t_batch_advance": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double, _DP]),
    "rh3t_batch_set_launch": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "rh3t_batch_set_launch_x": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    "rh3t_batch_set_arith": (ctypes.c_int, [_P, ctypes.c_int32]),
    "rh3t_batch_last_error": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(c_error)]),
    "rh3t_batch_profile": (ctypes.c_int, [_P, ctypes.c_int32]),
    "rh3t_batch_profile_read": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, _DP,
                                               ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
    "rh3t_ctx_set_physics": (ctypes.c_int, [_P, ctypes.POINTER(c_physics)]),
    "rh3t_l_vjp": (ctypes.c_int, [_P, _DP, _DP, _DP]),
    "rh3t_rk3_vjp": (ctypes.c_int, [_P, ctypes.c_double, _DP, _DP, _DP]),
    "rh3t_adjoint_create": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.POINTER(_P)]),
    "rh3t_adjoint_destroy": (None, [_P]),
    "rh3t_adjoint_forward": (ctypes.c_int, [_P, ctypes.c_int32, _DP]),
    "rh3t_adjoint_backward": (ctypes.c_int, [_P, _DP, _DP, _DP]),
    "rh3t_adjoint_device_ms": (ctypes.c_int, [_P, _DP, _DP]),
    "rh3t_last_error": (ctypes.c_int, [_P, ctypes.POINTER(c_error)]),
    "rh3t_profile": (ctypes.c_int, [_P, ctypes.c_int32]),
    "rh3t_profile_read": (ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_char_p, ctypes.c_int32, _DP,
                                         ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int32)]),
}

_lib = None


def lib():
    """The loaded librh3t.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RH3TError(ERR_UNSUPPORTED, f"{LIB_PATH} is not built; run 3t_rh_amd/build.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status, what=""):
    if status != OK:
        raise RH3TError(status, what)


def _dptr(a):
    return a.ctypes.data_as(_DP)


# ------------------------------------------------------------- data model

class Physics:
    """constant.rs as a runtime struct (src/constant.rs:1-25)."""

    FIELDS = ("gamma_e", "gamma_i", "gamma_r", "kappa_e", "kappa_i", "kappa_r", "cve", "cvi", "a_rad",
              "omega_ei", "omega_er", "lambda_", "weno_q", "eps")

    def __init__(self, **kw):
        c = c_physics()
        lib().rh3t_physics_default(ctypes.byref(c))
        for k, v in kw.items():
            if k not in self.FIELDS:
                raise TypeError(f"unknown physics constant {k}")
            setattr(c, k, float(v))
        self.c = c

    def __getattr__(self, k):
        if k in Physics.FIELDS:
            return getattr(self.c, k)
        raise AttributeError(k)

    def as_tuple(self):
        return tuple(getattr(self.c, k) for k in self.FIELDS)

    def replace(self, **kw):
        """A copy with some constants changed."""
        q = {k: getattr(self.c, k) for k in self.FIELDS}
        q.update(kw)
        return Physics(**q)

    def params(self):
        """The differentiable parameters, in PARAMS order."""
        return np.array([getattr(self.c, k) for k in PARAMS])

    def with_params(self, values):
        return self.replace(**dict(zip(PARAMS, (float(v) for v in values))))


# the parameters the reverse mode differentiates (src/constant.rs:17-25)
PARAMS = ("omega_ei", "omega_er", "kappa_e", "kappa_i", "kappa_r")


class BCType:
    """Factories for BCType (src/bc1.rs:16-28)."""

    @staticmethod
    def _mk(kind, state=None, p_inf=0.0, sigma=0.0, l_domain=0.0):
        b = c_bc()
        b.kind = kind
        if state is not None:
            for k in range(6):
                b.state[k] = float(state[k])
        b.p_inf, b.sigma, b.l_domain = p_inf, sigma, l_domain
        return b

    reflective_wall = staticmethod(lambda: BCType._mk(BC_REFLECTIVE_WALL))
    wall = staticmethod(lambda: BCType._mk(BC_WALL))
    periodic = staticmethod(lambda: BCType._mk(BC_PERIODIC))
    zeroth_order = staticmethod(lambda: BCType._mk(BC_ZEROTH_ORDER))
    constant = staticmethod(lambda state: BCType._mk(BC_CONSTANT, state))
    farfield = staticmethod(lambda state: BCType._mk(BC_FARFIELD, state))
    outflow = staticmethod(lambda p_inf, sigma, l_domain: BCType._mk(BC_OUTFLOW, None, p_inf, sigma, l_domain))

    @staticmethod
    def moving_front(behind, ahead, speed, front0=0.0):
        """TimeDependent (src/bc1.rs:634-636) with the built-in moving front
        profile: `behind` where P0.x <= front0 + speed * t, `ahead`
        elsewhere -- the moving Mach-10 shock closure of init_double_mach
        (src/init.rs:61-65) with behind = post, ahead = pre, speed = 10."""
        b = BCType._mk(BC_TIME_DEPENDENT, behind)
        b.profile = TD_MOVING_FRONT
        for k in range(6):
            b.state2[k] = float(ahead[k])
        b.speed, b.front0 = float(speed), float(front0)
        return b


class GridInfo:
    def __init__(self, nx, ny, dx, dy, x0, y0):
        self.nx, self.ny, self.dx, self.dy, self.x0, self.y0 = int(nx), int(ny), float(dx), float(dy), \
            float(x0), float(y0)

    def x(self, i):
        return self.x0 + float(i) * self.dx

    def y(self, j):
        return self.y0 + float(j) * self.dy


_DUMMY_INNER = [(-1002.0, -1002.0), (-1001.0, -1002.0), (-1001.0, -1001.0), (-1002.0, -1001.0)]


class Field:
    """Field + GhostGrid of the reference: grid, polygons, BC lists, arcs,
    fluid mask and (after :meth:`build_ghosts`) the ghost tables."""

    def __init__(self, grid, outer, bc_outer, inner=None, bc_inner=None, _handle=None):
        self._h = ctypes.c_void_p()
        if _handle is not None:
            self._h = _handle
        else:
            inner = inner if inner is not None else _DUMMY_INNER
            bc_inner = bc_inner if bc_inner is not None else [BCType.wall() for _ in inner]
            g = c_grid(grid.nx, grid.ny, grid.dx, grid.dy, grid.x0, grid.y0)
            o = np.ascontiguousarray(np.asarray(outer, dtype=np.float64).reshape(-1))
            n = np.ascontiguousarray(np.asarray(inner, dtype=np.float64).reshape(-1))
            bo = (c_bc * len(bc_outer))(*bc_outer)
            bi = (c_bc * len(bc_inner))(*bc_inner)
            _check(lib().rh3t_problem_create(ctypes.byref(g), _dptr(o), len(bc_outer), bo, _dptr(n),
                                             len(bc_inner), bi, ctypes.byref(self._h)), "Field::new")
        gi = c_grid()
        _check(lib().rh3t_problem_grid(self._h, ctypes.byref(gi)))
        self.grid = GridInfo(gi.nx, gi.ny, gi.dx, gi.dy, gi.x0, gi.y0)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and _lib is not None:
            _lib.rh3t_problem_destroy(h)
            self._h = None

    @classmethod
    def init_cylinder(cls, nx=121, mach=3.0):
        """init::init_cylinder (src/init.rs:116-510); ghosts built."""
        h = ctypes.c_void_p()
        _check(lib().rh3t_problem_init_cylinder(nx, mach, ctypes.byref(h)), "init_cylinder")
        f = cls(None, None, None, _handle=h)
        f.mach = mach
        return f

    @classmethod
    def init_double_mach(cls, h=1.0 / 320.0, top=None):
        """init::init_double_mach (src/init.rs:26-113) with mesh width h
        (1/320 in the reference); ghosts built.  top: the BC of the top side
        (None: the reference's ZerothOrder; e.g. double_mach_moving_shock()
        for the closure the reference defines there)."""
        hd = ctypes.c_void_p()
        if top is None:
            st = lib().rh3t_problem_init_double_mach(float(h), ctypes.byref(hd))
        else:
            st = lib().rh3t_problem_init_double_mach_top(float(h), ctypes.byref(top), ctypes.byref(hd))
        _check(st, "init_double_mach")
        return cls(None, None, None, _handle=hd)

    def double_mach_state(self):
        """Initial state of init_double_mach (src/init.rs:97-104)."""
        u = np.zeros((self.grid.nx, self.grid.ny, 6))
        _check(lib().rh3t_problem_double_mach_state(self._h, _dptr(u)), "double_mach_state")
        return u

    def clone(self):
        """Copy sharing the (geometry-only) heavy-ghost tables."""
        h = ctypes.c_void_p()
        _check(lib().rh3t_problem_clone(self._h, ctypes.byref(h)), "clone")
        f = Field(None, None, None, _handle=h)
        f.mach = getattr(self, "mach", None)
        return f

    def set_bc_state(self, kind, state):
        st = np.ascontiguousarray(state, dtype=np.float64)
        _check(lib().rh3t_problem_set_bc_state(self._h, kind, _dptr(st)), "set_bc_state")

    def cylinder_state(self, mach=None):
        u = np.zeros((self.grid.nx, self.grid.ny, 6))
        _check(lib().rh3t_problem_cylinder_state(self._h, self.mach if mach is None else mach, _dptr(u)))
        return u

    def add_arc(self, boundary, side_start, side_end, start, mid, end, fluid_outside=True):
        s, m, e = (np.asarray(p, dtype=np.float64) for p in (start, mid, end))
        _check(lib().rh3t_problem_add_arc(self._h, boundary, side_start, side_end, int(fluid_outside),
                                          _dptr(s), _dptr(m), _dptr(e)), "add_arc")

    def build_ghosts(self):
        _check(lib().rh3t_problem_build_ghosts(self._h), "GhostGrid::build")
        return self

    def counts(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().rh3t_problem_counts(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    @property
    def fluid(self):
        m = np.zeros(self.grid.nx * self.grid.ny, dtype=np.uint8)
        _check(lib().rh3t_problem_fluid_mask(self._h, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))))
        return m.reshape(self.grid.nx, self.grid.ny).astype(bool)

    def ghosts(self):
        """Structured array of the sorted ghost list (GhostInfo)."""
        n = self.counts()[1]
        arr = (c_ghost * max(n, 1))()
        _check(lib().rh3t_problem_ghosts(self._h, arr))
        dt = np.dtype([(f, np.int32) for f, _ in c_ghost._fields_[:10]] +
                      [(f, np.float64) for f, _ in c_ghost._fields_[10:]])
        return np.frombuffer(bytes(arr), dtype=dt, count=n).copy()

    def ghost_bc(self, heavy_id):
        g = c_ghost_bc()
        _check(lib().rh3t_problem_ghost_bc(self._h, heavy_id, ctypes.byref(g)))
        return g

    def bc(self, boundary, side):
        b = c_bc()
        _check(lib().rh3t_problem_bc(self._h, boundary, side, ctypes.byref(b)))
        return b

    def access_table(self):
        a = np.zeros(self.grid.nx * self.grid.ny * 17, dtype=np.uint32)
        _check(lib().rh3t_problem_access_table(self._h, a.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))),
               "access table")
        return a

    def solver(self, physics=None, device=0, arith=ARITH_EXACT):
        return Solver(self, physics, device, arith)


def device_count():
    n = ctypes.c_int32()
    lib().rh3t_device_count(ctypes.byref(n))
    return n.value


class Solver:
    """Device context: the l() / rk3_ssp() path on one GPU."""

    def __init__(self, field, physics=None, device=0, arith=ARITH_EXACT):
        self.field = field
        self.physics = physics if physics is not None else Physics()
        self._h = ctypes.c_void_p()
        st = lib().rh3t_ctx_create(field._h, ctypes.byref(self.physics.c), device, ctypes.byref(self._h))
        if st != OK:
            raise RH3TError(st, "rh3t_ctx_create")
        _check(lib().rh3t_ctx_set_arith(self._h, arith))
        self.nx, self.ny = field.grid.nx, field.grid.ny

    def close(self):
        if getattr(self, "_h", None):
            lib().rh3t_ctx_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def _raise(self, st, what):
        e = c_error()
        lib().rh3t_last_error(self._h, ctypes.byref(e))
        info = {"kind": e.kind, "stage": e.stage, "i": e.i, "j": e.j, "ghost_id": e.ghost_id,
                "state": list(e.state), "message": e.message.decode(errors="replace")}
        raise RH3TError(st, f"{what}: {info['message']}", info)

    def set_arith(self, mode):
        _check(lib().rh3t_ctx_set_arith(self._h, mode))

    def set_launch(self, bx, by, ghosts_per_block=1, x_shape=None):
        """Workgroup shape of the flux / assembly kernels and heavy ghosts per
        workgroup; x_shape = (bx, by) gives the x-flux kernel its own shape
        ((0, 0): the same as the others)."""
        _check(lib().rh3t_ctx_set_launch(self._h, bx, by, ghosts_per_block), "set_launch")
        if x_shape is not None:
            _check(lib().rh3t_ctx_set_launch_x(self._h, *x_shape), "set_launch x")

    def set_state(self, u, time=0.0):
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (self.nx, self.ny, 6):
            raise ValueError(f"state must have shape {(self.nx, self.ny, 6)}")
        _check(lib().rh3t_set_state(self._h, _dptr(u), float(time)), "set_state")

    def get_state(self):
        u = np.zeros((self.nx, self.ny, 6))
        t = ctypes.c_double()
        _check(lib().rh3t_get_state(self._h, _dptr(u), ctypes.byref(t)), "get_state")
        return u, t.value

    def l(self, with_ghosts=False):
        rhs = np.zeros((self.nx, self.ny, 6))
        ng = self.field.counts()[1]
        gv = np.zeros((max(ng, 1), 6)) if with_ghosts else None
        st = lib().rh3t_l(self._h, _dptr(rhs), _dptr(gv) if with_ghosts else None)
        if st != OK:
            self._raise(st, "l()")
        return (rhs, gv[:ng]) if with_ghosts else rhs

    def rk3_ssp(self, dt):
        st = lib().rh3t_rk3_step(self._h, float(dt))
        if st != OK:
            self._raise(st, "rk3_ssp")

    def calc_global_dt(self):
        d = ctypes.c_double()
        _check(lib().rh3t_calc_global_dt(self._h, ctypes.byref(d)), "calc_global_dt")
        return d.value

    def advance(self, nsteps, cfl_factor=0.05, dt_fixed=0.0):
        t = ctypes.c_double()
        st = lib().rh3t_advance(self._h, int(nsteps), float(cfl_factor), float(dt_fixed), ctypes.byref(t))
        if st != OK:
            self._raise(st, "advance")
        return t.value

    def set_physics(self, physics):
        """Replaces the physics constants (e.g. other exchange coefficients)."""
        self.physics = physics
        _check(lib().rh3t_ctx_set_physics(self._h, ctypes.byref(physics.c)), "set_physics")

    def _cot(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        if a.shape != (self.nx, self.ny, 6):
            raise ValueError(f"cotangent must have shape {(self.nx, self.ny, 6)}")
        return a

    def l_vjp(self, rbar):
        """((dl/du)^T rbar, d<rbar, l>/d PARAMS) at the current state."""
        rbar = self._cot(rbar)
        ubar, pg = np.zeros_like(rbar), np.zeros(len(PARAMS))
        st = lib().rh3t_l_vjp(self._h, _dptr(rbar), _dptr(ubar), _dptr(pg))
        if st != OK:
            self._raise(st, "l_vjp")
        return ubar, pg

    def rk3_vjp(self, dt, lam):
        """VJP of one rk3_ssp step of size dt from the current state (which
        is not advanced)."""
        lam = self._cot(lam)
        ubar, pg = np.zeros_like(lam), np.zeros(len(PARAMS))
        st = lib().rh3t_rk3_vjp(self._h, float(dt), _dptr(lam), _dptr(ubar), _dptr(pg))
        if st != OK:
            self._raise(st, "rk3_vjp")
        return ubar, pg

    def adjoint(self, checkpoint_every=5):
        return Adjoint(self, checkpoint_every)

    def profile(self, enable=True):
        _check(lib().rh3t_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        names = ctypes.create_string_buffer(1024)
        avg = (ctypes.c_double * 32)()
        cnt = (ctypes.c_int64 * 32)()
        n = ctypes.c_int32()
        _check(lib().rh3t_profile_read(self._h, 32, names, 1024, avg, cnt, ctypes.byref(n)))
        labels = names.raw.split(b"\0")[: n.value]
        return {labels[k].decode(): (avg[k], cnt[k]) for k in range(n.value)}


class Adjoint:
    """Checkpoint-and-recompute adjoint of a sequence of rk3_ssp steps with a
    frozen dt schedule (state kept every ``checkpoint_every`` steps)."""

    def __init__(self, solver, checkpoint_every=5):
        self.solver = solver
        self._h = ctypes.c_void_p()
        st = lib().rh3t_adjoint_create(solver._h, int(checkpoint_every), ctypes.byref(self._h))
        if st != OK:
            solver._raise(st, "rh3t_adjoint_create")

    def close(self):
        if getattr(self, "_h", None):
            lib().rh3t_adjoint_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def forward(self, dts):
        """len(dts) RK3 steps from the solver's state."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        st = lib().rh3t_adjoint_forward(self._h, len(dts), _dptr(dts))
        if st != OK:
            self.solver._raise(st, "adjoint forward")

    def backward(self, lam):
        """((du^N/du^0)^T lam, parameter cotangents) of the last forward."""
        lam = self.solver._cot(lam)
        out, pg = np.zeros_like(lam), np.zeros(len(PARAMS))
        st = lib().rh3t_adjoint_backward(self._h, _dptr(lam), _dptr(out), _dptr(pg))
        if st != OK:
            self.solver._raise(st, "adjoint backward")
        return out, pg

    def device_ms(self):
        """(forward, backward) device time in ms of the last forward and
        backward, without the backward's host transfers."""
        f, b = ctypes.c_double(), ctypes.c_double()
        st = lib().rh3t_adjoint_device_ms(self._h, ctypes.byref(f), ctypes.byref(b))
        if st != OK:
            self.solver._raise(st, "adjoint device_ms")
        return f.value, b.value


def _profile_read(fn, h):
    names = ctypes.create_string_buffer(1024)
    avg = (ctypes.c_double * 32)()
    cnt = (ctypes.c_int64 * 32)()
    n = ctypes.c_int32()
    _check(fn(h, 32, names, 1024, avg, cnt, ctypes.byref(n)))
    labels = names.raw.split(b"\0")[: n.value]
    return {labels[k].decode(): (avg[k], cnt[k]) for k in range(n.value)}


def cylinder_freestream(mach, gamma=1.4):
    """euler_to_three_energy of init_cylinder's freestream (src/init.rs:7-24,
    157-175): rho = 1, p = 1, u = mach * sqrt(gamma)."""
    rho, p = 1.0, 1.0
    ux = mach * np.sqrt(gamma * p / rho)
    e_total = p / (gamma - 1.0) + 0.5 * rho * (ux * ux)
    return np.array([rho, rho * ux, 0.0, e_total / 3.0, e_total / 3.0, e_total / 3.0])


class Batch:
    """Many independent solves advanced together (rh3t_batch).  Each member
    behaves as a Solver on its own Field; all members share one set of
    batched kernels and one hipGraph per step."""

    def __init__(self, fields, physics=None, device=0, arith=ARITH_EXACT):
        self.fields = list(fields)
        self.physics = physics if physics is not None else Physics()
        arr = (ctypes.c_void_p * len(self.fields))(*[f._h for f in self.fields])
        self._h = ctypes.c_void_p()
        st = lib().rh3t_batch_create(arr, len(self.fields), ctypes.byref(self.physics.c), device,
                                     ctypes.byref(self._h))
        if st != OK:
            raise RH3TError(st, "rh3t_batch_create")
        _check(lib().rh3t_batch_set_arith(self._h, arith), "batch set_arith")

    def close(self):
        if getattr(self, "_h", None):
            lib().rh3t_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        if _lib is not None:
            self.close()

    def __len__(self):
        return len(self.fields)

    def set_state(self, k, u, time=0.0):
        g = self.fields[k].grid
        u = np.ascontiguousarray(u, dtype=np.float64)
        if u.shape != (g.nx, g.ny, 6):
            raise ValueError("bad state shape")
        _check(lib().rh3t_batch_set_state(self._h, k, _dptr(u), float(time)), "batch set_state")

    def get_state(self, k):
        g = self.fields[k].grid
        u = np.zeros((g.nx, g.ny, 6))
        t = ctypes.c_double()
        _check(lib().rh3t_batch_get_state(self._h, k, _dptr(u), ctypes.byref(t)), "batch get_state")
        return u, t.value

    def advance(self, nsteps, cfl_factor=0.05):
        times = np.zeros(len(self.fields))
        st = lib().rh3t_batch_advance(self._h, int(nsteps), float(cfl_factor), _dptr(times))
        if st != OK:
            errs = []
            for k in range(len(self.fields)):
                e = c_error()
                lib().rh3t_batch_last_error(self._h, k, ctypes.byref(e))
                if e.kind:
                    errs.append((k, e.message.decode(errors="replace")))
            raise RH3TError(st, f"batch advance: {errs[:3]}", errs)
        return times

    def last_error(self, k):
        """Member k's first failure of the last call (the fields of
        rh3t_error), or None."""
        e = c_error()
        _check(lib().rh3t_batch_last_error(self._h, k, ctypes.byref(e)), "batch last_error")
        if not e.kind:
            return None
        return {"kind": e.kind, "stage": e.stage, "i": e.i, "j": e.j, "ghost_id": e.ghost_id,
                "state": list(e.state), "message": e.message.decode(errors="replace")}

    def set_launch(self, bx, by, ghosts_per_block=1, x_shape=None):
        _check(lib().rh3t_batch_set_launch(self._h, bx, by, ghosts_per_block), "batch set_launch")
        if x_shape is not None:
            _check(lib().rh3t_batch_set_launch_x(self._h, *x_shape), "batch set_launch x")

    def set_arith(self, mode):
        _check(lib().rh3t_batch_set_arith(self._h, mode), "batch set_arith")

    def profile(self, enable=True):
        _check(lib().rh3t_batch_profile(self._h, 1 if enable else 0))

    def profile_read(self):
        return _profile_read(lib().rh3t_batch_profile_read, self._h)
