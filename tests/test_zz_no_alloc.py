"""rh3t_calc_global_dt_noalloc against rh3t_calc_global_dt through the C ABI.

The no-allocation variant reads a state the caller holds in device memory
and keeps its per-workgroup minima in the library's fixed device array.
The test (the caller) allocates and fills its buffers with the HIP runtime
that librh3t.so itself loaded (libamdhip64.so.7, resolved by soname to the
instance already in the process).  Two calls in a row, the second on a
smaller input (which must not see the first call's partial minima), each
equal bit for bit to the context's own reduction of the same state; an
input above the fixed capacity is refused on the host before any launch.
"""
import ctypes

import numpy as np
import pytest

import oracle_bridge as ob

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _hip(rh3t):
    rh3t.lib()  # librh3t.so and its HIP runtime first
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    h.hipDeviceSynchronize.argtypes = []
    h.hipSetDevice.argtypes = [ctypes.c_int]
    for fn in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize, h.hipSetDevice):
        fn.restype = ctypes.c_int
    assert h.hipSetDevice(0) == 0
    return h


def _upload(h, a):
    a = np.ascontiguousarray(a)
    p = ctypes.c_void_p()
    assert h.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
    assert h.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
    return p


def _noalloc_dt(rh3t, h, field, u, phys=None):
    g = field.grid
    bufs = [_upload(h, np.asarray(u, dtype=np.float64).reshape(-1)),
            _upload(h, field.fluid.reshape(-1).astype(np.uint8)), _upload(h, np.array([-1.0]))]
    grid = rh3t.c_grid(g.nx, g.ny, g.dx, g.dy, g.x0, g.y0)
    pc = None if phys is None else ctypes.byref(phys.c)
    try:
        st = rh3t.lib().rh3t_calc_global_dt_noalloc(bufs[0], bufs[1], ctypes.byref(grid), pc, bufs[2], None)
        assert st == rh3t.OK, st
        assert h.hipDeviceSynchronize() == 0
        out = np.zeros(1)
        assert h.hipMemcpy(out.ctypes.data, bufs[2], 8, D2H) == 0
    finally:
        for b in bufs:
            h.hipFree(b)
    return float(out[0])


def test_noalloc_global_dt_matches_context(rh3t, workloads):
    h = _hip(rh3t)
    box = workloads.periodic_box(512)
    ub = workloads.smooth_modes_state(box, seed=3)
    cyl = workloads.cylinder(121, 3.0)
    uc = cyl.cylinder_state() * (1.0 + 0.1 * np.sin(np.arange(cyl.grid.ny) / 7.0))[None, :, None]
    uc[~cyl.fluid] = 0.0
    phys = rh3t.Physics(omega_ei=2.0, omega_er=0.3, kappa_e=0.01)
    for name, f, u in (("box 512", box, ub), ("cylinder 121x481", cyl, uc)):
        _say(f"noalloc dt: {name}")
        s = f.solver(physics=phys)
        s.set_state(u)
        want = s.calc_global_dt()
        s.close()
        got = _noalloc_dt(rh3t, h, f, u, phys)
        print(name, got, want)
        assert got == want
        ref = ob.global_dt(ob.OracleProblem(f), u, phys=ob.physics(omega_ei=2.0, omega_er=0.3, kappa_e=0.01))
        assert abs(got - ref) <= 1e-15 * ref
    _say("noalloc dt: capacity refusal")
    g = rh3t.c_grid(8192, 8192, 1.0, 1.0, 0.0, 0.0)
    dummy = ctypes.c_double(0.0)
    st = rh3t.lib().rh3t_calc_global_dt_noalloc(ctypes.addressof(dummy), ctypes.addressof(dummy),
                                                ctypes.byref(g), None, ctypes.addressof(dummy), None)
    assert st == rh3t.ERR_SIZE
