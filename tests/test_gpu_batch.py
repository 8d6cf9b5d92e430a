"""Batched solves on the GPU: every member of a ragged batch must evolve
exactly as the plain solver does on its own problem (same kernels, same
arithmetic: bitwise equality), whatever the launch shape."""
import numpy as np
import pytest

import oracle_bridge as ob

pytestmark = pytest.mark.gpu


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


@pytest.fixture(scope="module")
def members(rh3t, workloads):
    box = workloads.periodic_box(48)
    c121 = rh3t.Field.init_cylinder(121, 3.0)
    c121b = c121.clone()
    c121b.set_bc_state(rh3t.BC_FARFIELD, rh3t.cylinder_freestream(2.3))
    c161 = rh3t.Field.init_cylinder(161, 3.7)
    fields = [box, c121, c161, c121b]
    states = [workloads.smooth_modes_state(box, seed=4), c121.cylinder_state(3.0), c161.cylinder_state(3.7),
              c121b.cylinder_state(2.3)]
    return fields, states


def test_batch_matches_plain_bitwise(rh3t, members):
    fields, states = members
    _say("batch of 4 (box 48, cylinders 121 x2 Mach 3 / 2.3, 161 Mach 3.7), 3 steps")
    b = rh3t.Batch(fields)
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    times = b.advance(3, 0.05)
    for k, (f, u) in enumerate(zip(fields, states)):
        s = f.solver()
        s.set_state(u, 0.0)
        t = s.advance(3, 0.05)
        ref, _ = s.get_state()
        got, tb = b.get_state(k)
        assert tb == t == times[k]
        assert np.array_equal(got, ref), (k, np.abs(got - ref).max())
        s.close()
    b.close()


@pytest.mark.parametrize("shape", [(32, 8, 2), (128, 2, 4), (16, 16, 1)])
def test_batch_launch_shape_invariance(rh3t, members, shape):
    fields, states = members
    _say(f"batch launch shape {shape}")
    ref = []
    b = rh3t.Batch(fields)
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    b.advance(2)
    ref = [b.get_state(k)[0] for k in range(len(fields))]
    b.set_launch(*shape)
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    b.advance(2)
    for k in range(len(fields)):
        assert np.array_equal(b.get_state(k)[0], ref[k])
    b.close()


def test_batch_launch_shape_limits(rh3t, members):
    """Shapes outside the kernels' limits are refused with ERR_ARG and leave
    the previous shape in place: more than 256 lanes, a width that is not a
    multiple of 16, a workgroup that is not whole wavefronts, and more heavy
    ghosts per workgroup than the ghost kernels' LDS holds."""
    fields, states = members
    _say("batch launch shape limits")
    b = rh3t.Batch(fields)
    for shape in [(256, 2, 1), (24, 8, 1), (16, 2, 1), (64, 4, 64)]:
        with pytest.raises(rh3t.RH3TError) as ei:
            b.set_launch(*shape)
        assert ei.value.status == rh3t.ERR_ARG, shape
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    b.advance(1)
    assert all(np.isfinite(b.get_state(k)[0]).all() for k in range(len(fields)))
    b.close()


def test_batch_member_against_oracle(rh3t, members):
    fields, states = members
    _say("batch member (Mach 2.3 clone) against the oracle, 1 step")
    f, u = fields[3], states[3]
    p = ob.OracleProblem(f)
    dt = 0.05 * ob.global_dt(p, u)
    b = rh3t.Batch(fields)
    for k, uu in enumerate(states):
        b.set_state(k, uu, 0.0)
    b.advance(1)
    got, t = b.get_state(3)
    ref, tref = ob.rk3_step(p, u, dt)
    assert abs(t - tref) <= 1e-15 * tref
    err = ob.relerr(got, ref, f.fluid)
    print("relerr", err)
    assert err.max() <= 1e-11
    b.close()
