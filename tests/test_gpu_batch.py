"""Batched solves on the GPU: every member of a ragged batch must evolve
exactly as the plain solver does on its own problem (same kernels, same
arithmetic: bitwise equality), whatever the launch shape and in both
arithmetic variants; members against the oracle at 1e-12 (EXACT, the
default), including the full 64-solve workload of the benchmark."""
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


@pytest.mark.parametrize("arith", [0, 1], ids=["exact", "fast"])
def test_batch_matches_plain_bitwise(rh3t, members, arith):
    fields, states = members
    _say(f"batch of 4 (box 48, cylinders 121 x2 Mach 3 / 2.3, 161 Mach 3.7), 3 steps, arith={arith}")
    b = rh3t.Batch(fields, arith=arith)
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    times = b.advance(3, 0.05)
    for k, (f, u) in enumerate(zip(fields, states)):
        s = f.solver(arith=arith)
        s.set_state(u, 0.0)
        t = s.advance(3, 0.05)
        ref, _ = s.get_state()
        got, tb = b.get_state(k)
        assert tb == t == times[k]
        assert np.array_equal(got, ref), (k, np.abs(got - ref).max())
        s.close()
    b.close()


def test_batch_with_wall_ghosts_matches_plain_bitwise(rh3t, workloads):
    """A batch mixing Wall (ILW) ghosts with FarField ones -- two double Mach
    problems (h = 1/40; the second a clone, so the two share their tables)
    and a 121x481 cylinder -- evolves every member exactly as its plain
    solve (EXACT, 3 steps): the grouped Wall kernel indexes its solves,
    frames and packed points per member."""
    _say("batch of 3 (double Mach h=1/40 x2, cylinder 121), 3 steps, exact")
    dm = rh3t.Field.init_double_mach(1.0 / 40.0)
    dm2 = dm.clone()
    cyl = rh3t.Field.init_cylinder(121, 3.0)
    fields = [dm, cyl, dm2]
    u_dm = dm.double_mach_state()
    u_dm2 = u_dm + 0.01 * workloads.smooth_direction(dm, u_dm, 11)
    u_dm2[~dm.fluid] = u_dm[~dm.fluid]
    states = [u_dm, cyl.cylinder_state(3.0), u_dm2]
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


@pytest.mark.parametrize("shape", [(32, 8, 2), (128, 2, 4), (16, 16, 1), (64, 4, 1, (16, 16)),
                                   (32, 8, 1, (128, 2)), (64, 1, 2, (64, 4))])
def test_batch_launch_shape_invariance(rh3t, members, shape):
    """Launch shapes (the fused pass's, the x flux's own, heavy ghosts per
    workgroup) change no bit of the results."""
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
    for xs in [(256, 2), (24, 8), (16, 2), (0, 4)]:
        with pytest.raises(rh3t.RH3TError) as ei:
            b.set_launch(64, 4, 1, xs)
        assert ei.value.status == rh3t.ERR_ARG, xs
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
    assert err.max() <= 1e-12
    b.close()


def test_batch_failure_state(rh3t, members):
    """A negative density in the box member fails its stage 1 of the first
    step; the batch's time loop runs two more steps over the stage buffers,
    yet that member reports the failing cell's stage-1 state (the oracle's)
    and the cylinders report no failure."""
    fields, states = members
    bad = states[0].copy()
    bad[10, 20, 0] = -1.0
    _say("batch error path: negative density in member 0, advance(3)")
    b = rh3t.Batch(fields)
    for k, u in enumerate(states):
        b.set_state(k, bad if k == 0 else u, 0.0)
    with pytest.raises(rh3t.RH3TError) as ei:
        b.advance(3, 0.05)
    assert ei.value.status == rh3t.ERR_NONPHYSICAL
    p = ob.OracleProblem(fields[0])
    with pytest.raises(ob.OracleError) as eo:
        ob.rk3_step(p, bad, 0.05 * ob.global_dt(p, bad))
    g, o = b.last_error(0), eo.value.info
    print("gpu", g, "oracle", o)
    assert (g["stage"], g["i"], g["j"]) == (o["stage"], o["i"], o["j"])
    assert np.allclose(g["state"], o["state"], rtol=1e-12, atol=0.0, equal_nan=True)
    assert all(b.last_error(k) is None for k in range(1, len(fields)))
    b.close()


def test_batched_workload_64(rh3t, workloads):
    """The benchmark's batched workload itself: 64 half-cylinder solves drawn
    with seed 4 over all four grid classes (121x481 ... 241x961) and Mach
    ~ U[2, 4], EXACT arithmetic.  Three batched steps: every member bitwise
    equal to its plain solve; then one member per grid class, resynced
    every step, against the oracle's RK3 step at 1e-12 (the batch advances
    every member with its own dt = 0.05 dt_cfl)."""
    draws = workloads.batched_draw(64, seed=4)
    fields, states = workloads.batched_cylinders(draws)
    assert {nx for nx, _ in draws} == set(workloads.CYLINDER_CLASSES)
    _say("batched workload: 64 solves, 3 steps, bitwise against plain solves")
    b = rh3t.Batch(fields)
    for k, u in enumerate(states):
        b.set_state(k, u, 0.0)
    times = b.advance(3, 0.05)
    for k, (f, u) in enumerate(zip(fields, states)):
        s = f.solver()
        s.set_state(u, 0.0)
        t = s.advance(3, 0.05)
        got, tb = b.get_state(k)
        ref, _ = s.get_state()
        s.close()
        assert tb == t == times[k], k
        assert np.array_equal(got, ref), (k, draws[k], np.abs(got - ref).max())
    picks = [next(k for k, d in enumerate(draws) if d[0] == nx) for nx in workloads.CYLINDER_CLASSES]
    _say(f"batched workload: members {picks} resynced against the oracle, 3 steps")
    us = [states[k] for k in picks]
    probs = [ob.OracleProblem(fields[k]) for k in picks]
    for step in range(3):
        for k, u in zip(picks, us):
            b.set_state(k, u, 0.0)
        b.advance(1, 0.05)
        for q, k in enumerate(picks):
            dt = 0.05 * ob.global_dt(probs[q], us[q])
            ref, tref = ob.rk3_step(probs[q], us[q], dt, nthreads=16)
            got, t = b.get_state(k)
            assert abs(t - tref) <= 1e-15 * tref
            err = ob.relerr(got, ref, fields[k].fluid)
            print("step", step, "member", k, draws[k], "relerr", err.max(), flush=True)
            assert err.max() <= 1e-12, (step, k, err)
            us[q] = ref
    b.close()
