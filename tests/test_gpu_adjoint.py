"""GPU reverse mode: the HIP VJP kernels of l() and of the RK3 step, the
checkpointed adjoint of a step sequence, and the torch.autograd harness.

Checks, in increasing scope:
  * dot-product identity against the dual-number oracle (oracle/dual.cpp,
    itself pinned to central differences by test_oracle.py):
    <w, J v> (oracle, forward mode) == <J^T w, v> (GPU, reverse mode) for
    seeded state + parameter directions v and cotangents w, on periodic and
    on walled (ReflectiveWall / Constant / ZerothOrder) boxes, through the
    FarField ghosts of the cylinder, the Outflow ghosts of the translating
    shock and the Wall (ILW) ghosts of the double Mach reflection;
  * the checkpoint schedule does not change the gradient (bitwise);
  * the differentiation workload (512x512 smooth-mode box, exchange on,
    20 steps, checkpoints every 5, J = sum of a Gaussian window times the
    final E_r): torch.autograd gradient against central differences of the
    GPU forward (h = 1e-6) on 32 seeded directions over (u0, omega_ei,
    omega_er), relative agreement <= 1e-6.
Every case prints its name before touching the GPU.
"""
import numpy as np
import pytest

import oracle_bridge as ob

pytestmark = pytest.mark.gpu

DOT_TOL = 1e-10
KW_ACTIVE = dict(omega_ei=1.0, omega_er=0.1, kappa_e=0.01, kappa_i=0.02, kappa_r=0.005)
PSCALE = np.array([1.0, 0.1, 0.01, 0.02, 0.005])


def _say(msg):
    print(f"[gpu-case] {msg}", flush=True)


def _dot_check(lhs, terms, label):
    """lhs = <w, J v>; terms: the elementwise products summed by the
    reverse side (their magnitude sets the scale)."""
    rhs = sum(float(np.sum(t)) for t in terms)
    scale = sum(float(np.sum(np.abs(t))) for t in terms)
    rel = abs(lhs - rhs) / scale
    print(f"{label}: <w,Jv>={lhs:.15e} <J^T w,v>={rhs:.15e} rel={rel:.3e}")
    assert rel <= DOT_TOL


def _case(workloads, kind, n):
    f = workloads.periodic_box(n) if kind == "periodic" else workloads.walled_box(n)
    u = workloads.smooth_modes_state(f, seed=3)
    return f, u, ob.OracleProblem(f)


@pytest.mark.parametrize("kind", ["periodic", "walled"])
@pytest.mark.parametrize("kw", [KW_ACTIVE, {}], ids=["active", "default"])
def test_l_vjp_dot_product(rh3t, workloads, kind, kw):
    f, u, p = _case(workloads, kind, 32)
    _say(f"l_vjp dot product {kind} 32 physics={sorted(kw)}")
    phys = rh3t.Physics(**kw)
    s = f.solver(physics=phys)
    s.set_state(u)
    rng = np.random.default_rng(21)
    w = rng.standard_normal(u.shape)
    w[~f.fluid] = 0.0
    du = workloads.state_direction(u, f.fluid, 22)
    dpar = rng.standard_normal(5) * PSCALE
    ubar, pg = s.l_vjp(w)
    _, jv = ob.l_jvp(p, u, du, ob.physics(**kw), dpar)
    assert np.all(ubar[~f.fluid] == 0.0)
    _dot_check(float(np.sum(w * jv)), [ubar * du, pg * dpar], f"l_vjp {kind}")
    s.close()


def test_rk3_vjp_dot_product(rh3t, workloads):
    f, u, p = _case(workloads, "periodic", 32)
    _say("rk3_vjp dot product periodic 32")
    s = f.solver(physics=rh3t.Physics(**KW_ACTIVE))
    s.set_state(u)
    dt = 0.05 * s.calc_global_dt()
    rng = np.random.default_rng(31)
    w = rng.standard_normal(u.shape)
    du = workloads.state_direction(u, f.fluid, 32)
    dpar = rng.standard_normal(5) * PSCALE
    ubar, pg = s.rk3_vjp(dt, w)
    got, _ = s.get_state()
    assert np.array_equal(got, u), "rk3_vjp leaves the state alone"
    _, jv = ob.rk3_jvp(p, u, du, [dt], ob.physics(**KW_ACTIVE), dpar)
    _dot_check(float(np.sum(w * jv)), [ubar * du, pg * dpar], "rk3_vjp")
    s.close()


def _dts(solver, u, nsteps):
    """CFL schedule dt_n = 0.05 dt_cfl(u^n) of a base run (then frozen)."""
    solver.set_state(u)
    dts = []
    for _ in range(nsteps):
        dt = 0.05 * solver.calc_global_dt()
        solver.rk3_ssp(dt)
        dts.append(dt)
    return dts


def test_adjoint_checkpoint_schedule(rh3t, workloads):
    """12 steps on a 48x48 box: the gradient does not depend on the
    checkpoint interval (bitwise) nor on what an earlier backward left on the
    tape, and satisfies the dot-product identity against the oracle's
    forward mode through all 12 steps."""
    f, u, p = _case(workloads, "periodic", 48)
    _say("adjoint checkpoint schedule periodic 48, 12 steps")
    s = f.solver(physics=rh3t.Physics(**KW_ACTIVE))
    dts = _dts(s, u, 12)
    rng = np.random.default_rng(41)
    w = rng.standard_normal(u.shape)
    results = []
    for every in (1, 5, 12):
        s.set_state(u)
        adj = s.adjoint(every)
        adj.forward(dts)
        un, _ = s.get_state()
        lam0, pg = adj.backward(w)
        after, _ = s.get_state()
        assert np.array_equal(after, un), "backward leaves the context at u^N"
        # a second backward finds other segments on the tape and recomputes
        # the last ones too
        lam0b, pgb = adj.backward(w)
        assert np.array_equal(lam0b, lam0) and np.array_equal(pgb, pg)
        results.append((lam0, pg))
        adj.close()
    for lam0, pg in results[1:]:
        assert np.array_equal(lam0, results[0][0]) and np.array_equal(pg, results[0][1])
    du = workloads.state_direction(u, f.fluid, 42)
    dpar = rng.standard_normal(5) * PSCALE
    un_o, jv = ob.rk3_jvp(p, u, du, dts, ob.physics(**KW_ACTIVE), dpar)
    print("forward vs oracle", ob.relerr(un, un_o).max())
    assert ob.relerr(un, un_o).max() <= 1e-11
    lam0, pg = results[0]
    _dot_check(float(np.sum(w * jv)), [lam0 * du, pg * dpar], "adjoint 12 steps")
    s.close()


def test_interleaved_rollouts_keep_their_physics(rh3t, workloads):
    """Two torch rollouts with different exchange coefficients on one
    solver, the second's forward before the first's backward: the first
    gradient equals the gradient of the first rollout run alone (bitwise),
    i.e. each backward differentiates the forward it belongs to."""
    import torch

    from importlib import import_module

    ta = import_module("3t_rh_amd.torch_adjoint")
    _say("interleaved torch rollouts, box 32, 6 steps")
    f, u, _p = _case(workloads, "periodic", 32)
    phys = rh3t.Physics(**KW_ACTIVE)
    s = f.solver(physics=phys)
    dts = _dts(s, u, 6)
    w = torch.from_numpy(workloads.gaussian_window(f, seed=5))
    pa = phys.params()
    pb = pa * np.array([3.0, 0.2, 2.0, 0.5, 1.5])

    def run(pv):
        u0 = torch.tensor(u, requires_grad=True)
        params = torch.tensor(pv, requires_grad=True)
        un = ta.rollout(s, u0, params, dts, checkpoint_every=2)
        return u0, params, (w * un[..., 5]).sum()

    u0a, qa, la = run(pa)
    la.backward()
    solo = (u0a.grad.numpy().copy(), qa.grad.numpy().copy())
    u0a, qa, la = run(pa)
    u0b, qb, lb = run(pb)
    la.backward()
    lb.backward()
    assert np.array_equal(u0a.grad.numpy(), solo[0]) and np.array_equal(qa.grad.numpy(), solo[1])
    assert not np.array_equal(u0b.grad.numpy(), solo[0])
    # the solver keeps the constants of its last set_physics
    assert s.physics.params().tolist() == pb.tolist()
    s.close()


def _heavy_case(workloads, which):
    """FarField: the Mach-3 cylinder 121x481 (4,480 FarField ghosts), its
    freestream perturbed by 1% of a smooth direction so that the WENO
    weights and the characteristic projections vary along the boundary.
    Outflow: the translating Mach-2 shock on 64x64 (LODI relaxation active:
    subsonic flow at the right side), perturbed the same way.  The
    perturbation also keeps the max / min of the flux splitting and of the
    noncon limiter off ties, where the dual oracle and the reverse mode may
    take the derivatives of different (equal-valued) branches."""
    if which == "farfield":
        f = workloads.cylinder(121, 3.0)
        u0 = f.cylinder_state()
    else:
        f = workloads.translating_shock(64)
        u0 = workloads.translating_shock_state(f)
    u = u0 + 0.01 * workloads.smooth_direction(f, u0, 7)
    u[~f.fluid] = u0[~f.fluid]
    return f, u, ob.OracleProblem(f)


@pytest.mark.parametrize("which", ["farfield", "outflow"])
def test_l_vjp_dot_product_heavy_ghosts(rh3t, workloads, which):
    """The VJP of l() through the FarField / Outflow ghosts (their WENO
    extrapolation, characteristic BC branch and Taylor sum): dot-product
    identity against the dual-number oracle, which differentiates the same
    ghost code as the forward oracle."""
    f, u, p = _heavy_case(workloads, which)
    _say(f"l_vjp dot product through {which} ghosts {f.grid.nx}x{f.grid.ny}")
    s = f.solver(physics=rh3t.Physics(**KW_ACTIVE))
    s.set_state(u)
    rng = np.random.default_rng(51)
    w = rng.standard_normal(u.shape)
    w[~f.fluid] = 0.0
    du = workloads.state_direction(u, f.fluid, 52)
    dpar = rng.standard_normal(5) * PSCALE
    ubar, pg = s.l_vjp(w)
    _, jv = ob.l_jvp(p, u, du, ob.physics(**KW_ACTIVE), dpar)
    assert np.all(ubar[~f.fluid] == 0.0)
    _dot_check(float(np.sum(w * jv)), [ubar * du, pg * dpar], f"l_vjp {which}")
    s.close()


def test_rk3_vjp_dot_product_cylinder(rh3t, workloads):
    """One RK3 step of the perturbed cylinder: the three stage VJPs, each
    through the FarField ghosts of its stage state."""
    f, u, p = _heavy_case(workloads, "farfield")
    _say("rk3_vjp dot product cylinder 121x481")
    s = f.solver()
    s.set_state(u)
    dt = 0.05 * s.calc_global_dt()
    rng = np.random.default_rng(61)
    w = rng.standard_normal(u.shape)
    w[~f.fluid] = 0.0
    du = workloads.state_direction(u, f.fluid, 62)
    dpar = rng.standard_normal(5) * PSCALE
    ubar, pg = s.rk3_vjp(dt, w)
    _, jv = ob.rk3_jvp(p, u, du, [dt], ob.physics(), dpar)
    _dot_check(float(np.sum(w * jv)), [ubar * du, pg * dpar], "rk3_vjp cylinder")
    s.close()


def test_l_vjp_dot_product_wall_ghosts(rh3t, workloads):
    """The VJP of l() through the Wall (ILW) ghosts of the double Mach
    reflection at h = 1/40 (its two ILW linear systems per ghost), with the
    Constant / ZerothOrder sides around them; initial state perturbed by 1%
    of a smooth direction (off the ties of the piecewise-constant state)."""
    f = rh3t.Field.init_double_mach(1.0 / 40.0)
    u0 = f.double_mach_state()
    u = u0 + 0.01 * workloads.smooth_direction(f, u0, 9)
    u[~f.fluid] = u0[~f.fluid]
    p = ob.OracleProblem(f)
    _say(f"l_vjp dot product through wall ghosts {f.grid.nx}x{f.grid.ny}")
    s = f.solver(physics=rh3t.Physics(**KW_ACTIVE))
    s.set_state(u)
    rng = np.random.default_rng(71)
    w = rng.standard_normal(u.shape)
    w[~f.fluid] = 0.0
    du = workloads.state_direction(u, f.fluid, 72)
    dpar = rng.standard_normal(5) * PSCALE
    ubar, pg = s.l_vjp(w)
    _, jv = ob.l_jvp(p, u, du, ob.physics(**KW_ACTIVE), dpar)
    _dot_check(float(np.sum(w * jv)), [ubar * du, pg * dpar], "l_vjp wall")
    s.close()


def test_time_dependent_ghosts_are_rejected(rh3t, workloads):
    """The reverse mode does not cover TimeDependent ghosts: the double Mach
    problem with the moving-shock top must fail loudly rather than return a
    partial gradient."""
    _say("adjoint on the double Mach reflection with a moving-shock top: unsupported")
    f = rh3t.Field.init_double_mach(1.0 / 40.0, top=workloads.double_mach_moving_shock())
    s = f.solver()
    s.set_state(f.double_mach_state())
    with pytest.raises(rh3t.RH3TError) as ei:
        s.l_vjp(np.zeros((f.grid.nx, f.grid.ny, 6)))
    assert ei.value.status == rh3t.ERR_UNSUPPORTED
    s.close()


def test_adjoint_device_timing(rh3t, workloads):
    """rh3t_adjoint_device_ms is refused until a forward and a backward have
    run, then gives positive device times within the wall clock around each
    call (the backward's host transfers fall outside its events)."""
    import time

    _say("adjoint device timing, box 64, 6 steps")
    f, u, _p = _case(workloads, "periodic", 64)
    s = f.solver(physics=rh3t.Physics(**KW_ACTIVE))
    dts = _dts(s, u, 6)
    adj = s.adjoint(3)
    with pytest.raises(rh3t.RH3TError) as ei:
        adj.device_ms()
    assert ei.value.status == rh3t.ERR_ARG
    s.set_state(u)
    t0 = time.perf_counter()
    adj.forward(dts)
    wall_f = (time.perf_counter() - t0) * 1e3
    lam = np.random.default_rng(4).standard_normal(u.shape)
    lam[~f.fluid] = 0.0
    t0 = time.perf_counter()
    adj.backward(lam)
    wall_b = (time.perf_counter() - t0) * 1e3
    dev_f, dev_b = adj.device_ms()
    print(f"forward device {dev_f:.3f} ms (wall {wall_f:.3f}), backward device {dev_b:.3f} ms (wall {wall_b:.3f})")
    assert 0.0 < dev_f <= wall_f and 0.0 < dev_b <= wall_b
    adj.close()
    s.close()


def test_autograd_matches_central_differences(rh3t, workloads):
    """The differentiation workload through torch.autograd, against central
    differences of the GPU forward on 32 seeded directions (seeds
    5001..5032) over (u0, omega_ei, omega_er).  The state directions are
    smooth (random Fourier modes, like the initial state): a cell-wise white
    noise direction changes the local variation that the WENO weights
    (eps + beta)^-2 respond to nonlinearly at the 1e-6 scale, so its central
    difference carries a truncation error of ~1e-6..1e-5 while the
    dot-product checks above pin the reverse mode itself to rounding."""
    import torch

    from importlib import import_module

    ta = import_module("3t_rh_amd.torch_adjoint")
    n, nsteps = 512, 20
    _say(f"autograd vs FD: box {n}, {nsteps} steps, 32 directions")
    f = workloads.periodic_box(n)
    u = workloads.smooth_modes_state(f, seed=3)
    phys = rh3t.Physics(**workloads.ADJOINT_PHYSICS)
    s = f.solver(physics=phys)
    dts = _dts(s, u, nsteps)
    w = workloads.gaussian_window(f, seed=5)
    u0 = torch.tensor(u, requires_grad=True)
    params = torch.tensor(phys.params(), requires_grad=True)
    un = ta.rollout(s, u0, params, dts, checkpoint_every=5)
    loss = (torch.from_numpy(w) * un[..., 5]).sum()
    loss.backward()
    gu, gp = u0.grad.numpy(), params.grad.numpy()

    def J(x, pv):
        s.set_physics(phys.with_params(pv))
        s.set_state(x)
        for d in dts:
            s.rk3_ssp(d)
        y, _ = s.get_state()
        return float(np.sum(w * y[..., 5]))

    h = 1e-6
    worst = 0.0
    p0 = phys.params()
    for seed in range(5001, 5033):
        rng = np.random.default_rng(seed)
        du = workloads.smooth_direction(f, u, seed)
        dp = np.zeros(5)
        dp[:2] = rng.standard_normal(2) * PSCALE[:2]
        fd = (J(u + h * du, p0 + h * dp) - J(u - h * du, p0 - h * dp)) / (2 * h)
        ad = float(np.sum(gu * du) + np.sum(gp * dp))
        rel = abs(fd - ad) / abs(ad)
        worst = max(worst, rel)
        print(f"direction {seed}: ad={ad:.12e} fd={fd:.12e} rel={rel:.2e}", flush=True)
    print("worst relative FD disagreement", worst)
    assert worst <= 1e-6
    s.close()
