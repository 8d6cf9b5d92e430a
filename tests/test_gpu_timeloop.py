"""The time-loop driver on the device solver: a 32x32 smooth-mode periodic
box run to t_final through two output times.  The snapshots hold the device
state at their times, the dt sequence is the CFL rule of the oracle's state
clipped to the output times, and the final state agrees with the oracle
stepped with the same dts (1e-11 relative, max-norm per field)."""
import os
from importlib import import_module

import numpy as np
import pytest

import oracle_bridge as ob

pytestmark = pytest.mark.gpu


class Recorder:
    """Passes the driver's calls through to the solver, keeping the dts."""

    def __init__(self, solver):
        self.s, self.field, self.dts = solver, solver.field, []

    def get_state(self):
        return self.s.get_state()

    def set_state(self, u, t=0.0):
        self.s.set_state(u, t)

    def calc_global_dt(self):
        return self.s.calc_global_dt()

    def rk3_ssp(self, dt):
        self.dts.append(dt)
        self.s.rk3_ssp(dt)


def test_driver_on_device(rh3t, workloads, tmp_path):
    tl = import_module("3t_rh_amd.timeloop")
    print("[gpu-case] time loop, periodic box 32, two outputs", flush=True)
    f = workloads.periodic_box(32)
    u0 = workloads.smooth_modes_state(f, seed=3)
    s = f.solver()
    s.set_state(u0, 0.0)
    dt_cfl0 = s.calc_global_dt()
    interval = 4.5 * 0.05 * dt_cfl0  # an output every ~4.5 steps
    rec = Recorder(s)
    out = str(tmp_path / "run")
    res = tl.run(rec, t_final=2.2 * interval, t_store_interval=interval, out_dir=out, log=lambda m: None)
    assert sorted(os.listdir(out)) == ["solution_0000.bin", "solution_0001.bin", "solution_0002.bin",
                                       "solution_0003.bin"]
    un, tn = s.get_state()
    assert tn == res["t"]
    v, tv = tl.load_data(f, os.path.join(out, "solution_0003.bin"))
    assert tv == tn and np.array_equal(v, un)
    # the oracle with the recorded dts; each dt is the CFL rule of the
    # oracle's own state (clipped at output times)
    p = ob.OracleProblem(f)
    u, t = u0.copy(), 0.0
    nxt = interval
    for k, dt in enumerate(rec.dts):
        want = 0.05 * ob.global_dt(p, u)
        if t + want > nxt:
            want = nxt - t
        if t + want > 2.2 * interval:
            want = 2.2 * interval - t
        assert abs(dt - want) <= 1e-11 * want, (k, dt, want)
        u, _ = ob.rk3_step(p, u, dt)
        t = t + dt
        if t >= nxt - 1e-12:
            nxt += interval
    err = ob.relerr(un, u).max()
    print("final state vs oracle", err)
    assert err <= 1e-11
    # the readers on the run's own output: the energy split of the last
    # snapshot is that of the device state, and on a periodic box the
    # totals of rho, the momenta and ee + ei + er stay put (the flux
    # differences telescope, the noncon and exchange terms sum to zero)
    pp = import_module("3t_rh_amd.postprocess")
    paths = pp.list_snapshots(out)
    rows = pp.energy_split_history(paths, str(tmp_path / "split.csv"))
    assert [r[0] for r in rows] == [pp.read_snapshot(p).time for p in paths]
    d = np.abs(un[..., 3] - un[..., 5])
    assert rows[-1][2] == d.max()
    assert abs(rows[-1][5] - d.mean()) <= 1e-14 * d.mean()
    tots = np.array([pp.totals(p) for p in paths])
    scale = np.abs(un).sum(axis=(0, 1))
    cons = np.column_stack([tots[:, :3], tots[:, 3:].sum(axis=1)])
    drift = np.abs(cons - cons[0]).max(axis=0) / np.array([*scale[:3], scale[3:].sum()])
    print("conservation drift (rho, mx, my, E)", drift)
    assert np.all(drift <= 1e-13)
    s.close()
