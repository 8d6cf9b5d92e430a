"""Time-loop driver and RH3TBIN1 snapshots (src/main.rs:410-487,
src/io.rs:25-124), on the CPU: the byte layout of the format, the round
trip, the loader's refusals, and the driver's dt clipping, output ids and
restart against a stub solver whose steps are plain host arithmetic."""
import math
import os
import struct
from importlib import import_module

import numpy as np
import pytest

tl = import_module("3t_rh_amd.timeloop")


def _state(field, seed):
    g = field.grid
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.5, 1.5, size=(g.nx, g.ny, 6))
    u[~field.fluid] = -7.0  # never written to a snapshot
    return u


def test_snapshot_byte_layout(rh3t, tmp_path):
    """Header fields and one record at a known offset, against struct."""
    f = rh3t.Field.init_cylinder(121, 3.0)
    g = f.grid
    u = _state(f, 1)
    p = str(tmp_path / "s.bin")
    tl.save_data(f, u, 0.125, p)
    raw = open(p, "rb").read()
    assert raw[:8] == b"RH3TBIN1"
    assert struct.unpack("<IIQQd", raw[8:40]) == (1, 8, g.nx, g.ny, 0.125)
    assert len(raw) == 40 + g.nx * g.ny * 8 * 8
    fluid = f.fluid
    i, j = map(int, np.argwhere(fluid)[17])
    off = 40 + (j * g.nx + i) * 64  # j-major, i fastest
    rec = struct.unpack("<8d", raw[off:off + 64])
    assert rec[0] == g.x0 + i * g.dx and rec[1] == g.y0 + j * g.dy
    assert rec[2:] == tuple(u[i, j])
    i, j = map(int, np.argwhere(~fluid)[3])
    off = 40 + (j * g.nx + i) * 64
    rec = struct.unpack("<8d", raw[off:off + 64])
    assert rec[0] == g.x0 + i * g.dx and all(math.isnan(v) for v in rec[2:])
    assert not os.path.exists(p + ".tmp")


def test_snapshot_round_trip(rh3t, tmp_path):
    f = rh3t.Field.init_cylinder(121, 3.0)
    u = _state(f, 2)
    p = str(tmp_path / "s.bin")
    tl.save_data(f, u, 3.5e-3, p)
    base = np.full_like(u, 9.0)
    v, t = tl.load_data(f, p, base)
    assert t == 3.5e-3
    assert np.array_equal(v[f.fluid], u[f.fluid])
    assert np.all(v[~f.fluid] == 9.0), "points outside the fluid keep the caller's values"


def test_snapshot_refusals(rh3t, workloads, tmp_path):
    f = rh3t.Field.init_cylinder(121, 3.0)
    u = _state(f, 3)
    p = str(tmp_path / "s.bin")
    tl.save_data(f, u, 0.0, p)
    with pytest.raises(tl.SnapshotError, match="grid mismatch"):
        tl.load_data(workloads.periodic_box(16), p)
    raw = bytearray(open(p, "rb").read())
    bad = str(tmp_path / "bad.bin")
    open(bad, "wb").write(b"RH3TBIN2" + raw[8:])
    with pytest.raises(tl.SnapshotError, match="not a RH3TBIN1"):
        tl.load_data(f, bad)
    i, j = map(int, np.argwhere(f.fluid)[5])
    off = 40 + (j * f.grid.nx + i) * 64 + 3 * 8
    raw[off:off + 8] = struct.pack("<d", float("inf"))
    open(bad, "wb").write(bytes(raw))
    with pytest.raises(tl.SnapshotError, match=f"non-finite restart state at \\({i}, {j}\\)"):
        tl.load_data(f, bad)
    open(bad, "wb").write(bytes(raw[:1000]))
    with pytest.raises(tl.SnapshotError, match="truncated"):
        tl.load_data(f, bad)


class StubSolver:
    """Host stand-in with the Solver calls the driver makes: the CFL dt is a
    fixed function of the step count, a step adds dt to the density."""

    def __init__(self, field, u, t=0.0):
        self.field, self.u, self.t, self.dts = field, u.copy(), t, []

    def get_state(self):
        return self.u.copy(), self.t

    def set_state(self, u, t=0.0):
        self.u, self.t = u.copy(), t

    def calc_global_dt(self):
        return 0.08 + 0.01 * (len(self.dts) % 3)  # dt = 0.05 x (0.08, 0.09, 0.10)

    def rk3_ssp(self, dt):
        self.dts.append(dt)
        self.u[..., 0] += dt
        self.t = self.t + dt


def test_driver_clips_and_stores(workloads, tmp_path):
    f = workloads.periodic_box(8)
    u0 = workloads.smooth_modes_state(f, seed=3)
    s = StubSolver(f, u0)
    out = str(tmp_path / "run")
    res = tl.run(s, t_final=0.023, t_store_interval=0.01, out_dir=out, log=lambda m: None)
    # the dt rule, replayed: 0.05 dt_cfl clipped to the next output time
    t, nxt, expect = 0.0, 0.01, []
    while t < 0.023 - 1e-14:
        dt = 0.05 * (0.08 + 0.01 * (len(expect) % 3))
        if nxt <= 0.023 and t + dt > nxt:
            dt = nxt - t
        if t + dt > 0.023:
            dt = 0.023 - t
        expect.append(dt)
        t = t + dt
        if t >= nxt - 1e-12:
            nxt += 0.01
    assert s.dts == expect
    assert res["steps"] == len(expect) and res["t"] == t
    names = sorted(os.listdir(out))
    assert names == ["solution_0000.bin", "solution_0001.bin", "solution_0002.bin", "solution_0003.bin"]
    times = [tl.load_data(f, os.path.join(out, n))[1] for n in names]
    assert abs(times[1] - 0.01) < 1e-12 and abs(times[2] - 0.02) < 1e-12 and times[3] == res["t"]
    v, _ = tl.load_data(f, os.path.join(out, names[-1]))
    assert np.array_equal(v, s.u)


def test_driver_restart_continues(workloads, tmp_path):
    """A restart from solution_0001 reaches the same final state and time as
    the uninterrupted run, storing the same later snapshots."""
    f = workloads.periodic_box(8)
    u0 = workloads.smooth_modes_state(f, seed=3)
    out = str(tmp_path / "run")
    full = StubSolver(f, u0)
    tl.run(full, t_final=0.023, t_store_interval=0.01, out_dir=out, log=lambda m: None)
    final_full = open(os.path.join(out, "solution_0003.bin"), "rb").read()
    os.remove(os.path.join(out, "solution_0003.bin"))
    os.remove(os.path.join(out, "solution_0002.bin"))
    again = StubSolver(f, u0)
    # the stub's dt pattern depends on its step count: start it where the
    # uninterrupted run was at t = 0.01
    n1 = next(k for k in range(len(full.dts)) if sum(full.dts[:k + 1]) >= 0.01 - 1e-12) + 1
    again.dts = [0.0] * n1
    res = tl.run(again, t_final=0.023, t_store_interval=0.01, out_dir=out, restart_id=1, log=lambda m: None)
    assert res["last_id"] == 3
    assert open(os.path.join(out, "solution_0003.bin"), "rb").read() == final_full
