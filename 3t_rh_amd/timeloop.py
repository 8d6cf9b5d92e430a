"""Time-loop driver and RH3TBIN1 snapshots around the device solver.

Mirrors ``main()`` of colinhu123/3T_RH (src/main.rs:410-487) and its
snapshot I/O (src/io.rs:25-124):

* :func:`save_data` / :func:`load_data`: the RH3TBIN1 format, little endian,
  header ``"RH3TBIN1"``, u32 version 1, u32 nvar 8, u64 nx, u64 ny, f64 time,
  then nx*ny records of 8 f64 ``(x, y, rho, mom_x, mom_y, ee, ei, er)`` in
  j-major / i-fastest order, NaN for the state of points outside the fluid;
  written to ``<name>.tmp`` and renamed into place.
* :func:`run`: the output-time loop.  Every step takes dt = 0.05 dt_cfl
  from the device CFL reduction, clipped so that the step lands on the next
  output time and on t_final, then stores ``solution_NNNN.bin`` when an
  output time is reached; a restart continues from ``solution_<id>.bin``.

The state stays on the device between steps; the host reads one scalar (the
CFL dt) per step and the whole state only at output times.
"""
import math
import os
import shutil
import struct

import numpy as np

MAGIC = b"RH3TBIN1"
VERSION = 1
NVAR = 8
_HEADER = struct.Struct("<8sIIQQd")


class SnapshotError(ValueError):
    pass


def _coords(grid):
    x = grid.x0 + np.arange(grid.nx, dtype=np.float64) * grid.dx
    y = grid.y0 + np.arange(grid.ny, dtype=np.float64) * grid.dy
    return x, y


def save_data(field, u, time, path):
    """Writes state u (nx, ny, 6) of `field` at `time` to `path` (RH3TBIN1,
    src/io.rs:25-92)."""
    g = field.grid
    u = np.asarray(u, dtype=np.float64)
    if u.shape != (g.nx, g.ny, 6):
        raise ValueError(f"state must have shape {(g.nx, g.ny, 6)}")
    rec = np.empty((g.ny, g.nx, NVAR), dtype="<f8")
    x, y = _coords(g)
    rec[:, :, 0] = x[None, :]
    rec[:, :, 1] = y[:, None]
    rec[:, :, 2:] = np.transpose(u, (1, 0, 2))
    rec[~field.fluid.T, 2:] = np.nan
    tmp = path + ".tmp"
    with open(tmp, "wb") as fh:
        fh.write(_HEADER.pack(MAGIC, VERSION, NVAR, g.nx, g.ny, float(time)))
        rec.tofile(fh)
    os.replace(tmp, path)


def load_data(field, path, u=None):
    """Reads a RH3TBIN1 snapshot of `field`'s grid (src/io.rs:104-124).
    Returns (state, time): the fluid cells from the file, the other cells
    from `u` (zeros if None).  Raises SnapshotError on a foreign file, a
    grid mismatch or a non-finite fluid state."""
    g = field.grid
    with open(path, "rb") as fh:
        head = fh.read(_HEADER.size)
        if len(head) != _HEADER.size:
            raise SnapshotError(f"{path}: truncated header")
        magic, version, nvar, nx, ny, time = _HEADER.unpack(head)
        if magic != MAGIC or version != VERSION or nvar != NVAR:
            raise SnapshotError(f"{path}: not a RH3TBIN1 v{VERSION} file with {NVAR} variables")
        if (nx, ny) != (g.nx, g.ny):
            raise SnapshotError(f"{path}: restart grid mismatch ({nx}, {ny}) vs ({g.nx}, {g.ny})")
        rec = np.fromfile(fh, dtype="<f8", count=nx * ny * NVAR)
    if rec.size != nx * ny * NVAR:
        raise SnapshotError(f"{path}: truncated payload")
    rec = rec.reshape(ny, nx, NVAR)
    out = np.zeros((g.nx, g.ny, 6)) if u is None else np.array(u, dtype=np.float64, copy=True)
    fluid = field.fluid
    vals = np.transpose(rec[:, :, 2:], (1, 0, 2))
    bad = fluid & ~np.all(np.isfinite(vals), axis=2)
    if bad.any():
        i, j = np.argwhere(bad)[0]
        raise SnapshotError(f"{path}: non-finite restart state at ({i}, {j})")
    out[fluid] = vals[fluid]
    return out, float(time)


def clear_data_folder(out_dir):
    """src/io.rs:94-102."""
    if os.path.exists(out_dir):
        shutil.rmtree(out_dir)
    os.makedirs(out_dir, exist_ok=True)


def _round_half_away(x):
    # f64::round: halves away from zero (Python's round() rounds to even)
    r = math.floor(abs(x))
    if abs(x) - r >= 0.5:
        r += 1
    return int(math.copysign(r, x))


def run(solver, t_final=1.0, t_store_interval=0.001, out_dir="data_new", restart_id=None, cfl_factor=0.05,
        log=print):
    """The output-time loop of src/main.rs:410-487 on `solver` (a Solver whose
    state is the initial state).  A fresh run clears `out_dir` and stores
    solution_0000.bin; restart_id continues from solution_<id>.bin.  Returns
    {"t", "steps", "last_id"}."""
    field = solver.field
    os.makedirs(out_dir, exist_ok=True)
    name = lambda k: os.path.join(out_dir, f"solution_{k:04d}.bin")  # noqa: E731
    if restart_id is None:
        clear_data_folder(out_dir)
        u, t = solver.get_state()
        store_id = 0
    else:
        u0, _ = solver.get_state()
        u, t = load_data(field, name(restart_id), u0)
        solver.set_state(u, t)
        store_id = _round_half_away(t / t_store_interval)
    next_store = (store_id + 1) * t_store_interval
    n = 0
    if restart_id is None:
        save_data(field, u, t, name(0))
        log(f"stored solution_0000.bin at t = {t:.8e}")
    else:
        log(f"Restarting from id={store_id}, t={t:.8e}; next output t={next_store:.8e}")
    while t < t_final - 1e-14:
        dt_cfl = solver.calc_global_dt()
        dt = cfl_factor * dt_cfl
        if next_store <= t_final and t + dt > next_store:
            dt = next_store - t
        if t + dt > t_final:
            dt = t_final - t
        if not dt > 0.0:
            raise RuntimeError(f"non-positive dt at t={t}")
        solver.rk3_ssp(dt)
        t = t + dt  # Field.time advances the same way on the device
        n += 1
        log(f"step={n}, t={t:.8e}, dt={dt:.8e}, dt_cfl={dt_cfl:.8e}")
        if next_store <= t_final and t >= next_store - 1e-12:
            store_id += 1
            u, _ = solver.get_state()
            save_data(field, u, t, name(store_id))
            log(f"stored solution_{store_id:04d}.bin at t={t:.8e}")
            next_store = (store_id + 1) * t_store_interval
    last_regular = store_id * t_store_interval
    if abs(t - last_regular) > 1e-12:
        store_id += 1
        u, _ = solver.get_state()
        save_data(field, u, t, name(store_id))
        log(f"stored final solution_{store_id:04d}.bin at t={t:.8e}")
    log(f"Finished: t={t:.8e}, restart-local steps={n}, last id={store_id}")
    return {"t": t, "steps": n, "last_id": store_id}
