"""Readers of RH3TBIN1 snapshots: the data side of colinhu123/3T_RH's
post-processing scripts, on the files :func:`timeloop.run` writes.

* :func:`read_snapshot` and :func:`list_snapshots`: the checked
  memory-mapped reader of visualize_sol.py:36-96 / video_gen.py:68-125
  (header, magic, version, nvar and byte-size checks, (ny, nx, 8) records,
  NaN outside the fluid) and the ``solution_*.bin`` listing, which never
  sees a ``.bin.tmp`` file still being written.
* :func:`density_frame` and :func:`density_range`: the masked density of
  one frame (visualize_sol.py:80-95) and the one colour range of a whole
  sequence (video_gen.py:150-181).
* :func:`energy_split_row` and :func:`energy_split_history`: the
  three-energy symmetry diagnostic of py_utils/energy_split.py:12-35
  (max / mean / rms of |Ee-Ei|, |Ee-Er|, |Ei-Er| over the valid cells and
  the location of max |Ee-Er|), with its CSV.
* :func:`primitives`: velocities and the three pressures of
  py_utils/contour_gen.py:112-187 (the kinetic share rho (u^2+v^2)/6 of
  src/state.rs:245-256), with the adiabatic indices of the run's Physics
  (1.4, 1.4, 1.4 by default, src/constant.rs:11-13).
* :func:`totals`: the sums of the six conserved fields over the valid
  cells, the quantity py_utils/conservation_check.py:14-31 follows per
  output.

The plotting and movie writing of those scripts (matplotlib, ffmpeg) are
not part of this module.
"""
import glob
import os
import struct
from collections import namedtuple

import numpy as np

MAGIC = b"RH3TBIN1"
VERSION = 1
NVAR = 8
HEADER = struct.Struct("<8sIIQQd")
COLUMNS = ("x", "y", "rho", "mom_x", "mom_y", "ee", "ei", "er")
ENERGY_SPLIT_HEADER = ("time,max_Ee_Ei,max_Ee_Er,max_Ei_Er,mean_Ee_Ei,mean_Ee_Er,mean_Ei_Er,"
                       "rms_Ee_Ei,rms_Ee_Er,rms_Ei_Er,x_max_Ee_Er,y_max_Ee_Er")


class SnapshotFormatError(RuntimeError):
    """A file that is not a complete RH3TBIN1 v1 snapshot."""


Snapshot = namedtuple("Snapshot", "path nx ny time data")
Snapshot.__doc__ = """One snapshot: `data` is a read-only (ny, nx, 8) float64
memory map of the records (x, y, rho, mom_x, mom_y, ee, ei, er), j-major."""


def read_snapshot(path):
    """The checked reader of visualize_sol.py:36-79."""
    with open(path, "rb") as fh:
        raw = fh.read(HEADER.size)
    if len(raw) != HEADER.size:
        raise SnapshotFormatError(f"Incomplete binary header in {path}")
    magic, version, nvar, nx, ny, time = HEADER.unpack(raw)
    if magic != MAGIC:
        raise SnapshotFormatError(f"Wrong file magic in {path}: {magic!r}; expected {MAGIC!r}")
    if version != VERSION:
        raise SnapshotFormatError(f"Unsupported binary version {version} in {path}")
    if nvar != NVAR:
        raise SnapshotFormatError(f"Unexpected variable count {nvar} in {path}; expected {NVAR}")
    expected = HEADER.size + nx * ny * nvar * 8
    actual = os.path.getsize(path)
    if actual != expected:
        raise SnapshotFormatError(f"Binary size mismatch in {path}: expected={expected} bytes, "
                                  f"actual={actual} bytes")
    data = np.memmap(path, dtype="<f8", mode="r", offset=HEADER.size, shape=(ny, nx, nvar), order="C")
    return Snapshot(path, nx, ny, time, data)


def list_snapshots(out_dir):
    """The sorted ``solution_*.bin`` files of a run folder."""
    return sorted(glob.glob(os.path.join(out_dir, "solution_*.bin")))


def _snap(s):
    return s if isinstance(s, Snapshot) else read_snapshot(s)


def field(snap, name):
    """One record column as a (ny, nx) array view."""
    return _snap(snap).data[:, :, COLUMNS.index(name)]


def density_frame(snap):
    """(x (nx,), y (ny,), rho masked outside the fluid, time):
    visualize_sol.py:80-95."""
    s = _snap(snap)
    x = np.asarray(s.data[0, :, 0])
    y = np.asarray(s.data[:, 0, 1])
    rho = np.ma.masked_invalid(s.data[:, :, 2])
    return x, y, rho, s.time


def density_range(paths):
    """Global (min, max) of the density over every frame, as video_gen.py
    fixes one colour range for a whole movie (video_gen.py:150-181); frames
    without a valid cell are skipped, (inf, -inf) if none has one."""
    lo, hi = np.inf, -np.inf
    for p in paths:
        rho = density_frame(p)[2]
        if rho.count() > 0:
            lo = min(lo, float(rho.min()))
            hi = max(hi, float(rho.max()))
    return lo, hi


def _split_stats(d, valid):
    m = np.where(valid, d, -np.inf)
    ij = np.unravel_index(int(np.argmax(m)), d.shape)
    dv = d[valid]
    return float(d[ij]), float(np.mean(dv)), float(np.sqrt(np.mean(dv ** 2))), ij


def energy_split_row(snap):
    """[t, max|Ee-Ei|, max|Ee-Er|, max|Ei-Er|, the three means, the three
    rms, x and y of max|Ee-Er|] of one snapshot (energy_split.py:14-31). A
    valid cell has all three energies finite; the location of a maximum is
    its first cell in record (j-major) order."""
    s = _snap(snap)
    a = s.data
    ee, ei, er = a[:, :, 5], a[:, :, 6], a[:, :, 7]
    valid = np.isfinite(ee) & np.isfinite(ei) & np.isfinite(er)
    if not valid.any():
        raise SnapshotFormatError(f"{s.path}: no valid cell")
    m1, a1, r1, _ = _split_stats(np.abs(ee - ei), valid)
    m2, a2, r2, i2 = _split_stats(np.abs(ee - er), valid)
    m3, a3, r3, _ = _split_stats(np.abs(ei - er), valid)
    return [s.time, m1, m2, m3, a1, a2, a3, r1, r2, r3, float(a[i2][0]), float(a[i2][1])]


def energy_split_history(paths, csv_path=None):
    """energy_split_row of every snapshot, in the given order, as an
    (n, 12) array; with csv_path, also written as energy_split.py writes
    energy_split_history.csv (one header line, "%.16e", comma separated)."""
    rows = np.asarray([energy_split_row(p) for p in paths], dtype=np.float64).reshape(-1, 12)
    if csv_path is not None:
        np.savetxt(csv_path, rows, delimiter=",", header=ENERGY_SPLIT_HEADER, comments="", fmt="%.16e")
    return rows


def primitives(snap, gamma=(1.4, 1.4, 1.4)):
    """{u, v, pe, pi, pr, p} as (ny, nx) arrays (contour_gen.py:112-187,
    src/state.rs:245-256); NaN outside the fluid.  gamma: the run's
    (gamma_e, gamma_i, gamma_r)."""
    s = _snap(snap)
    a = s.data
    rho, mx, my = a[:, :, 2], a[:, :, 3], a[:, :, 4]
    with np.errstate(divide="ignore", invalid="ignore"):
        u = mx / rho
        v = my / rho
    kin = rho * (u * u + v * v) / 6.0
    pe = (gamma[0] - 1.0) * (a[:, :, 5] - kin)
    pi = (gamma[1] - 1.0) * (a[:, :, 6] - kin)
    pr = (gamma[2] - 1.0) * (a[:, :, 7] - kin)
    return {"u": u, "v": v, "pe": pe, "pi": pi, "pr": pr, "p": pe + pi + pr}


def totals(snap):
    """Sums of (rho, mom_x, mom_y, ee, ei, er) over the cells whose six
    values are finite (math.fsum: exact, so independent of cell order)."""
    import math

    a = _snap(snap).data
    st = a[:, :, 2:8].reshape(-1, 6)
    ok = np.all(np.isfinite(st), axis=1)
    return [math.fsum(st[ok, k]) for k in range(6)]
