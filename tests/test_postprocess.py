"""RH3TBIN1 readers (py_utils/energy_split.py, visualize_sol.py,
video_gen.py, py_utils/contour_gen.py, py_utils/conservation_check.py), on
the CPU, against snapshots written by the time loop's writer: the reader's
checks, and each diagnostic against a brute-force loop over the records in
file order, a closed form, or the oracle's Derived values."""
import math
import os
import struct
from importlib import import_module

import numpy as np
import pytest

import oracle_bridge as ob

pp = import_module("3t_rh_amd.postprocess")
tl = import_module("3t_rh_amd.timeloop")


def _state(field, seed):
    g = field.grid
    rng = np.random.default_rng(seed)
    u = rng.uniform(0.5, 1.5, size=(g.nx, g.ny, 6))
    u[..., 1:3] -= 1.0
    u[~field.fluid] = -7.0
    return u


@pytest.fixture
def cyl(rh3t):
    return rh3t.Field.init_cylinder(121, 3.0)


def _write(field, u, t, path):
    tl.save_data(field, u, t, str(path))
    return str(path)


def test_reader_layout(cyl, tmp_path):
    g = cyl.grid
    u = _state(cyl, 1)
    s = pp.read_snapshot(_write(cyl, u, 0.25, tmp_path / "solution_0000.bin"))
    assert (s.nx, s.ny, s.time) == (g.nx, g.ny, 0.25)
    assert s.data.shape == (g.ny, g.nx, 8)
    x, y, rho, t = pp.density_frame(s)
    assert np.array_equal(x, g.x0 + np.arange(g.nx) * g.dx) and np.array_equal(y, g.y0 + np.arange(g.ny) * g.dy)
    assert t == 0.25
    assert np.array_equal(rho.mask, ~cyl.fluid.T)
    assert np.array_equal(rho.compressed(), u[..., 0].T[cyl.fluid.T])
    for k, name in enumerate(pp.COLUMNS[2:]):
        col = pp.field(s, name)
        assert np.array_equal(col[cyl.fluid.T], u[..., k].T[cyl.fluid.T])
        assert np.all(np.isnan(col[~cyl.fluid.T]))


def test_reader_refusals(cyl, tmp_path):
    p = _write(cyl, _state(cyl, 2), 0.0, tmp_path / "s.bin")
    raw = open(p, "rb").read()
    bad = str(tmp_path / "bad.bin")
    cases = [
        (raw[:30], "Incomplete binary header"),
        (b"RH3TBIN2" + raw[8:], "Wrong file magic"),
        (raw[:8] + struct.pack("<I", 2) + raw[12:], "Unsupported binary version 2"),
        (raw[:12] + struct.pack("<I", 7) + raw[16:], "Unexpected variable count 7"),
        (raw[:-8], "Binary size mismatch"),
        (raw + b"\0", "Binary size mismatch"),
    ]
    for blob, msg in cases:
        open(bad, "wb").write(blob)
        with pytest.raises(pp.SnapshotFormatError, match=msg):
            pp.read_snapshot(bad)
    assert issubclass(pp.SnapshotFormatError, RuntimeError)


def test_listing_skips_partial_files(cyl, tmp_path):
    u = _state(cyl, 3)
    for k in (2, 0, 1):
        _write(cyl, u, k * 0.1, tmp_path / f"solution_{k:04d}.bin")
    open(tmp_path / "solution_0003.bin.tmp", "wb").write(b"partial")
    names = [os.path.basename(p) for p in pp.list_snapshots(str(tmp_path))]
    assert names == ["solution_0000.bin", "solution_0001.bin", "solution_0002.bin"]


def _brute_split(u, field):
    """The diagnostic over the records in file order (j outer, i inner),
    strict '>' keeping the first maximum, exact sums."""
    g = field.grid
    best = [-math.inf] * 3
    at = [None] * 3
    vals = [[], [], []]
    for j in range(g.ny):
        for i in range(g.nx):
            if not field.fluid[i, j]:
                continue
            ee, ei, er = u[i, j, 3], u[i, j, 4], u[i, j, 5]
            for k, d in enumerate((abs(ee - ei), abs(ee - er), abs(ei - er))):
                vals[k].append(d)
                if d > best[k]:
                    best[k], at[k] = d, (i, j)
    n = len(vals[0])
    mean = [math.fsum(v) / n for v in vals]
    rms = [math.sqrt(math.fsum(d * d for d in v) / n) for v in vals]
    i2, j2 = at[1]
    return best, mean, rms, (g.x0 + i2 * g.dx, g.y0 + j2 * g.dy)


def test_energy_split_against_brute_force(cyl, tmp_path):
    u = _state(cyl, 4)
    # a tie for max|Ee-Er| at two cells, A first in record (j-major) order
    # and B first in i-major order: A wins
    ja = 7
    ia = int(np.flatnonzero(cyl.fluid[:, ja])[-1])
    jb = ja + 50
    ib = int(np.flatnonzero(cyl.fluid[:, jb])[0])
    assert ib < ia
    for i, j in ((ia, ja), (ib, jb)):
        u[i, j, 3], u[i, j, 5] = 5.0, 1.0
    row = pp.energy_split_row(_write(cyl, u, 0.5, tmp_path / "s.bin"))
    best, mean, rms, xy = _brute_split(u, cyl)
    assert row[0] == 0.5
    assert row[1:4] == best
    assert row[10:12] == [cyl.grid.x0 + ia * cyl.grid.dx, cyl.grid.y0 + ja * cyl.grid.dy] == list(xy)
    np.testing.assert_allclose(row[4:7], mean, rtol=1e-14)
    np.testing.assert_allclose(row[7:10], rms, rtol=1e-14)


def test_energy_split_closed_form(cyl, tmp_path):
    """Ee - Ei = 0.25, Ee - Er = -0.5 on every fluid cell: max = mean = rms."""
    u = _state(cyl, 5)
    u[..., 3] = np.round(u[..., 3] * 64.0) / 64.0  # dyadic: the differences below are exact
    u[..., 4] = u[..., 3] - 0.25
    u[..., 5] = u[..., 3] + 0.5
    row = pp.energy_split_row(_write(cyl, u, 0.0, tmp_path / "s.bin"))
    for k, c in enumerate((0.25, 0.5, 0.75)):
        for v in (row[1 + k], row[4 + k], row[7 + k]):
            assert v == c
    # every cell ties: the first fluid cell in record order
    j0 = int(np.argmax(cyl.fluid.any(axis=0)))
    i0 = int(np.argmax(cyl.fluid[:, j0]))
    assert row[10:12] == [cyl.grid.x0 + i0 * cyl.grid.dx, cyl.grid.y0 + j0 * cyl.grid.dy]


def test_energy_split_history_csv(cyl, tmp_path):
    paths = [_write(cyl, _state(cyl, 10 + k), 0.1 * k, tmp_path / f"solution_{k:04d}.bin") for k in range(3)]
    csv = str(tmp_path / "energy_split_history.csv")
    rows = pp.energy_split_history(pp.list_snapshots(str(tmp_path)), csv)
    assert rows.shape == (3, 12)
    for k, p in enumerate(paths):
        assert list(rows[k]) == pp.energy_split_row(p)
    lines = open(csv).read().splitlines()
    assert lines[0] == pp.ENERGY_SPLIT_HEADER and len(lines) == 4
    assert np.array_equal(np.loadtxt(csv, delimiter=",", skiprows=1), rows), "%.16e round-trips every double"


def test_density_range(cyl, tmp_path):
    us = [_state(cyl, 20 + k) for k in range(3)]
    us[1][cyl.fluid, 0] += 3.0
    paths = [_write(cyl, u, 0.0, tmp_path / f"solution_{k:04d}.bin") for k, u in enumerate(us)]
    lo, hi = pp.density_range(paths)
    allr = np.concatenate([u[cyl.fluid, 0] for u in us])
    assert (lo, hi) == (float(allr.min()), float(allr.max()))


def test_primitives_match_oracle_derived(cyl, tmp_path):
    u = _state(cyl, 6)
    u[..., 0] += 0.5
    s = pp.read_snapshot(_write(cyl, u, 0.0, tmp_path / "s.bin"))
    prim = pp.primitives(s)
    for i, j in np.argwhere(cyl.fluid)[::997]:
        d = ob.derived(u[i, j])
        assert (prim["u"][j, i], prim["v"][j, i]) == (d.u, d.v)
        assert (prim["pe"][j, i], prim["pi"][j, i], prim["pr"][j, i]) == (d.pe, d.pi, d.pr)
        assert prim["p"][j, i] == d.pe + d.pi + d.pr
    assert np.all(np.isnan(prim["p"][~cyl.fluid.T]))


def test_totals(cyl, tmp_path):
    u = _state(cyl, 7)
    tot = pp.totals(_write(cyl, u, 0.0, tmp_path / "s.bin"))
    for k in range(6):
        assert tot[k] == math.fsum(u[cyl.fluid, k])
