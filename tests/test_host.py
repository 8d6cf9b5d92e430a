"""Host-side checks (CPU): the C ABI surface and the GhostGrid bookkeeping.

Ghost discovery, ghost order, the access table and the cheap-BC source
cells must agree with the reference bit for bit (src/ghost.rs:50-118,
352-376; src/bc1.rs:547-607, 930-1037).  They are pinned here by an
independent brute-force recount in Python (crossing-number mask, cross
offsets, sorted set) and by the fluid/ghost counts of the cylinder grids
that an arithmetic recount gave before this code existed (55,132 / 4,914
at 121x481, 97,987 / 6,560, 153,084 / 8,204, 220,444 / 9,852).
"""
import ctypes
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "rh3t.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rh3t_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_header_symbol(rh3t):
    names = header_functions()
    assert len(names) >= 25
    L = ctypes.CDLL(rh3t.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python mirror binds the whole ABI with explicit signatures
    assert set(names) == set(rh3t.SIGNATURES), set(names) ^ set(rh3t.SIGNATURES)


def _rust_block():
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    blocks = re.findall(r"```rust\n(.*?)```", text, flags=re.S)
    return next(b for b in blocks if 'extern "C"' in b)


def header_structs():
    """typedef struct { ... } name; of include/rh3t.h: name -> field names."""
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(ROOT, "include", "rh3t.h")).read(), flags=re.S)
    out = {}
    for body, name in re.findall(r"typedef struct\s*\{(.*?)\}\s*(rh3t_[a-z_]+)\s*;", text, flags=re.S):
        fields = []
        for decl in body.split(";"):
            decl = decl.strip()
            if not decl:
                continue
            names = decl.split(None, 1)[1]
            fields += [re.sub(r"\[.*?\]", "", n).strip().lstrip("*") for n in names.split(",")]
        out[name] = fields
    return out


def test_rust_binding_matches_header():
    """The extern "C" block a maintainer would add to the reference crate
     binds exactly the header's functions, and its
    #[repr(C)] structs carry the header's fields in the header's order."""
    rust = _rust_block()
    fns = set(re.findall(r"pub fn (rh3t_[a-z0-9_]+)\(", rust))
    assert fns == set(header_functions()), fns ^ set(header_functions())
    rstructs = {}
    for name, body in re.findall(r"pub struct (rh3t_[a-z_]+)\s*\{(.*?)\}", rust, flags=re.S):
        rstructs[name] = re.findall(r"pub ([a-z_0-9]+)\s*:", body)
    for name, fields in header_structs().items():
        assert name in rstructs, name
        assert rstructs[name] == fields, (name, rstructs[name], fields)
    for opaque in ("rh3t_problem", "rh3t_ctx", "rh3t_batch", "rh3t_adjoint"):
        assert opaque in rstructs


def test_no_device_fails_loudly(rh3t, workloads):
    if rh3t.device_count() > 0:
        pytest.skip("a GPU is visible")
    f = workloads.periodic_box(16)
    with pytest.raises(rh3t.RH3TError) as e:
        f.solver()
    assert e.value.status == rh3t.ERR_NO_DEVICE


# ------------------------------------------------------------ brute force

def is_fluid_poly(pts, px, py, inside=True):
    """Crossing-number test of src/geometry.rs:254-278, vectorised."""
    inside_mask = np.zeros(np.shape(px), dtype=bool)
    n = len(pts)
    for k in range(n):
        ax, ay = pts[k]
        bx, by = pts[(k + 1) % n]
        crosses = (ay > py) != (by > py)
        with np.errstate(divide="ignore", invalid="ignore"):
            xi = ax + (py - ay) * (bx - ax) / (by - ay)
        inside_mask ^= crosses & (px < xi)
    return inside_mask if inside else ~inside_mask


def cylinder_polygon(h):
    pts = [(-3.0, -6.0), (0.0, -6.0), (0.0, -1.0 + 0.0125)]
    for k in range(1, 361):
        th = -0.5 * math.pi - math.pi * (k / 360.0)
        pts.append((math.cos(th), math.sin(th) + 0.0125))
    pts += [(0.0, 6.0), (-3.0, 6.0)]
    return pts


def brute_ghosts(fluid_fn, nx, ny):
    ii, jj = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    fl = fluid_fn(ii, jj)
    ghosts = set()
    fi, fj = np.nonzero(fl)
    for d in range(-4, 5):
        for (ti, tj) in ((fi + d, fj), (fi, fj + d)):
            ok = fluid_fn(ti, tj)
            for a, b in zip(ti[~ok], tj[~ok]):
                ghosts.add((int(a), int(b)))
    return fl, sorted(ghosts)


@pytest.mark.parametrize("nx,nfluid,nghost", [(121, 55132, 4914), (161, 97987, 6560)])
def test_cylinder_fluid_and_ghost_sets(rh3t, nx, nfluid, nghost):
    f = rh3t.Field.init_cylinder(nx, 3.0)
    g = f.grid
    assert (g.nx, g.ny) == (nx, 4 * (nx - 1) + 1)
    pts = cylinder_polygon(g.dx)
    inner = [(-1002.0, -1002.0), (-1001.0, -1002.0), (-1001.0, -1001.0), (-1002.0, -1001.0)]

    def fluid_fn(i, j):
        x = g.x0 + np.asarray(i, dtype=np.float64) * g.dx
        y = g.y0 + np.asarray(j, dtype=np.float64) * g.dy
        return is_fluid_poly(pts, x, y) & is_fluid_poly(inner, x, y, inside=False)

    fl, ghosts = brute_ghosts(fluid_fn, g.nx, g.ny)
    assert np.array_equal(f.fluid, fl)
    gh = f.ghosts()
    assert list(zip(gh["i"].tolist(), gh["j"].tolist())) == ghosts
    assert f.counts()[:2] == (nfluid, nghost)
    # cylinder BCs: FarField on the straight sides, ReflectiveWall on the arc
    kinds = set(gh["bc_kind"].tolist())
    assert kinds == {rh3t.BC_FARFIELD, rh3t.BC_REFLECTIVE_WALL}


def double_mach_polygon():
    s3 = math.sqrt(3.0)
    a = (-0.5 - s3 / 12.0, 0.0)
    c = (23.0 * s3 / 12.0, 23.0 / 12.0)
    d = (c[0], 23.0 / 12.0 + s3 / 2.0)
    return [a, (0.0, 0.0), c, d, (a[0], d[1])]


@pytest.mark.parametrize("h", [1.0 / 40.0, 1.0 / 320.0], ids=["h40", "h320"])
def test_double_mach_fluid_and_ghost_sets(rh3t, h):
    """init_double_mach (src/init.rs:26-113): grid size, fluid mask and
    ghost set against the brute-force crossing rule and radius-4 cross
    discovery; BC kinds; the initial shock at x = 0 (src/init.rs:97-104)."""
    f = rh3t.Field.init_double_mach(h)
    g = f.grid
    pts = double_mach_polygon()
    lx, ly = pts[2][0] - pts[0][0], pts[3][1]
    assert (g.nx, g.ny) == (math.ceil(lx / h) + 1, math.ceil(ly / h) + 1)
    assert (g.x0, g.y0, g.dx, g.dy) == (pts[0][0], 0.0, h, h)
    inner = [(-1002.0, -1002.0), (-1001.0, -1002.0), (-1001.0, -1001.0), (-1002.0, -1001.0)]

    def fluid_fn(i, j):
        x = g.x0 + np.asarray(i, dtype=np.float64) * g.dx
        y = g.y0 + np.asarray(j, dtype=np.float64) * g.dy
        return is_fluid_poly(pts, x, y) & is_fluid_poly(inner, x, y, inside=False)

    fl, ghosts = brute_ghosts(fluid_fn, g.nx, g.ny)
    assert np.array_equal(f.fluid, fl)
    gh = f.ghosts()
    assert list(zip(gh["i"].tolist(), gh["j"].tolist())) == ghosts
    assert set(gh["bc_kind"].tolist()) == {rh3t.BC_CONSTANT, rh3t.BC_WALL, rh3t.BC_ZEROTH_ORDER}
    # Wall ghosts are the heavy ones; every one of them has its tables
    assert f.counts()[2] == int(np.sum(gh["bc_kind"] == rh3t.BC_WALL))
    u = f.double_mach_state()
    x = g.x0 + np.arange(g.nx, dtype=np.float64)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    post = x <= 0.0
    assert np.all(u[fl & post, 0] == 8.0) and np.all(u[fl & ~post, 0] == 1.4)
    assert np.all(u[fl & post, 1] == 8.0 * 8.25)
    assert np.all(u[~fl] == 0.0)


@pytest.mark.parametrize("nx,nfluid,nghost", [(201, 153084, 8204), (241, 220444, 9852)])
def test_cylinder_counts_large(rh3t, nx, nfluid, nghost):
    f = rh3t.Field.init_cylinder(nx, 2.5)
    assert f.counts()[:2] == (nfluid, nghost)


def test_access_table_is_point_addressing(rh3t):
    """The reference's access[l*17+k] (src/ghost.rs:68-107) is a function of
    the target point only -- the property the halo-padded device layout
    relies on."""
    f = rh3t.Field.init_cylinder(121, 3.0)
    g = f.grid
    acc = f.access_table().reshape(g.nx * g.ny, 17)
    gh = f.ghosts()
    gid = {(int(a), int(b)): k for k, (a, b) in enumerate(zip(gh["i"], gh["j"]))}
    offs = sorted(set([(d, 0) for d in range(-4, 5)] + [(0, d) for d in range(-4, 5)]))
    assert len(offs) == 17
    fluid = f.fluid.reshape(-1)
    ncell = g.nx * g.ny
    for l in np.nonzero(fluid)[0][::7]:
        i, j = divmod(int(l), g.ny)
        for k, (di, dj) in enumerate(offs):
            ti, tj = i + di, j + dj
            t = int(acc[l, k])
            if 0 <= ti < g.nx and 0 <= tj < g.ny and fluid[ti * g.ny + tj]:
                assert t == ti * g.ny + tj
            else:
                assert t == ncell + gid[(ti, tj)]


def test_periodic_sources(rh3t, workloads):
    f = workloads.periodic_box(32)
    gh = f.ghosts()
    assert len(gh) == 16 * 32
    assert np.array_equal(gh["src_i"], gh["i"] % 32)
    assert np.array_equal(gh["src_j"], gh["j"] % 32)


def test_reflective_sources_are_nearest_mirror_cells(rh3t):
    """reflective_wall_value's interior cell (src/bc1.rs:930-1003): nearest
    fluid candidate of the mirror point in the fixed 3x3 order."""
    f = rh3t.Field.init_cylinder(121, 3.0)
    g = f.grid
    fl = f.fluid
    gh = f.ghosts()
    refl = gh[gh["bc_kind"] == rh3t.BC_REFLECTIVE_WALL]
    assert len(refl) > 300
    for row in refl[::5]:
        px, py = g.x(row["i"]), g.y(row["j"])
        mx, my = 2.0 * row["p0x"] - px, 2.0 * row["p0y"] - py
        ic, jc = int(round((mx - g.x0) / g.dx)), int(round((my - g.y0) / g.dy))
        cand = [(ic, jc), (ic - 1, jc), (ic + 1, jc), (ic, jc - 1), (ic, jc + 1), (ic - 1, jc - 1),
                (ic - 1, jc + 1), (ic + 1, jc - 1), (ic + 1, jc + 1)]
        best, bd = None, math.inf
        for a, b in cand:
            if 0 <= a < g.nx and 0 <= b < g.ny and fl[a, b]:
                d2 = (g.x(a) - mx) ** 2 + (g.y(b) - my) ** 2
                if d2 < bd:
                    best, bd = (a, b), d2
        assert (row["src_i"], row["src_j"]) == best


def test_cylinder_arc_geometry(rh3t):
    """Arc override of init_cylinder (src/init.rs:318-344): centre (0, 0.0125),
    radius 1; normals near the stagnation point are ~+x; mirror symmetry about
    y = 0.0125.  (The reference's own tests assume a centre at (0,0) and are
    stale; these are the corrected expectations.)"""
    f = rh3t.Field.init_cylinder(121, 3.0)
    g = f.grid
    gh = f.ghosts()
    arc = gh[gh["bc_kind"] == rh3t.BC_REFLECTIVE_WALL]
    r = np.hypot(arc["p0x"], arc["p0y"] - 0.0125)
    assert np.abs(r - 1.0).max() < 1e-12
    d = np.hypot(arc["p0x"] + 1.0, arc["p0y"] - 0.0125)
    k = int(np.argmin(d))
    # y = 0.0125 falls halfway between grid lines: the ghosts nearest the
    # stagnation point come as a mirror pair straddling it, with normals
    # tilted by +-atan(h/2) from +x.
    assert d[k] < g.dy
    assert arc["nx"][k] > math.cos(g.dy) and abs(arc["ny"][k]) < g.dy
    # mirror pairs: y_j + y_j' = 0.025  <=>  j + j' = 481
    lookup = {(int(a), int(b)): n for n, (a, b) in enumerate(zip(gh["i"], gh["j"]))}
    i_g = int(round((-0.975 - g.x0) / g.dx))
    pairs = 0
    for j in range(241, 300):
        a, b = lookup.get((i_g, j)), lookup.get((i_g, 481 - j))
        if a is None or b is None:
            continue
        pa, pb = gh[a], gh[b]
        assert abs(pa["p0x"] - pb["p0x"]) < 1e-9
        assert abs((pa["p0y"] - 0.0125) + (pb["p0y"] - 0.0125)) < 1e-9
        assert abs(pa["nx"] - pb["nx"]) < 1e-9 and abs(pa["ny"] + pb["ny"]) < 1e-9
        assert abs(pa["distance"] - pb["distance"]) < 1e-9
        pairs += 1
    assert pairs >= 5


# ------------------------------------------------ GhostBC least squares

def rect_wall_field(rh3t, n=80):
    grid = rh3t.GridInfo(n, n, 1.0 / n, 1.0 / n, 0.0, 0.0)
    f = rh3t.Field(grid, [(0.0, 0.0), (1.0, 0.0), (1.0, 1.0), (0.0, 1.0)], [rh3t.BCType.wall() for _ in range(4)])
    return f.build_ghosts()


def _heavy_of(f, i, j):
    gh = f.ghosts()
    k = [n for n, (a, b) in enumerate(zip(gh["i"], gh["j"])) if (a, b) == (i, j)][0]
    return gh[k], f.ghost_bc(int(gh[k]["heavy_id"]))


OFF = [0, 1, 5, 14, 30, 55]
SOFF = [0, 1, 17, 98, 354, 979]
WOFF = [0, 1, 9, 36, 100, 225]


def test_er_stencil_structure_axis_walls(rh3t):
    """E_r cardinality and structure (src/bc1.rs:1828-1960): vertical wall at
    ghost (-1, 40), horizontal wall at (40, -1)."""
    f = rect_wall_field(rh3t)
    ny = f.grid.ny
    for (gi, gj), vertical in (((-1, 40), True), ((40, -1), False)):
        g, bc = _heavy_of(f, gi, gj)
        for r in range(5):
            w = r + 1
            idx = [divmod(int(s), ny) for s in bc.stencil[OFF[r]:OFF[r + 1]]]
            assert len(idx) == w * w
            for l in range(w):
                s = idx[l * w:(l + 1) * w]
                if vertical:
                    assert all(q[0] == s[0][0] for q in s)
                    assert all(s[a][1] == s[a - 1][1] + 1 for a in range(1, w))
                else:
                    assert all(q[1] == s[0][1] for q in s)
                    assert all(s[a][0] == s[a - 1][0] + 1 for a in range(1, w))
            for l in range(1, w):
                a, b = idx[(l - 1) * w], idx[l * w]
                assert (b[0] == a[0] + 1) if vertical else (b[1] == a[1] + 1)


def test_oblique_wall_stencils_are_fluid(rh3t):
    """Oblique boundary (src/bc1.rs:2051-2149): every E_r of every heavy ghost
    consists of fluid cells in Cartesian rows/columns.  The reference checks
    one point outside the hypotenuse of a triangle; a full GhostGrid::build
    on that triangle cannot fit the 5-wide substencils into its acute tips
    (the reference panics there too), so the whole-boundary check uses a
    square with a 45-degree cut corner."""
    n = 100
    grid = rh3t.GridInfo(n, n, 1.0 / n, 1.0 / n, 0.0, 0.0)
    pts = [(0.0, 0.0), (1.0, 0.0), (1.0, 0.6), (0.6, 1.0), (0.0, 1.0)]
    f = rh3t.Field(grid, pts, [rh3t.BCType.wall() for _ in range(5)]).build_ghosts()
    fl = f.fluid.reshape(-1)
    nh = f.counts()[2]
    assert nh > 100
    for h in range(0, nh, 3):
        bc = f.ghost_bc(h)
        st = np.array(list(bc.stencil))
        assert fl[st].all()
        for r in range(1, 5):
            w = r + 1
            idx = [divmod(int(s), n) for s in st[OFF[r]:OFF[r + 1]]]
            for l in range(w):
                s = idx[l * w:(l + 1) * w]
                vert = s[0][0] == s[1][0]
                if vert:
                    assert all(q[0] == s[0][0] for q in s) and all(s[a][1] == s[a - 1][1] + 1 for a in range(1, w))
                else:
                    assert all(q[1] == s[0][1] for q in s) and all(s[a][0] == s[a - 1][0] + 1 for a in range(1, w))


def _local(f, g, lin):
    i, j = divmod(int(lin), f.grid.ny)
    dx, dy = f.grid.x(i) - g["p0x"], f.grid.y(j) - g["p0y"]
    return g["nx"] * dx + g["ny"] * dy, -dx * g["ny"] + dy * g["nx"]


@pytest.mark.parametrize("ghost", [(-1, 40), (-3, 17), (40, -2), (82, 61)])
def test_ls_weights_reproduce_normal_derivatives(rh3t, ghost):
    """w[r][k] . p(E_r) = d^k p / dn^k at P0 for any p of degree <= r in the
    local (n, t) frame (Tan et al. 2012, Sec. 2.4; src/bc1.rs:176-185)."""
    f = rect_wall_field(rh3t)
    g, bc = _heavy_of(f, *ghost)
    w = np.frombuffer(bc.w, dtype=np.float64)
    rng = np.random.default_rng(11)
    for r in range(5):
        m = (r + 1) ** 2
        loc = [_local(f, g, s) for s in bc.stencil[OFF[r]:OFF[r + 1]]]
        coef = {(a, b): rng.uniform(-1, 1) for a in range(r + 1) for b in range(r + 1 - a)}
        data = np.array([sum(c * x ** a * y ** b for (a, b), c in coef.items()) for x, y in loc])
        for k in range(r + 1):
            got = w[WOFF[r] + k * m: WOFF[r] + (k + 1) * m] @ data
            want = math.factorial(k) * coef[(k, 0)]
            assert abs(got - want) <= 1e-7 * max(1.0, abs(want)), (r, k, got, want)


@pytest.mark.parametrize("ghost", [(-1, 40), (40, -2)])
def test_smoothness_forms_closed_form(rh3t, ghost):
    """beta_r = d^T S_r d on the E_r data of p(n,t) (src/bc1.rs:62-85,
    1070-1196): 0 for a constant, h^2 for p = n, 13 h^4 / 3 for p = n^2
    (r >= 2); 2h^2 for r = 0."""
    f = rect_wall_field(rh3t)
    h = f.grid.dx
    g, bc = _heavy_of(f, *ghost)
    S = np.frombuffer(bc.s, dtype=np.float64)
    assert abs(S[0] - 2 * h * h) <= 1e-18
    d = np.array(list(bc.d))
    assert np.allclose(d[:4], [2 * h ** 4, 2 * h ** 3, 2 * h ** 2, 2 * h], rtol=0, atol=1e-18)
    assert abs(d.sum() - 1.0) < 1e-15
    for r in range(1, 5):
        m = (r + 1) ** 2
        Sr = S[SOFF[r]:SOFF[r + 1]].reshape(m, m)
        loc = np.array([_local(f, g, s) for s in bc.stencil[OFF[r]:OFF[r + 1]]])
        one = np.ones(m)
        assert abs(one @ Sr @ one) <= 1e-9 * np.abs(Sr).sum() * h ** 2
        n1 = loc[:, 0]
        assert abs(n1 @ Sr @ n1 - h * h) <= 1e-9 * h * h
        if r >= 2:
            n2 = loc[:, 0] ** 2
            assert abs(n2 @ Sr @ n2 - 13.0 * h ** 4 / 3.0) <= 1e-8 * h ** 4


def test_time_dependent_profile_refused_on_the_host(rh3t, workloads):
    """A TimeDependent BC without a built-in profile of this build is
    refused when the problem is created (no device needed); the moving front
    profile builds, and its top-side ghosts are TimeDependent."""
    bad = workloads.double_mach_moving_shock()
    bad.profile = 7
    with pytest.raises(rh3t.RH3TError) as ei:
        rh3t.Field.init_double_mach(1.0 / 40.0, top=bad)
    assert ei.value.status == rh3t.ERR_UNSUPPORTED
    f = rh3t.Field.init_double_mach(1.0 / 40.0, top=workloads.double_mach_moving_shock())
    gh = f.ghosts()
    td = gh["bc_kind"] == rh3t.BC_TIME_DEPENDENT
    assert td.sum() > 50 and np.all(gh["heavy_id"][td] == -1)
    # the same ghost set as the reference's ZerothOrder top
    g0 = rh3t.Field.init_double_mach(1.0 / 40.0).ghosts()
    assert np.array_equal(g0[["i", "j"]], gh[["i", "j"]])
