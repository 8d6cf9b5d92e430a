"""Problem set-ups of the benchmarked workloads (geometry + initial state).

* ``periodic_box(N)``: the unit square [0,1)^2 with N x N cells (dx = 1/N),
  Periodic on all four sides; ``smooth_modes_state``: seeded sum of 8
  Fourier modes per primitive field (rho in [0.5,1.5], u,v in [-0.5,0.5],
  Te,Ti,Tr in [0.8,1.2]).
* ``translating_shock(N)``: Mach-2 planar shock at x = 0.25 on [0,1)^2,
  Constant inflow (post-shock) on the left, Outflow on the right, Periodic
  in y; front perturbed by a seeded 1e-3 sinusoid.
* ``cylinder(nx, mach)``: init::init_cylinder of the reference
  (src/init.rs:116-510) on the grid family h = 3/(nx-1).

Random draws use the reference test suite's LCG (src/main.rs:493-498).
"""
import math

import numpy as np

from . import BCType, Field, GridInfo

TWO_PI = 2.0 * math.pi


class LCG:
    """rng() of src/main.rs:493-498: x <- x*6364136223846793005 +
    1442695040888963407 (mod 2^64), value (x >> 11) / 2^53."""

    def __init__(self, seed):
        self.s = seed & 0xFFFFFFFFFFFFFFFF

    def __call__(self):
        self.s = (self.s * 6364136223846793005 + 1442695040888963407) & 0xFFFFFFFFFFFFFFFF
        return float(self.s >> 11) / float(1 << 53)


def _square(lx=1.0, ly=1.0):
    return [(0.0, 0.0), (lx, 0.0), (lx, ly), (0.0, ly)]


def periodic_box(n):
    grid = GridInfo(n, n, 1.0 / n, 1.0 / n, 0.0, 0.0)
    f = Field(grid, _square(), [BCType.periodic() for _ in range(4)])
    return f.build_ghosts()


def primitive_to_state(rho, u, v, te, ti, tr, cve=1.0, cvi=1.0, a_rad=1.0):
    """Conservative 3-T state from (rho, u, v, Te, Ti, Tr), the inverse of
    Derived::from_state (src/state.rs:241-275): E_k = rho c_k T_k + rho w^2/6
    for e, i and E_r = A Tr^4 + rho w^2/6."""
    w2 = u * u + v * v
    kin = rho * w2 / 6.0
    return np.stack([rho, rho * u, rho * v, rho * cve * te + kin, rho * cvi * ti + kin,
                     a_rad * tr ** 4 + kin], axis=-1)


def fourier_field(x, y, rng, n_modes=8, kmax=4):
    ks, amps, phs = [], [], []
    for _ in range(n_modes):
        while True:
            kx = int(rng() * (2 * kmax + 1)) - kmax
            ky = int(rng() * (2 * kmax + 1)) - kmax
            if kx != 0 or ky != 0:
                break
        ks.append((kx, ky))
        amps.append(rng())
        phs.append(TWO_PI * rng())
    s = np.zeros_like(x)
    for (kx, ky), a, p in zip(ks, amps, phs):
        s += a * np.sin(TWO_PI * (kx * x + ky * y) + p)
    return s / sum(amps)


def smooth_modes_state(field, seed=3):
    g = field.grid
    x = g.x0 + np.arange(g.nx, dtype=np.float64)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    y = g.y0 + np.arange(g.ny, dtype=np.float64)[None, :] * g.dy + 0.0 * np.arange(g.nx)[:, None]
    rng = LCG(seed)
    mids = [(1.0, 0.5), (0.0, 0.5), (0.0, 0.5), (1.0, 0.2), (1.0, 0.2), (1.0, 0.2)]
    q = [mid + half * fourier_field(x, y, rng) for mid, half in mids]
    u = primitive_to_state(*q)
    u[~field.fluid] = 0.0
    return np.ascontiguousarray(u)


# --------------------------------------------------------- shock problem

def rankine_hugoniot_mach2(gamma=1.4):
    """Pre state rho=1, u=v=0, Te=Ti=Tr=1 (E_k = 1 each, p = 1.2); post
    state from the gamma-gas jump at M = 2 with Te=Ti=Tr in equilibrium."""
    rho1, p1 = 1.0, 1.2
    c1 = math.sqrt(1.68)  # gamma*ge*1 summed over the three energies
    m = 2.0
    s = m * c1
    rho2 = rho1 * (gamma + 1.0) * m * m / ((gamma - 1.0) * m * m + 2.0)
    p2 = p1 * (2.0 * gamma * m * m - (gamma - 1.0)) / (gamma + 1.0)
    u2 = s * (1.0 - rho1 / rho2)
    # 2T + T^4/rho2 = p2 / (0.4 rho2): specific internal energies (e_e = e_i
    # = T, e_r = T^4/rho) times 0.4 give p / rho.
    t = 1.0
    for _ in range(100):
        fval = 2.0 * t + t ** 4 / rho2 - p2 / (0.4 * rho2)
        t -= fval / (2.0 + 4.0 * t ** 3 / rho2)
    return (rho1, 0.0, 1.0), (rho2, u2, t), s


def translating_shock(n, sigma=0.25, l_domain=1.0):
    pre, post, _ = rankine_hugoniot_mach2()
    rho2, u2, t2 = post
    post_state = primitive_to_state(np.array(rho2), np.array(u2), np.array(0.0), np.array(t2),
                                    np.array(t2), np.array(t2))
    grid = GridInfo(n, n, 1.0 / n, 1.0 / n, 0.0, 0.0)
    # bottom, right, top, left
    bcs = [BCType.periodic(), BCType.outflow(1.2, sigma, l_domain), BCType.periodic(),
           BCType.constant(post_state.tolist())]
    f = Field(grid, _square(), bcs)
    return f.build_ghosts()


def translating_shock_state(field, seed=2):
    g = field.grid
    (rho1, u1, t1), (rho2, u2, t2), _ = rankine_hugoniot_mach2()
    rng = LCG(seed)
    a = [rng() for _ in range(4)]
    ph = [TWO_PI * rng() for _ in range(4)]
    x = g.x0 + np.arange(g.nx)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    y = g.y0 + np.arange(g.ny)[None, :] * g.dy + 0.0 * np.arange(g.nx)[:, None]
    front = 0.25 + 1e-3 * sum(a[m] * np.sin(TWO_PI * (m + 1) * y + ph[m]) for m in range(4)) / sum(a)
    behind = x < front
    rho = np.where(behind, rho2, rho1)
    u = np.where(behind, u2, u1)
    t = np.where(behind, t2, t1)
    # radiation temperature from E_r = A Tr^4 with Tr = T
    st = primitive_to_state(rho, u, np.zeros_like(u), t, t, t)
    st[~field.fluid] = 0.0
    return np.ascontiguousarray(st)


def cylinder(nx=121, mach=3.0):
    return Field.init_cylinder(nx, mach)


def euler_to_three_energy(rho, ux, uy, p, gamma=1.4):
    """src/init.rs:7-24: conservative 3-T state of an Euler state, the total
    energy split equally into thirds."""
    kinetic = 0.5 * rho * (ux * ux + uy * uy)
    e_total = p / (gamma - 1.0) + kinetic
    return [rho, rho * ux, rho * uy, e_total / 3.0, e_total / 3.0, e_total / 3.0]


def double_mach_moving_shock():
    """The TimeDependent BC that init_double_mach defines for its top side
    (src/init.rs:57-65): post-shock where P0.x <= 10 t, pre-shock elsewhere
    (Mach-10 states of src/init.rs:45-49), as the built-in moving front
    profile."""
    from . import BCType
    pre = euler_to_three_energy(1.4, 0.0, 0.0, 1.0)
    post = euler_to_three_energy(8.0, 8.25, 0.0, 116.5)
    return BCType.moving_front(post, pre, 10.0, 0.0)


CYLINDER_CLASSES = (121, 161, 201, 241)


def batched_draw(n_solves=64, seed=4):
    """The batched workload's draw: per solve a grid class uniform over
    {121x481, 161x641, 201x801, 241x961} (h = 3/(nx-1)) and an inflow Mach
    number ~ U[2, 4], from the LCG."""
    rng = LCG(seed)
    draws = []
    for _ in range(n_solves):
        nx = CYLINDER_CLASSES[min(3, int(rng() * 4))]
        draws.append((nx, 2.0 + 2.0 * rng()))
    return draws


def batched_cylinders(draws):
    """Fields and initial states of a batched draw: one init_cylinder per
    grid class, cloned per solve (the heavy-ghost tables are shared) with
    the solve's freestream on the FarField sides."""
    from . import BC_FARFIELD, cylinder_freestream
    base = {nx: Field.init_cylinder(nx, 3.0) for nx in sorted(set(d[0] for d in draws))}
    fields, states = [], []
    for nx, mach in draws:
        f = base[nx].clone()
        f.set_bc_state(BC_FARFIELD, cylinder_freestream(mach))
        fields.append(f)
        states.append(f.cylinder_state(mach))
    return fields, states


# ----------------------------------------------- differentiation problem

# Exchange coefficients of the differentiated workload: switched on (the
# reference defaults are 0, src/constant.rs:17-20) so that the source term
# and its parameters enter the gradient.
ADJOINT_PHYSICS = dict(omega_ei=1.0, omega_er=0.1)


def gaussian_window(field, seed=5):
    """w = exp(-|x - x_c|^2 / (2 sigma^2)) on the cell nodes, with
    x_c ~ U(0.25, 0.75)^2 and sigma ~ U(0.05, 0.15) drawn from the LCG."""
    g = field.grid
    rng = LCG(seed)
    xc, yc = 0.25 + 0.5 * rng(), 0.25 + 0.5 * rng()
    sig = 0.05 + 0.1 * rng()
    x = g.x0 + np.arange(g.nx)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    y = g.y0 + np.arange(g.ny)[None, :] * g.dy + 0.0 * np.arange(g.nx)[:, None]
    return np.exp(-((x - xc) ** 2 + (y - yc) ** 2) / (2.0 * sig * sig))


def window_loss_cotangent(field, seed=5):
    """dJ/du^N of J = sum(w * E_r) (the sixth state field)."""
    lam = np.zeros((field.grid.nx, field.grid.ny, 6))
    lam[..., 5] = gaussian_window(field, seed)
    lam[~field.fluid] = 0.0
    return lam


def state_direction(u, fluid, seed):
    """Seeded direction in state space, scaled per field by the mean
    magnitude of that field, zero on non-fluid cells."""
    rng = np.random.default_rng(seed)
    du = rng.standard_normal(u.shape) * np.abs(u[fluid]).mean(axis=0)
    du[~fluid] = 0.0
    return du


def smooth_direction(field, u, seed):
    """Seeded smooth direction in state space: per field a sum of 8 random
    Fourier modes (as smooth_modes_state) scaled by the mean magnitude of
    that field, zero on non-fluid cells."""
    g = field.grid
    x = g.x0 + np.arange(g.nx, dtype=np.float64)[:, None] * g.dx + 0.0 * np.arange(g.ny)[None, :]
    y = g.y0 + np.arange(g.ny, dtype=np.float64)[None, :] * g.dy + 0.0 * np.arange(g.nx)[:, None]
    rng = LCG(seed)
    scale = np.abs(u[field.fluid]).mean(axis=0)
    du = np.stack([scale[k] * fourier_field(x, y, rng) for k in range(6)], axis=-1)
    du[~field.fluid] = 0.0
    return du


def walled_box(n):
    """Unit square with ReflectiveWall bottom and top, Constant inflow on
    the left and ZerothOrder on the right (the cheap ghost kinds)."""
    grid = GridInfo(n, n, 1.0 / n, 1.0 / n, 0.0, 0.0)
    inflow = primitive_to_state(np.array(1.0), np.array(0.3), np.array(0.0), np.array(1.0), np.array(1.0),
                                np.array(1.0))
    bcs = [BCType.reflective_wall(), BCType.zeroth_order(), BCType.reflective_wall(),
           BCType.constant(inflow.tolist())]
    f = Field(grid, _square(), bcs)
    return f.build_ghosts()
