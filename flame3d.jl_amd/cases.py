"""Synthetic inputs: grids with analytic metrics and initial states.

Everything is built with torch on a chosen device so the 256^3 / 512x256x256
inputs are generated in HBM directly; small cases run on the CPU for the
oracle tests.  Arrays follow the library layout: padded fields are
(NV, Nz+8, Ny+8, Nx+8), i fastest.

Metrics use the reference's convention (parse_mesh.jl:210-221): the nine
"dxi/dx" arrays are cofactors of the coordinate Jacobian, i.e. J^-1-weighted
metric terms, and J = 1/det(d(x,y,z)/d(xi,eta,zeta)).  Here the coordinate
derivatives are analytic instead of the reference's CD6 differences.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, List, Optional, Tuple

import numpy as np
import torch

NG = 4


# ---------------------------------------------------------------------------
# thermo helpers (vectorised NASA-7, thermo.jl:20-35, :136-160, :219-228)
# ---------------------------------------------------------------------------
def _coef(mech, dev, dtype=torch.float64):
    lo = torch.tensor(np.asarray(mech.coeffs_lo).T, dtype=dtype, device=dev)  # [Ns, 7]
    hi = torch.tensor(np.asarray(mech.coeffs_hi).T, dtype=dtype, device=dev)
    sep = torch.tensor(np.asarray(mech.coeffs_sep), dtype=dtype, device=dev)
    mw = torch.tensor(np.asarray(mech.mw), dtype=dtype, device=dev)
    return lo, hi, sep, mw


def internal_energy(mech, T: torch.Tensor, rhoi: torch.Tensor) -> torch.Tensor:
    """InternalEnergy(T, rho_i) [J/m^3] for T of shape S and rhoi (Ns, *S)."""
    lo, hi, sep, mw = _coef(mech, T.device)
    T2 = 0.5 * T * T
    T3 = 2.0 / 3.0 * T2 * T
    T4 = T2 * T2
    T5 = 0.8 * T4 * T
    out = torch.zeros_like(T)
    for n in range(mech.Ns):
        el = (lo[n, 0] - 1) * T + lo[n, 1] * T2 + lo[n, 2] * T3 + lo[n, 3] * T4 + lo[n, 4] * T5
        eh = ((hi[n, 0] - 1) * T + hi[n, 1] * T2 + hi[n, 2] * T3 + hi[n, 3] * T4 + hi[n, 4] * T5
              + (hi[n, 5] - lo[n, 5]))
        e = torch.where(T < sep[n], el, eh)
        out = out + rhoi[n] * e / mw[n]
    return out * mech.Ru


def cp_mix(mech, T: torch.Tensor, Y: torch.Tensor) -> torch.Tensor:
    """Mass-specific cp [J/kg/K] of a mixture with mass fractions Y."""
    lo, hi, sep, mw = _coef(mech, T.device)
    out = torch.zeros_like(T)
    for n in range(mech.Ns):
        c = torch.where(T < sep[n], torch.ones_like(T), torch.zeros_like(T))
        cpl = lo[n, 0] + lo[n, 1] * T + lo[n, 2] * T ** 2 + lo[n, 3] * T ** 3 + lo[n, 4] * T ** 4
        cph = hi[n, 0] + hi[n, 1] * T + hi[n, 2] * T ** 2 + hi[n, 3] * T ** 3 + hi[n, 4] * T ** 4
        out = out + Y[n] * (c * cpl + (1 - c) * cph) * mech.Ru / mw[n]
    return out


def mix_W(mech, Y: torch.Tensor) -> torch.Tensor:
    mw = torch.tensor(np.asarray(mech.mw), dtype=Y.dtype, device=Y.device)
    s = torch.zeros_like(Y[0])
    for n in range(mech.Ns):
        s = s + Y[n] / mw[n]
    return 1.0 / s


def state_from_TPY(mech, T, P, Y, u, v, w):
    """Primitive Q (7, *S) and rho_i (Ns, *S) from T, P, Y and velocity."""
    W = mix_W(mech, Y)
    rho = P * W / (mech.Ru * T)
    rhoi = Y * rho.unsqueeze(0)
    ei = internal_energy(mech, T, rhoi)
    Q = torch.stack([rho, u, v, w, P, T, ei])
    return Q, rhoi


def prim2c(Q: torch.Tensor) -> torch.Tensor:
    """U from Q (utils.jl:71-92), for host-side initial conditions."""
    rho, u, v, w, ei = Q[0], Q[1], Q[2], Q[3], Q[6]
    return torch.stack([rho, u * rho, v * rho, w * rho, ei + 0.5 * rho * (u * u + v * v + w * w)])


# ---------------------------------------------------------------------------
# grids and metrics
# ---------------------------------------------------------------------------
@dataclass
class Grid:
    Nx: int
    Ny: int
    Nz: int
    x: torch.Tensor  # (Nz+8, Ny+8, Nx+8)
    y: torch.Tensor
    z: torch.Tensor
    metrics: List[torch.Tensor]  # 9 arrays, order dxi/dx .. dzeta/dz
    J: torch.Tensor

    @property
    def shape3(self):
        return self.x.shape


def _index_grids(Nx, Ny, Nz, dev):
    i = torch.arange(Nx + 2 * NG, dtype=torch.float64, device=dev) - NG
    j = torch.arange(Ny + 2 * NG, dtype=torch.float64, device=dev) - NG
    k = torch.arange(Nz + 2 * NG, dtype=torch.float64, device=dev) - NG
    K, Jg, I = torch.meshgrid(k, j, i, indexing="ij")
    return I, Jg, K


def metrics_from_jacobian(xd, yd, zd) -> Tuple[List[torch.Tensor], torch.Tensor]:
    """xd = (dx/dxi, dx/deta, dx/dzeta) etc. -> cofactor metrics and J
    (parse_mesh.jl:210-221)."""
    dxdxi, dxdeta, dxdzeta = xd
    dydxi, dydeta, dydzeta = yd
    dzdxi, dzdeta, dzdzeta = zd
    det = (dxdxi * (dydeta * dzdzeta - dydzeta * dzdeta)
           - dxdeta * (dydxi * dzdzeta - dydzeta * dzdxi)
           + dxdzeta * (dydxi * dzdeta - dydeta * dzdxi))
    J = 1.0 / det
    m = [dydeta * dzdzeta - dydzeta * dzdeta,
         dxdzeta * dzdeta - dxdeta * dzdzeta,
         dxdeta * dydzeta - dxdzeta * dydeta,
         dydzeta * dzdxi - dydxi * dzdzeta,
         dxdxi * dzdzeta - dxdzeta * dzdxi,
         dxdzeta * dydxi - dxdxi * dydzeta,
         dydxi * dzdeta - dydeta * dzdxi,
         dxdeta * dzdxi - dxdxi * dzdeta,
         dxdxi * dydeta - dxdeta * dydxi]
    return [t.contiguous() for t in m], J.contiguous()


def uniform_grid(Nx, Ny, Nz, Lx, Ly, Lz, dev="cpu", periodic=True) -> Grid:
    """Uniform Cartesian grid; periodic spacing L/N (nodes at 0..L-h)."""
    I, Jg, K = _index_grids(Nx, Ny, Nz, dev)
    hx = Lx / (Nx if periodic else Nx - 1)
    hy = Ly / (Ny if periodic else Ny - 1)
    hz = Lz / (Nz if periodic else Nz - 1)
    one = torch.ones_like(I)
    zero = torch.zeros_like(I)
    m, J = metrics_from_jacobian((hx * one, zero, zero), (zero, hy * one, zero),
                                 (zero, zero, hz * one))
    return Grid(Nx, Ny, Nz, I * hx, Jg * hy, K * hz, m, J)


def skewed_grid(Nx, Ny, Nz, h=1e-3, amp=0.15, dev="cpu") -> Grid:
    """Smooth curvilinear grid with all nine metric terms non-zero (for
    parity tests that must catch transposed metric indices)."""
    I, Jg, K = _index_grids(Nx, Ny, Nz, dev)
    a = 2 * math.pi / 16.0
    x = h * (I + amp * torch.sin(a * Jg) * torch.cos(a * K))
    y = h * (Jg + amp * torch.sin(a * I + a * K))
    z = h * (K + amp * torch.sin(a * I) * torch.sin(a * Jg))
    xd = (h * torch.ones_like(I), h * amp * a * torch.cos(a * Jg) * torch.cos(a * K),
          -h * amp * a * torch.sin(a * Jg) * torch.sin(a * K))
    yd = (h * amp * a * torch.cos(a * I + a * K), h * torch.ones_like(I),
          h * amp * a * torch.cos(a * I + a * K))
    zd = (h * amp * a * torch.cos(a * I) * torch.sin(a * Jg),
          h * amp * a * torch.sin(a * I) * torch.cos(a * Jg), h * torch.ones_like(I))
    m, J = metrics_from_jacobian(xd, yd, zd)
    return Grid(Nx, Ny, Nz, x, y, z, m, J)


def stretched_y_grid(Nx, Ny, Nz, Lx, Ly, Lz, beta=2.5, dev="cpu") -> Grid:
    """Uniform in x and z, tanh-stretched in y with stretch factor beta,
    clustered at the centre line y = 0:
        y(s) = Ly/2 * sign(s) * (1 - tanh(beta (1 - |s|)) / tanh(beta)),
        s = -1 + 2 eta / (Ny - 1)  (nodes on both walls)."""
    I, Jg, K = _index_grids(Nx, Ny, Nz, dev)
    hx, hz = Lx / Nx, Lz / Nz
    s = -1.0 + 2.0 * Jg / (Ny - 1)
    tb = math.tanh(beta)
    sg = torch.sign(s)
    y = 0.5 * Ly * sg * (1 - torch.tanh(beta * (1 - s.abs())) / tb)
    dyds = 0.5 * Ly * beta / tb / torch.cosh(beta * (1 - s.abs())) ** 2
    dydeta = dyds * 2.0 / (Ny - 1)
    one = torch.ones_like(I)
    zero = torch.zeros_like(I)
    m, J = metrics_from_jacobian((hx * one, zero, zero), (zero, dydeta, zero),
                                 (zero, zero, hz * one))
    return Grid(Nx, Ny, Nz, I * hx, y, K * hz, m, J)


def parse_mesh_coords(Nx=300, Ny=127, Nz=127, Lx=0.02, ymin=-0.0032, ymax=0.0032, ystar=0.0,
                      zmin=-0.0032, zmax=0.0032, zstar=0.0, alpha=4e-4, n_uniform=None,
                      dev="cpu"):
    """Interior node coordinates of the reference's mesh generator
    (parse_mesh.jl:24-50, its defaults): x uniform over the first
    n_uniform = Nx - 20 nodes with spacing Lx/(Nx-1), then growing by
    (0.5 + m/2) spacings; y and z clustered around ystar / zstar by
    y = ystar + alpha sinh(c1 (1 - s) + c2 s), s = j/(Ny-1),
    c1 = asinh((ymin - ystar)/alpha), c2 = asinh((ymax - ystar)/alpha).
    Returns padded (Nz+8, Ny+8, Nx+8) arrays with zero ghosts (f3d_mesh_ghosts
    fills them)."""
    n_uniform = Nx - 20 if n_uniform is None else n_uniform
    h = Lx / (Nx - 1)
    xs = [i * h for i in range(n_uniform)]
    for i in range(n_uniform + 1, Nx + 1):  # 1-based i, as the loop it follows
        xs.append(xs[-1] + h * (0.5 + (i - n_uniform) / 2))
    c1, c2 = math.asinh((ymin - ystar) / alpha), math.asinh((ymax - ystar) / alpha)
    c3, c4 = math.asinh((zmin - zstar) / alpha), math.asinh((zmax - zstar) / alpha)
    ys = [ystar + alpha * math.sinh(c1 * (1 - j / (Ny - 1)) + c2 * j / (Ny - 1)) for j in range(Ny)]
    zs = [zstar + alpha * math.sinh(c3 * (1 - k / (Nz - 1)) + c4 * k / (Nz - 1)) for k in range(Nz)]
    sh = (Nz + 2 * NG, Ny + 2 * NG, Nx + 2 * NG)
    x = torch.zeros(sh, dtype=torch.float64)
    y = torch.zeros(sh, dtype=torch.float64)
    z = torch.zeros(sh, dtype=torch.float64)
    it = slice(NG, NG + Nz), slice(NG, NG + Ny), slice(NG, NG + Nx)
    x[it] = torch.tensor(xs, dtype=torch.float64)[None, None, :]
    y[it] = torch.tensor(ys, dtype=torch.float64)[None, :, None]
    z[it] = torch.tensor(zs, dtype=torch.float64)[:, None, None]
    return x.to(dev), y.to(dev), z.to(dev)


def grid_from_coords(ctx, x, y, z) -> Grid:
    """Ghost nodes and metrics of padded node coordinates on the GPU
    (f3d_mesh_ghosts, f3d_mesh_metrics); x, y, z are device arrays whose
    ghosts are overwritten."""
    ctx.mesh_ghosts(x, y, z)
    m, J = ctx.mesh_metrics(x, y, z)
    return Grid(ctx.Nx, ctx.Ny, ctx.Nz, x, y, z, m, J)


# ---------------------------------------------------------------------------
# initial states
# ---------------------------------------------------------------------------
AIR_Y = {"O2": 0.233, "N2": 0.767}


def _Y_field(mech, comp: dict, like: torch.Tensor) -> torch.Tensor:
    Y = torch.zeros((mech.Ns,) + tuple(like.shape), dtype=torch.float64, device=like.device)
    for name, val in comp.items():
        Y[mech.index(name)] = val
    return Y


def random_state(grid: Grid, mech, seed=12345, shock=True, T0=1200.0, P0=101325.0,
                 noise=0.02):
    """Smooth random state with an optional pressure/density step (to drive
    the WENO and NND branches).  Species composition varies smoothly."""
    g = torch.Generator(device="cpu").manual_seed(seed)
    dev = grid.x.device
    sh = grid.shape3

    def smooth(scale):
        # a few random Fourier modes
        out = torch.zeros(sh, dtype=torch.float64, device=dev)
        I, Jg, K = _index_grids(grid.Nx, grid.Ny, grid.Nz, dev)
        for _ in range(4):
            kx, ky, kz, ph = (torch.rand(4, generator=g, dtype=torch.float64) * torch.tensor(
                [0.6, 0.6, 0.6, 6.28], dtype=torch.float64)).tolist()
            out = out + torch.sin(kx * I + ky * Jg + kz * K + ph)
        return scale * out / 4

    T = T0 * (1 + smooth(0.15))
    P = P0 * (1 + smooth(0.1))
    u = 200.0 * smooth(1.0) + 50.0
    v = 150.0 * smooth(1.0)
    w = 100.0 * smooth(1.0)
    if shock:
        I, Jg, K = _index_grids(grid.Nx, grid.Ny, grid.Nz, dev)
        step = (I + 0.5 * Jg - 0.3 * K > 0.45 * grid.Nx).to(torch.float64)
        P = P * (1 + 1.5 * step)
        T = T * (1 + 0.4 * step)
    Y = torch.zeros((mech.Ns,) + tuple(sh), dtype=torch.float64, device=dev)
    for n in range(mech.Ns):
        Y[n] = 0.05 + 0.04 * smooth(1.0) + 0.02 * n / max(1, mech.Ns)
    Y = Y / Y.sum(0, keepdim=True)
    Q, rhoi = state_from_TPY(mech, T, P, Y, u, v, w)
    return Q.contiguous(), rhoi.contiguous()


@dataclass
class Case:
    name: str
    grid: Grid
    Q: torch.Tensor
    rhoi: torch.Tensor
    bc: Tuple[int, ...]
    dt: float
    inflow_Q: Optional[list] = None
    inflow_rhoi: Optional[list] = None


def sound_speed(Q):
    gam = Q[4] / Q[6] + 1
    return torch.sqrt(gam * Q[4] / Q[0])


def cfl_dt(grid: Grid, Q: torch.Tensor, cfl: float) -> float:
    """dt from the inviscid contravariant speeds |U_xi| + c |grad xi| (J weighting)."""
    c = sound_speed(Q)
    J = grid.J
    lam = torch.zeros_like(c)
    for d in range(3):
        ax, ay, az = grid.metrics[3 * d], grid.metrics[3 * d + 1], grid.metrics[3 * d + 2]
        un = (ax * Q[1] + ay * Q[2] + az * Q[3]) * J
        gn = torch.sqrt(ax * ax + ay * ay + az * az) * J
        lam = lam + un.abs() + c * gn
    s = (slice(NG, -NG),) * 3
    return float(cfl / lam[s].max())


def mixing_layer(N=(256, 256, 256), mech=None, seed=12345, dev="cuda", beta=2.5,
                 L=0.01) -> Case:
    """Shock-laden mixing layer (air, non-reacting): Ma 1.5 upper stream over
    a Ma 0.3 lower stream at equal pressure with density ratio 2 (T 600 K /
    300 K), tanh profile of thickness L/40 across y = 0, a normal shock at
    x = L/3 imposed on the supersonic stream by the Rankine-Hugoniot jump
    (gamma = 1.4), 1% seeded white-noise velocity perturbations.  x: fixed
    inflow ghosts / zero-gradient outflow, y: zero gradient, z: periodic."""
    from . import abi
    Nx, Ny, Nz = N
    grid = stretched_y_grid(Nx, Ny, Nz, L, L, L, beta=beta, dev=dev)
    P0 = 101325.0
    T_up, T_lo = 600.0, 300.0
    gam = 1.4
    W = 1.0 / (0.233 / mech.mw[mech.index("O2")] + 0.767 / mech.mw[mech.index("N2")])
    Rg = mech.Ru / W
    c_up, c_lo = math.sqrt(gam * Rg * T_up), math.sqrt(gam * Rg * T_lo)
    U_up, U_lo = 1.5 * c_up, 0.3 * c_lo
    delta = L / 40
    th = 0.5 * (1 + torch.tanh(grid.y / delta))  # 1 upper, 0 lower
    T = T_lo + (T_up - T_lo) * th
    u = U_lo + (U_up - U_lo) * th
    P = torch.full_like(T, P0)
    # normal shock on the supersonic part
    c = torch.sqrt(gam * Rg * T)
    M1 = u / c
    post = (grid.x > L / 3) & (M1 > 1.0)
    M1s = torch.clamp(M1, min=1.0)
    p_ratio = 1 + 2 * gam / (gam + 1) * (M1s ** 2 - 1)
    r_ratio = (gam + 1) * M1s ** 2 / ((gam - 1) * M1s ** 2 + 2)
    P = torch.where(post, P * p_ratio, P)
    T = torch.where(post, T * p_ratio / r_ratio, T)
    u = torch.where(post, u / r_ratio, u)
    g = torch.Generator(device=dev).manual_seed(seed)
    amp = 0.01 * U_up
    u = u + amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    v = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    w = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    Y = _Y_field(mech, AIR_Y, T)
    Q, rhoi = state_from_TPY(mech, T, P, Y, u, v, w)
    dt = cfl_dt(grid, Q, 0.5)
    bc = (abi.BC_KEEP, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_PERIODIC, abi.BC_PERIODIC)
    return Case("mixing_layer", grid, Q.contiguous(), rhoi.contiguous(), bc, dt)


def reacting_mixing_layer(N=(256, 256, 256), mech=None, seed=12345, dev="cuda", beta=2.5,
                          dt=2e-7, cfl=0.5, P_atm=3.0, T_air=1200.0, T_fuel=1000.0) -> Case:
    """Reacting H2/air shock-laden mixing layer (the 9-species H2/O2 mechanism
    with N2): a Ma 1.5 hot-air upper stream (T 1200 K) over a Ma 0.3 hydrogen
    lower stream (T 1000 K) at 3 atm, mixture fraction Z = 1 - th(y) with the
    tanh profile of thickness L/40 (Y = Z Y_H2 + (1 - Z) Y_air, T linear in Z,
    the states of the H2/O2 reactor distribution), a normal shock at x = L/3 on
    the supersonic stream (Rankine-Hugoniot, gamma 1.4; 7.4 atm, 1580 K behind
    it: the radical pool builds up during the run, ignition is tens of
    microseconds away), 1% seeded white-noise velocity.  The box edge L is
    chosen so that the step `dt` is at CFL `cfl`, so that Strang half-steps of
    chemistry are dt/2.  Boundaries as mixing_layer.  (With 1500 K air the
    post-shock band ignites within the run and the backward-Euler step jumps
    to the burnt root in some cells, which makes the field chaotic: FP32
    rounding alone then moves T by 10^3 K there.)"""
    from . import abi
    Nx, Ny, Nz = N

    def build(L):
        grid = stretched_y_grid(Nx, Ny, Nz, L, L, L, beta=beta, dev=dev)
        gam = 1.4
        W_air = 1.0 / (0.233 / mech.mw[mech.index("O2")] + 0.767 / mech.mw[mech.index("N2")])
        W_h2 = float(mech.mw[mech.index("H2")])
        c_up = math.sqrt(gam * mech.Ru / W_air * T_air)
        c_lo = math.sqrt(gam * mech.Ru / W_h2 * T_fuel)
        U_up, U_lo = 1.5 * c_up, 0.3 * c_lo
        delta = L / 40
        th = 0.5 * (1 + torch.tanh(grid.y / delta))  # 1 air (upper), 0 fuel (lower)
        Z = 1.0 - th
        T = T_fuel * Z + T_air * th
        u = U_lo + (U_up - U_lo) * th
        P = torch.full_like(T, 101325.0 * P_atm)
        c = c_up * torch.sqrt(T / T_air)
        M1 = u / c
        post = (grid.x > L / 3) & (M1 > 1.0) & (th > 0.5)
        M1s = torch.clamp(M1, min=1.0)
        p_ratio = 1 + 2 * gam / (gam + 1) * (M1s ** 2 - 1)
        r_ratio = (gam + 1) * M1s ** 2 / ((gam - 1) * M1s ** 2 + 2)
        P = torch.where(post, P * p_ratio, P)
        T = torch.where(post, T * p_ratio / r_ratio, T)
        u = torch.where(post, u / r_ratio, u)
        g = torch.Generator(device=dev).manual_seed(seed)
        amp = 0.01 * U_up
        u = u + amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
        v = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
        w = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
        Y = torch.zeros((mech.Ns,) + tuple(T.shape), dtype=torch.float64, device=dev)
        Y[mech.index("H2")] = Z
        Y[mech.index("O2")] = 0.233 * th
        Y[mech.index("N2")] = 0.767 * th
        Q, rhoi = state_from_TPY(mech, T, P, Y, u, v, w)
        return grid, Q, rhoi

    L0 = 0.05
    grid, Q, rhoi = build(L0)
    L = L0 * dt / cfl_dt(grid, Q, cfl)
    grid, Q, rhoi = build(L)
    bc = (abi.BC_KEEP, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_PERIODIC, abi.BC_PERIODIC)
    return Case("reacting_mixing_layer", grid, Q.contiguous(), rhoi.contiguous(), bc, dt)


def reacting_jet(N=(512, 256, 256), mech=None, seed=12345, dev="cuda", beta=1.0, D=0.01,
                 Ma=1.5, T_jet=5000.0, T_air=300.0, P0=3596.0, noise=0.02, cfl=0.5,
                 dt=None) -> Case:
    """Supersonic hot-air jet, the coupled reacting workload of 5-species air:
    a round jet of undissociated air (O2 0.233, N2 0.767) at T_jet and Mach
    Ma of its own sound speed along x through still air at T_air, both at
    P0; the jet fills r < D/2 around the axis (y = 0, z = Lz/2) with a tanh
    edge of thickness D/20, over the whole domain at t = 0 and held in the
    x-min ghost cells (inflow); seeded white-noise velocity perturbations of
    noise * U_jet inside the jet.  Grid Lx = 20 D, Ly = Lz = 10 D,
    tanh-stretched in y around the axis (beta), uniform in x and z.  BCs:
    x fixed inflow ghosts / zero-gradient outflow, y and z zero gradient.
    The time step is the CFL step unless `dt` fixes it (a surrogate trained
    for one chemistry step needs a fixed one; it must not exceed the CFL
    step).  At 512x256x256 and the fixed 1.2e-8 s step, beta = 2 (the
    smallest y cells on the axis, convective CFL 0.44) runs away on the jet
    axis after ~15 steps (T 8200 K, then NaN at step 20, in FP64); beta = 1
    (CFL 0.23) holds over 50 steps."""
    from . import abi
    Nx, Ny, Nz = N
    grid = stretched_y_grid(Nx, Ny, Nz, 20 * D, 10 * D, 10 * D, beta=beta, dev=dev)
    r = torch.sqrt(grid.y ** 2 + (grid.z - 5 * D) ** 2)
    th = 0.5 * (1 - torch.tanh((r - 0.5 * D) / (D / 20)))  # 1 inside the jet
    Y = _Y_field(mech, AIR_Y, grid.x)
    W = 1.0 / (0.233 / mech.mw[mech.index("O2")] + 0.767 / mech.mw[mech.index("N2")])
    Rg = mech.Ru / W
    Tj = torch.tensor([T_jet], dtype=torch.float64, device=dev)
    cp = float(cp_mix(mech, Tj, _Y_field(mech, AIR_Y, Tj))[0])
    gam = cp / (cp - Rg)
    U_jet = Ma * math.sqrt(gam * Rg * T_jet)
    T = T_air + (T_jet - T_air) * th
    g = torch.Generator(device=dev).manual_seed(seed)
    amp = noise * U_jet * th
    u = U_jet * th + amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    v = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    w = amp * torch.randn(T.shape, generator=g, device=dev, dtype=torch.float64)
    P = torch.full_like(T, P0)
    Q, rhoi = state_from_TPY(mech, T, P, Y, u, v, w)
    dt_cfl = cfl_dt(grid, Q, cfl)
    if dt is None:
        dt = dt_cfl
    elif dt > dt_cfl:
        raise ValueError(f"reacting_jet: dt {dt:g} exceeds the CFL step {dt_cfl:g}")
    bc = (abi.BC_KEEP, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY, abi.BC_COPY)
    return Case("reacting_jet", grid, Q.contiguous(), rhoi.contiguous(), bc, dt)


def tgv_gas(Ma=0.1, Re=1600.0, T0=300.0, P0=101325.0, R_gas=287.0, gamma=1.4, Pr=0.72):
    """The calorically perfect gas of the Taylor-Green case (gamma 1.4, Pr 0.72)
    with the constant viscosity that gives Re = rho0 V0 L / mu at L = 1 m."""
    from . import mech as mm
    c0 = math.sqrt(gamma * R_gas * T0)
    rho0 = P0 / (R_gas * T0)
    return mm.perfect_gas(R_gas=R_gas, gamma=gamma, mu=rho0 * Ma * c0 / Re, Pr=Pr)


def taylor_green(N=64, mech=None, Ma=0.1, Re=1600.0, T0=300.0, cfl=0.5, dev="cpu") -> Case:
    """Taylor-Green vortex on the periodic [0, 2 pi L)^3 box with V0 = Ma c0,
    rho0 from p0 = rho0 R T0, mu from Re = rho0 V0 L / mu (L = 1 m); mech
    None: tgv_gas(Ma, Re, T0)."""
    from . import abi
    if mech is None:
        mech = tgv_gas(Ma, Re, T0)
    L = 1.0
    grid = uniform_grid(N, N, N, 2 * math.pi * L, 2 * math.pi * L, 2 * math.pi * L, dev=dev)
    W = float(mech.mw[0])
    Rg = mech.Ru / W
    gam = 1.4
    c0 = math.sqrt(gam * Rg * T0)
    V0 = Ma * c0
    P0 = 101325.0
    rho0 = P0 / (Rg * T0)
    x, y, z = grid.x / L, grid.y / L, grid.z / L
    u = V0 * torch.sin(x) * torch.cos(y) * torch.cos(z)
    v = -V0 * torch.cos(x) * torch.sin(y) * torch.cos(z)
    w = torch.zeros_like(u)
    P = P0 + rho0 * V0 ** 2 / 16 * (torch.cos(2 * x) + torch.cos(2 * y)) * (torch.cos(2 * z) + 2)
    T = torch.full_like(P, T0)
    Y = torch.ones((1,) + tuple(P.shape), dtype=torch.float64, device=P.device)
    Q, rhoi = state_from_TPY(mech, T, P, Y, u, v, w)
    dt = cfl_dt(grid, Q, cfl)
    bc = (abi.BC_PERIODIC,) * 6
    return Case("taylor_green", grid, Q.contiguous(), rhoi.contiguous(), bc, dt)


def chem_cells(N=(128, 128, 128), mech=None, seed=12345, dev="cpu",
               T_range=(900.0, 2500.0), P_atm=(1.0, 10.0), z_pure=0.0):
    """Independent H2/O2/N2 reactors: T ~ U[900, 2500] K, P ~ U[1, 10] atm,
    mixture fraction Z ~ Beta(2, 2); Y = Z Y_fuel(H2) + (1 - Z) Y_ox(O2 0.233,
    N2 0.767).  With z_pure > 0 a fraction z_pure of the cells is pure
    oxidizer (Z = 0) and another z_pure pure fuel (Z = 1), the two streams of
    a non-premixed flow.  Returns padded Q, rho_i (ghost cells hold copies of
    the interior pattern and are never read by the chemistry kernels)."""
    Nx, Ny, Nz = N
    g = torch.Generator(device="cpu").manual_seed(seed)
    sh = (Nz + 2 * NG, Ny + 2 * NG, Nx + 2 * NG)
    T = T_range[0] + (T_range[1] - T_range[0]) * torch.rand(sh, generator=g, dtype=torch.float64)
    P = 101325.0 * (P_atm[0] + (P_atm[1] - P_atm[0]) * torch.rand(sh, generator=g,
                                                                   dtype=torch.float64))
    beta = torch.distributions.Beta(torch.tensor(2.0, dtype=torch.float64),
                                    torch.tensor(2.0, dtype=torch.float64))
    torch.manual_seed(seed)
    Z = beta.sample(sh)
    if z_pure > 0:
        u = torch.rand(sh, generator=g, dtype=torch.float64)
        Z = torch.where(u < z_pure, torch.zeros_like(Z), Z)
        Z = torch.where((u >= z_pure) & (u < 2 * z_pure), torch.ones_like(Z), Z)
    Y = torch.zeros((mech.Ns,) + sh, dtype=torch.float64)
    Y[mech.index("H2")] = Z
    Y[mech.index("O2")] = (1 - Z) * 0.233
    Y[mech.index("N2")] = (1 - Z) * 0.767
    zero = torch.zeros(sh, dtype=torch.float64)
    Q, rhoi = state_from_TPY(mech, T, P, Y, zero, zero, zero)
    return Q.to(dev).contiguous(), rhoi.to(dev).contiguous()


def chem_cells_air(N=(16, 8, 8), mech=None, seed=4, T_range=(3000.0, 6000.0),
                   P_atm=(0.03, 1.0)):
    """Dissociating 5-species air reactors: T ~ U[3000, 6000] K, P ~ U[0.03,
    1] atm, composition air with a random partial dissociation."""
    Nx, Ny, Nz = N
    g = torch.Generator(device="cpu").manual_seed(seed)
    sh = (Nz + 2 * NG, Ny + 2 * NG, Nx + 2 * NG)
    T = T_range[0] + (T_range[1] - T_range[0]) * torch.rand(sh, generator=g, dtype=torch.float64)
    P = 101325.0 * (P_atm[0] + (P_atm[1] - P_atm[0]) * torch.rand(sh, generator=g,
                                                                   dtype=torch.float64))
    a = 0.3 * torch.rand(sh, generator=g, dtype=torch.float64)  # O2 dissociation fraction
    b = 0.05 * torch.rand(sh, generator=g, dtype=torch.float64)  # N2 -> NO / N
    Y = torch.zeros((mech.Ns,) + sh, dtype=torch.float64)
    Y[mech.index("O2")] = 0.233 * (1 - a)
    Y[mech.index("O")] = 0.233 * a
    Y[mech.index("N2")] = 0.767 * (1 - b)
    Y[mech.index("NO")] = 0.767 * b * 0.5
    Y[mech.index("N")] = 0.767 * b * 0.5
    zero = torch.zeros(sh, dtype=torch.float64)
    return state_from_TPY(mech, T, P, Y, zero, zero, zero)
