"""Problem generators (initial conditions) for the benchmark configurations.

Host-side setup, run once per problem in FP64; the same array feeds the GPU
path and the oracle. Cell centres follow State::get_pos (State.hpp:123-134):
x_i = i*dx + (-(ng - 0.5)*dx) + loc.x, ghost cells included, and primitive
states are converted with prim_to_cons's grouping (Eos.hpp:81-99). Each
generator sets every cell, ghosts included, like the reference's generators
(e.g. ShockTube2d.cpp:26-47); the caller fills ghosts afterwards.

Random perturbations use a counter-based generator (splitmix64 of
seed + counter), mirroring the reference's per-cell seeding idea
(HillierSnowKelvinHelmholtz.cpp:66, 103) without its YAKL dependency.
"""
from __future__ import annotations

import numpy as np

from . import Config

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def splitmix64(x: np.ndarray) -> np.ndarray:
    """Steele, Lea & Flood (2014) SplitMix64 finaliser of the counter x."""
    with np.errstate(over="ignore"):
        z = (np.asarray(x, dtype=np.uint64) + np.uint64(0x9E3779B97F4A7C15))
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform01(seed: int, counter: np.ndarray) -> np.ndarray:
    """Uniform doubles in [0, 1) from the top 53 bits of splitmix64(seed + counter)."""
    with np.errstate(over="ignore"):
        z = splitmix64(np.uint64(seed) + np.asarray(counter, dtype=np.uint64))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def cell_centres(cfg: Config):
    """State::get_pos for every cell (ghosts included): x[i], y[j]."""
    ghost_offset = -(cfg.ng - 0.5) * cfg.dx
    i = np.arange(cfg.xc, dtype=np.float64)
    x = i * cfg.dx + ghost_offset + cfg.x_start
    if cfg.num_dim > 1:
        j = np.arange(cfg.yc, dtype=np.float64)
        y = j * cfg.dx + ghost_offset + cfg.y_start
    else:
        y = np.zeros(1)
    return x, y


def prim_to_cons(cfg: Config, rho, vx, vy, p) -> np.ndarray:
    """prim_to_cons (Eos.hpp:81-99) on whole fields, same grouping; tracer planes zero."""
    q = np.zeros(cfg.shape)
    q[0, 0] = rho
    q[1, 0] = rho * vx
    v2 = vx * vx
    if cfg.num_dim > 1:
        q[2, 0] = rho * vy
        v2 = v2 + vy * vy
    e_kin = 0.5 * rho * v2
    e_int = p / (cfg.gamma - 1.0)
    q[1 + cfg.num_dim, 0] = e_int + e_kin
    return q


def shock_tube_2d(cfg: Config) -> np.ndarray:
    """Sod tube along x (ShockTube2d.cpp:10-49): (rho, p) = (1, 1) for x < 0.5, else (0.125, 0.1)."""
    x, _ = cell_centres(cfg)
    left = x < 0.5
    shape = (cfg.yc, cfg.xc)
    rho = np.broadcast_to(np.where(left, 1.0, 0.125), shape).copy()
    p = np.broadcast_to(np.where(left, 1.0, 0.1), shape).copy()
    zero = np.zeros(shape)
    return prim_to_cons(cfg, rho, zero, zero, p)


def kh_band(cfg: Config, seed: int = 20261202, amp: float = 0.01, rho_perturb: np.ndarray | None = None):
    """Kelvin-Helmholtz shear band on [0,1]^2: p = 2.5; |y - 0.5| < 0.25 has rho = 2, vx = +0.5,
    elsewhere rho = 1, vx = -0.5; vx and vy get U(-amp, amp) from the per-cell counter
    2*(j*xc + i) (vx) and 2*(j*xc + i) + 1 (vy). `rho_perturb` (interior-shaped) is added
    to the interior density (the adjoint's control field). Tracer fields, if any, are a dye
    of mass fraction 1 inside the band and 0 outside."""
    x, y = cell_centres(cfg)
    Y = np.broadcast_to(y[:, None], (cfg.yc, cfg.xc))
    band = np.abs(Y - 0.5) < 0.25
    rho = np.where(band, 2.0, 1.0)
    vx = np.where(band, 0.5, -0.5)
    idx = np.arange(cfg.yc * cfg.xc, dtype=np.uint64).reshape(cfg.yc, cfg.xc)
    vx = vx + amp * (2.0 * uniform01(seed, 2 * idx) - 1.0)
    vy = amp * (2.0 * uniform01(seed, 2 * idx + 1) - 1.0)
    if rho_perturb is not None:
        g = cfg.ng
        rho = rho.copy()
        rho[g:g + cfg.ny, g:g + cfg.nx] += rho_perturb
    p = np.full((cfg.yc, cfg.xc), 2.5)
    q = prim_to_cons(cfg, rho, vx, vy, p)
    for k in range(cfg.num_tracers):  # passive dye marking the band
        q[4 + k, 0] = rho * np.where(band, 1.0, 0.0)
    return q


def sedov(cfg: Config, energy: float = 1.0, radius_cells: float = 3.5, p_ambient: float = 1e-5):
    """Sedov blast on [-1,1]^2: rho = 1, v = 0, p = p_ambient, and `energy` spread uniformly over
    the cells whose centre lies within radius_cells*dx of the origin: p = (gamma-1)*E/(N*dx^2)."""
    x, y = cell_centres(cfg)
    X = np.broadcast_to(x[None, :], (cfg.yc, cfg.xc))
    Y = np.broadcast_to(y[:, None], (cfg.yc, cfg.xc))
    inside = (X * X + Y * Y) < (radius_cells * cfg.dx) ** 2
    g = cfg.ng
    n_in = int(inside[g:g + cfg.ny, g:g + cfg.nx].sum())
    p_in = (cfg.gamma - 1.0) * energy / (n_in * cfg.dx * cfg.dx)
    p = np.where(inside, p_in, p_ambient)
    rho = np.ones((cfg.yc, cfg.xc))
    zero = np.zeros((cfg.yc, cfg.xc))
    return prim_to_cons(cfg, rho, zero, zero, p)


def cell_centres_3d(cfg: Config):
    """State::get_pos of every cell of a 3-D grid: x[i], y[j], z[k]."""
    off = -(cfg.ng - 0.5) * cfg.dx
    x = np.arange(cfg.xc, dtype=np.float64) * cfg.dx + off + cfg.x_start
    y = np.arange(cfg.yc, dtype=np.float64) * cfg.dx + off + cfg.y_start
    z = np.arange(cfg.zc, dtype=np.float64) * cfg.dx + off + cfg.z_start
    return x, y, z


def prim_to_cons_3d(cfg: Config, rho, vx, vy, vz, p) -> np.ndarray:
    """prim_to_cons (Eos.hpp:81-99) for 3-D fields shaped [z][y][x], same grouping."""
    q = np.zeros(cfg.shape)
    q[0] = rho
    q[1] = rho * vx
    q[2] = rho * vy
    q[3] = rho * vz
    v2 = vx * vx
    v2 = v2 + vy * vy
    v2 = v2 + vz * vz
    q[4] = p / (cfg.gamma - 1.0) + 0.5 * rho * v2
    return q


def blast_3d(cfg: Config, seed: int = 7, amp: float = 0.05, p_ratio: float = 10.0, radius: float = 0.2):
    """Smooth seeded 3-D state (ghosts included): density, velocity and pressure are sums of low
    Fourier modes with seeded phases, plus a spherical over-pressure of `p_ratio` within `radius` of
    the domain centre."""
    x, y, z = cell_centres_3d(cfg)
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    L = cfg.x_dim
    ph = 2 * np.pi * uniform01(seed, np.arange(12, dtype=np.uint64))

    def mode(a, b, c, k):
        return np.sin(2 * np.pi * (a * X + b * Y + c * Z) / L + ph[k])

    rho = 1.0 + amp * (mode(1, 0, 1, 0) + mode(0, 1, 1, 1))
    vx = amp * mode(0, 1, 1, 2)
    vy = amp * mode(1, 0, 1, 3)
    vz = amp * mode(1, 1, 0, 4)
    cx = cfg.x_start + 0.5 * cfg.nx * cfg.dx
    cy = cfg.y_start + 0.5 * cfg.ny * cfg.dx
    cz = cfg.z_start + 0.5 * cfg.nz * cfg.dx
    r2 = (X - cx) ** 2 + (Y - cy) ** 2 + (Z - cz) ** 2
    p = (1.0 + amp * mode(1, 1, 1, 5)) * np.where(r2 < (radius * L) ** 2, p_ratio, 1.0)
    return prim_to_cons_3d(cfg, rho, vx, vy, vz, p)


# The benchmark workloads (Sod 256^2, KH 2048^2, Sedov 4096^2), with the reference's key names.
def sod_config() -> Config:
    return Config(num_dim=2, nx=256, ny=256, ng=2, x_dim=1.0, bc=("zerograd",) * 4, reconstruction="muscl",
                  slope_limiter="minmod", riemann_solver="hllc", time_stepper="rk2", gamma=1.4, max_cfl=0.4,
                  max_time=0.2, output_delta=0.1)


def kh_config(n: int = 2048) -> Config:
    return Config(num_dim=2, nx=n, ny=n, ng=3, x_dim=1.0, bc=("periodic",) * 4, reconstruction="weno5z",
                  riemann_solver="hllc", time_stepper="ssprk3", gamma=1.4, max_cfl=0.4, max_time=1.0,
                  output_delta=0.1)


def sedov_config(n: int = 4096) -> Config:
    return Config(num_dim=2, nx=n, ny=n, ng=3, x_dim=2.0, x_start=-1.0, y_start=-1.0, bc=("zerograd",) * 4,
                  reconstruction="ppm", slope_limiter="monotonizedcentral", riemann_solver="hll",
                  time_stepper="ssprk3", gamma=1.4, max_cfl=0.3, max_time=0.05, output_delta=0.1)


def riemann_batch(nprob: int = 4096, seed: int = 4096, n_min: int = 128, n_max: int = 4096, ng: int = 3,
                  gamma: float = 1.4):
    """Ensemble of 1-D Riemann problems on [0, 1]: lengths log-uniform in [n_min, n_max], interface at
    x = 0.5; left/right rho log-uniform in [0.1, 10], p log-uniform in [0.01, 100], vx uniform in
    [-1, 1], drawn per problem from the counter stream of `seed`. Returns (lengths, packed Q, states)
    with Q the back-to-back [3][n + 2 ng] blocks, every cell (ghosts included) set."""
    k = np.arange(nprob, dtype=np.uint64) * np.uint64(8)
    u = [uniform01(seed, k + np.uint64(j)) for j in range(7)]
    lengths = np.floor(np.exp(np.log(n_min) + u[0] * (np.log(n_max + 1) - np.log(n_min)))).astype(np.int64)
    lengths = np.clip(lengths, n_min, n_max)
    rho_l = 10.0 ** (-1.0 + 2.0 * u[1])
    rho_r = 10.0 ** (-1.0 + 2.0 * u[2])
    p_l = 10.0 ** (-2.0 + 4.0 * u[3])
    p_r = 10.0 ** (-2.0 + 4.0 * u[4])
    v_l = 2.0 * u[5] - 1.0
    v_r = 2.0 * u[6] - 1.0
    blocks = []
    for i in range(nprob):
        n = int(lengths[i])
        dx = 1.0 / n
        x = np.arange(n + 2 * ng) * dx + (-(ng - 0.5) * dx)
        left = x < 0.5
        rho = np.where(left, rho_l[i], rho_r[i])
        v = np.where(left, v_l[i], v_r[i])
        p = np.where(left, p_l[i], p_r[i])
        q = np.empty((3, n + 2 * ng))
        q[0] = rho
        q[1] = rho * v
        q[2] = p / (gamma - 1.0) + 0.5 * rho * (v * v)
        blocks.append(q.ravel())
    states = np.stack([rho_l, v_l, p_l, rho_r, v_r, p_r], axis=1)
    return lengths.astype(np.int32), np.concatenate(blocks), states
