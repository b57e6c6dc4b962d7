"""Pins for the oracle's full step, ghost fill and dt (CPU only).

Invariants and closed forms that a transcription mistake in the sequencing
(stage coefficients, flux differences, ghost index maps, the dt clamp) would
break: the exact Sod solution, exact conservation, y-invariance of a 2-D
tube (and equality with the 1-D run), preservation of a uniform state, a
brute-force ghost-fill map, and the dt clamp's nextafter loop.
"""
import itertools

import numpy as np
import pytest

from mosscap_amd import Config
from mosscap_amd import problems
from tests.exact_riemann import sample

G = 1.4


def _rel_l1(a, b, cfg):
    g = cfg.ng
    sl = (slice(None), 0, slice(g, g + cfg.ny), slice(g, g + cfg.nx)) if cfg.num_dim > 1 else \
        (slice(None), 0, 0, slice(g, g + cfg.nx))
    a, b = a[sl], b[sl]
    return np.abs(a - b).reshape(a.shape[0], -1).sum(1) / np.abs(b).reshape(b.shape[0], -1).sum(1)


def sod_1d(n=256, recon="muscl", lim="minmod", solver="hllc", stepper="rk2", ng=2):
    return Config(num_dim=1, nx=n, ng=ng, bc=("zerograd", "zerograd", "wall", "wall"), reconstruction=recon,
                  slope_limiter=lim, riemann_solver=solver, time_stepper=stepper, max_cfl=0.4, max_time=0.2,
                  output_delta=0.1)


def sod_ic_1d(cfg):
    x, _ = problems.cell_centres(cfg)
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)[None, :]
    p = np.where(left, 1.0, 0.1)[None, :]
    return problems.prim_to_cons(cfg, rho, 0 * rho, None, p)


@pytest.mark.parametrize("recon,lim,solver,stepper,tol", [
    ("muscl", "minmod", "hllc", "rk2", 0.012),
    ("weno5z", "minmod", "hllc", "ssprk3", 0.008),
    ("ppm", "monotonizedcentral", "hll", "ssprk3", 0.012),
    ("muscl", "vanleer", "rusanov", "ssprk4", 0.015),
])
def test_sod_matches_exact_solution(orc, recon, lim, solver, stepper, tol):
    cfg = sod_1d(recon=recon, lim=lim, solver=solver, stepper=stepper, ng=3)
    sim = orc.OracleSim(cfg)
    sim.set_q(sod_ic_1d(cfg))
    sim.fill_bcs()
    t, _, dts = sim.run()
    assert t == pytest.approx(0.2, abs=0)  or t >= 0.2
    assert t - dts[-1] < 0.2
    q = sim.get_q()[:, 0, 0, cfg.ng:cfg.ng + cfg.nx]
    x, _ = problems.cell_centres(cfg)
    ex = sample(x[cfg.ng:cfg.ng + cfg.nx], t, 0.5, 1.0, 0.0, 1.0, 0.125, 0.0, 0.1)
    err = np.abs(q[0] - ex[0]).mean()
    assert err < tol, err


def test_sod_2d_rows_identical_and_equal_1d(orc):
    cfg2 = problems.sod_config()
    cfg2.nx = cfg2.ny = 64
    sim2 = orc.OracleSim(cfg2)
    sim2.set_q(problems.shock_tube_2d(cfg2))
    sim2.fill_bcs()
    t2, _, _ = sim2.run(max_steps=40)
    q2 = sim2.get_q()
    g = cfg2.ng
    rows = q2[:, 0, g:g + cfg2.ny, :]
    assert (rows == rows[:, :1, :]).all(), "2-D Sod rows differ"
    assert (q2[2] == 0.0).all()
    cfg1 = sod_1d(n=64)
    sim1 = orc.OracleSim(cfg1)
    sim1.set_q(sod_ic_1d(cfg1))
    sim1.fill_bcs()
    t1, _, _ = sim1.run(max_steps=40)
    q1 = sim1.get_q()
    assert t1 == t2
    np.testing.assert_array_equal(q1[[0, 1, 2], 0, 0, :], q2[[0, 1, 3], 0, g, :])


@pytest.mark.parametrize("recon,lim,solver,stepper", [
    ("weno5z", "minmod", "hllc", "ssprk3"), ("ppm", "vanleer", "hll", "rk2"),
    ("muscl", "monotonizedcentral", "rusanov", "ssprk4"), ("fog", "minmod", "hllc", "ssprk3")])
def test_periodic_conservation(orc, recon, lim, solver, stepper):
    cfg = problems.kh_config(48)
    cfg.reconstruction, cfg.slope_limiter, cfg.riemann_solver, cfg.time_stepper = recon, lim, solver, stepper
    sim = orc.OracleSim(cfg)
    q0 = problems.kh_band(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    g = cfg.ng
    tot0 = sim.get_q()[:, 0, g:-g, g:-g].sum(axis=(1, 2))
    sim.run(max_steps=10)
    tot = sim.get_q()[:, 0, g:-g, g:-g].sum(axis=(1, 2))
    np.testing.assert_allclose(tot, tot0, rtol=1e-13, atol=1e-13 * np.abs(tot0).max())


def test_uniform_state_preserved(orc):
    for recon, lim, solver, stepper in itertools.product(["muscl", "ppm", "weno5z"], ["minmod"],
                                                         ["rusanov", "hll", "hllc"], ["rk2", "ssprk3"]):
        cfg = Config(num_dim=2, nx=12, ny=10, ng=3, bc=("periodic",) * 4, reconstruction=recon, slope_limiter=lim,
                     riemann_solver=solver, time_stepper=stepper, max_cfl=0.5)
        shape = (cfg.yc, cfg.xc)
        q = problems.prim_to_cons(cfg, np.full(shape, 1.3), np.full(shape, 0.4), np.full(shape, -0.2),
                                  np.full(shape, 2.0))
        sim = orc.OracleSim(cfg)
        sim.set_q(q)
        sim.run(max_steps=3)
        np.testing.assert_allclose(sim.get_q(), q, rtol=2e-15, atol=0)


def _bc_brute(q, cfg):
    """Ghost fill from the physical meaning of each boundary: mirror, periodic
    wrap, edge copy or constant; wall/diode adjust the normal momentum. x is
    applied over all rows first, then y over all columns."""
    q = q.copy()
    g, nd = cfg.ng, cfg.num_dim
    n_x, n_y = cfg.xc, cfg.yc

    def src(k, n, kind):
        start = k < g
        if kind == "periodic":
            return k + (n - 2 * g) if start else k - (n - 2 * g)
        if kind == "zerograd":
            return g if start else n - 1 - g
        # mirror across the boundary face
        return 2 * g - 1 - k if start else 2 * (n - g) - 1 - k

    def fix(v, kind, start, axis, cv):
        v = v.copy()
        m = 1 + axis
        if kind == "wall":
            v[m] = -v[m]
        elif kind == "symmetricoutflowdiode":
            pre = v[m]
            v[m] = min(pre, 0.0) if start else max(pre, 0.0)
            v[1 + nd] -= (v[m] - pre) ** 2 / v[0]
        elif kind == "constant":
            v[:] = cv[:len(v)]
        return v

    for i in [*range(g), *range(n_x - g, n_x)]:
        kind = cfg.bc[0] if i < g else cfg.bc[1]
        cv = np.array(cfg.bc_const[0] if i < g else cfg.bc_const[1])
        for j in range(n_y):
            q[:, 0, j, i] = fix(q[:, 0, j, src(i, n_x, kind)], kind, i < g, 0, cv)
    if nd > 1:
        for j in [*range(g), *range(n_y - g, n_y)]:
            kind = cfg.bc[2] if j < g else cfg.bc[3]
            cv = np.array(cfg.bc_const[2] if j < g else cfg.bc_const[3])
            for i in range(n_x):
                q[:, 0, j, i] = fix(q[:, 0, src(j, n_y, kind), i], kind, j < g, 1, cv)
    return q


KINDS = ["wall", "periodic", "symmetric", "symmetricoutflowdiode", "zerograd", "constant"]


@pytest.mark.parametrize("ng", [2, 3])
def test_ghost_fill_brute_force(orc, ng):
    rng = np.random.default_rng(ng)
    pairs = [(a, b) for a in KINDS for b in KINDS if (a == "periodic") == (b == "periodic")]
    for k, (xs, xe) in enumerate(pairs):
        ys, ye = pairs[(k * 7 + 3) % len(pairs)]
        cfg = Config(num_dim=2, nx=6, ny=5, ng=ng, bc=(xs, xe, ys, ye),
                     bc_const=tuple(tuple(rng.uniform(0.5, 2.0, 5)) for _ in range(4)))
        q = rng.uniform(0.5, 2.0, size=cfg.shape)
        q[1:3] -= 1.25
        sim = orc.OracleSim(cfg)
        sim.set_q(q)
        sim.fill_bcs()
        np.testing.assert_allclose(sim.get_q(), _bc_brute(q, cfg), rtol=1e-15, atol=0,
                                   err_msg=str(cfg.bc))


def test_dt_closed_form_and_clamp(orc):
    cfg = Config(num_dim=2, nx=8, ny=8, ng=2, bc=("periodic",) * 4, max_cfl=0.4)
    shape = (cfg.yc, cfg.xc)
    q = problems.prim_to_cons(cfg, np.full(shape, 1.4), np.full(shape, 0.6), np.full(shape, 0.8),
                              np.full(shape, 1.0))
    sim = orc.OracleSim(cfg)
    sim.set_q(q)
    a = 1.0 + 1.0  # c_s = sqrt(1.4*1/1.4) = 1, |v| = 1
    dt = sim.compute_dt(0.0, 0.0, max_time=10.0, delta=10.0)
    assert dt == pytest.approx(0.4 * cfg.dx / a, rel=1e-15)
    # the clamp lands exactly on the next output time
    for t in (0.1 - 1e-4, 0.3 - 7e-5, 0.7 - 3.3e-5):
        prev = round(t, 1) - 0.1
        dt = sim.compute_dt(t, prev, max_time=10.0, delta=0.1)
        target = min(10.0, prev + 0.1)
        assert t + dt >= target
        assert t + np.nextafter(dt, 0.0) < target or dt == target - t
    # max_time clamps too
    dt = sim.compute_dt(0.9999, 0.95, max_time=1.0, delta=10.0)
    assert 0.9999 + dt >= 1.0
