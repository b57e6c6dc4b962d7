"""Oracle pins for the gravity source term (Gravity.cpp:14-40): S is rebuilt
from the stage's own Q every stage (SourceTerms.hpp:8-16) and enters each stage
lambda as S*dt (TimeStepping.cpp). A uniform gas in a periodic box only feels
the source, so the stage arithmetic is fixed bit for bit, and SSPRK3, exact for
polynomial right-hand sides up to degree 2, gives free fall in closed form:
m = rho g t and E = E0 + rho |g t|^2 / 2."""
import numpy as np
import pytest

from mosscap_amd import Config, problems


def _uniform(cfg, rho=1.5, p=2.0):
    shape = (cfg.yc, cfg.xc)
    z = np.zeros(shape)
    return problems.prim_to_cons(cfg, np.full(shape, rho), z, z, np.full(shape, p))


def _cfg(**kw):
    base = dict(num_dim=2, nx=12, ny=10, ng=3, bc=("periodic",) * 4, reconstruction="weno5z",
                riemann_solver="hllc", time_stepper="ssprk3", gravity=(-0.7, 0.3, 0.0))
    base.update(kw)
    return Config(**base)


def test_ssprk3_stage_arithmetic_on_uniform_state(orc):
    cfg = _cfg()
    gx, gy, _ = cfg.gravity
    q0 = _uniform(cfg)
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    dt = 0.0125
    sim.time_step(dt)
    got = sim.get_q()
    # the same lambdas by hand: every flux difference is exactly 0 on a uniform state
    rho, mx, my, E = (q0[v, 0, 5, 5] for v in range(4))

    def src(m_x, m_y):
        eu = m_x * gx
        eu += m_y * gy
        return (0.0 * dt, (0.0 + rho * gx) * dt, (0.0 + rho * gy) * dt, (0.0 + eu) * dt)

    q = [rho, mx, my, E]
    s = src(q[1], q[2])
    q1 = [q[v] + (0.0 + s[v]) for v in range(4)]
    s = src(q1[1], q1[2])
    q2 = [0.75 * q[v] + 0.25 * (q1[v] + 0.0 + s[v]) for v in range(4)]
    s = src(q2[1], q2[2])
    q3 = [(1.0 / 3.0) * q[v] + (2.0 / 3.0) * (q2[v] + 0.0 + s[v]) for v in range(4)]
    for v in range(4):
        assert np.all(got[v, 0] == q3[v]), v


@pytest.mark.parametrize("stepper", ["ssprk3", "rk2", "ssprk4"])
def test_free_fall_closed_form(orc, stepper):
    cfg = _cfg(time_stepper=stepper, reconstruction="muscl", riemann_solver="hll")
    gx, gy, _ = cfg.gravity
    rho = 1.5
    q0 = _uniform(cfg, rho=rho)
    E0 = q0[3, 0, 0, 0]
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    dt, n = 0.01, 25
    for _ in range(n):
        sim.time_step(dt)
    q = sim.get_q()
    t = n * dt
    np.testing.assert_allclose(q[1, 0], rho * gx * t, rtol=1e-13)
    np.testing.assert_allclose(q[2, 0], rho * gy * t, rtol=1e-13)
    # RK2 (2nd order) is exact for the linear momentum but not the quadratic energy
    e_tol = 1e-13 if stepper != "rk2" else 1e-4
    np.testing.assert_allclose(q[3, 0], E0 + 0.5 * rho * (gx * gx + gy * gy) * t * t, rtol=e_tol)
    assert np.all(q[0, 0] == rho)


def test_no_gravity_means_zero_source(orc):
    """gravity = None reproduces the source-free run bit for bit (S*dt = +0)."""
    cfg = _cfg(gravity=None)
    cfg0 = _cfg(gravity=(0.0, 0.0, 0.0))
    q0 = problems.kh_band(problems.kh_config(16))
    cfg.nx = cfg.ny = cfg0.nx = cfg0.ny = 16
    out = []
    for c in (cfg, cfg0):
        sim = orc.OracleSim(c)
        sim.set_q(q0)
        sim.fill_bcs()
        sim.run(max_steps=3)
        out.append(sim.get_q())
    np.testing.assert_array_equal(out[0], out[1])
