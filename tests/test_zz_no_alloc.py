"""mh_step_noalloc (one FP64 step with no device allocation, scratch in a fixed device array)
against the context entry points it restates, mh_compute_dt + mh_time_step: same dt, same state,
bit for bit, on calls in a row whose grids shrink (the scratch array keeps the previous call's
contents, so a stale CFL slot or stage state would show), plus the argument checks (CPU: they refuse
before any device access)."""
import ctypes as C

import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import Config, HydroContext, problems

def _case(name):
    if name == "kh":
        cfg = problems.kh_config(192)
        cfg.ny = 160
        return cfg, problems.kh_band(cfg)
    if name == "sod":
        cfg = problems.sod_config()
        cfg.nx, cfg.ny = 96, 64
        return cfg, problems.shock_tube_2d(cfg)
    cfg = Config(num_dim=1, nx=300, ng=3, bc=("wall", "zerograd", "periodic", "periodic"), reconstruction="ppm",
                 slope_limiter="vanleer", riemann_solver="hll", time_stepper="ssprk4", max_cfl=0.5)
    x = (np.arange(cfg.xc) + 0.5) / cfg.xc
    rho = 1.0 + 0.5 * np.sin(6.0 * x)
    q = np.zeros(cfg.shape)
    q[0, 0, 0] = rho
    q[1, 0, 0] = rho * 0.3
    q[2, 0, 0] = 1.0 / 0.4 + 0.5 * rho * 0.09
    return cfg, q


def _with_context(cfg, q0, dt_given):
    with HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        q_in = ctx.download()
        dt = dt_given if dt_given > 0 else ctx.compute_dt(0.0, 0.0, max_time=1e300, delta=1e300)
        ctx.time_step(dt)
        return q_in, dt, ctx.download()


def _device_buffer(ctx):
    """Device address of a context's current state buffer (device memory owned by the library, so the
    test needs no other allocator in the process)."""
    v = M.MhDeviceView()
    assert M.lib().mh_device_state(ctx.handle, C.byref(v)) == M.MH_OK
    return v.q


def _no_alloc(cfg, q_in, dt_given):
    with HydroContext(cfg) as a, HydroContext(cfg) as b:
        a.upload(q_in)
        b.upload(np.full(cfg.shape, np.nan))
        dt = M.step_noalloc(cfg, _device_buffer(a), _device_buffer(b), dt=dt_given)
        assert np.array_equal(a.download(), q_in)  # the input is read only
        return dt, b.download()


@pytest.mark.gpu
@pytest.mark.parametrize("order", [("kh", "sod", "line"), ("line", "kh", "sod")])
def test_step_noalloc_equals_time_step(order):
    sizes = []
    for k, name in enumerate(order):
        cfg, q0 = _case(name)
        dt_given = 0.0 if k != 1 else 1e-4  # CFL dt on the first and last call, a given dt between
        q_in, dt_ref, q_ref = _with_context(cfg, q0, dt_given)
        dt, q = _no_alloc(cfg, q_in, dt_given)
        assert dt == dt_ref, (name, dt, dt_ref)
        np.testing.assert_array_equal(q, q_ref, err_msg=name)
        sizes.append(q_in.size)
    if order[0] == "kh":
        assert sizes[0] > sizes[1] > sizes[2]  # each call smaller than the one before


def test_step_noalloc_refuses_oversized_grids():
    cfg = problems.kh_config(2048)
    with pytest.raises(M.MoscapError, match="too large for the scratch array"):
        M.step_noalloc(cfg, 1 << 20, 2 << 20)  # refused before any device access


def test_step_noalloc_refuses_aliasing_and_3d():
    cfg = problems.kh_config(64)
    with pytest.raises(M.MoscapError, match="distinct"):
        M.step_noalloc(cfg, 1 << 20, 1 << 20)
    cfg3 = Config(num_dim=3, nx=16, ny=16, nz=16, ng=3, bc=("periodic",) * 6, reconstruction="weno5z")
    with pytest.raises(M.MoscapError, match="1-D and 2-D"):
        M.step_noalloc(cfg3, 1 << 20, 2 << 20)
