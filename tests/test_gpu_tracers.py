"""Tracer fields on the GPU through the C ABI: the FP64 path equals the oracle
bitwise (hydro and tracer planes, ghosts included) for several schemes and
boundary sets, on the graph-replayed device loop and the host loop; the
compensated form stays within its tolerance; unsupported combinations are
refused."""
import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import problems

pytestmark = pytest.mark.gpu

CASES = [
    dict(reconstruction="weno5z", riemann_solver="hllc", time_stepper="ssprk3", bc=("periodic",) * 4),
    dict(reconstruction="muscl", slope_limiter="monotonizedcentral", riemann_solver="hll", time_stepper="rk2",
         bc=("wall", "zerograd", "symmetric", "symmetricoutflowdiode")),
    dict(reconstruction="ppm", slope_limiter="vanleer", riemann_solver="rusanov", time_stepper="ssprk4",
         bc=("zerograd", "wall", "periodic", "periodic")),
]


def _setup(case, ntr=2, nx=70, ny=44):
    cfg = problems.kh_config(nx)
    cfg.ny = ny
    cfg.num_tracers = ntr
    cfg.max_cfl = 0.4
    for k, v in case.items():
        setattr(cfg, k, v)
    q0 = problems.kh_band(cfg)
    if ntr > 1:
        x = (np.arange(cfg.xc) + 0.5) / cfg.xc
        y = (np.arange(cfg.yc) + 0.5) / cfg.yc
        q0[5, 0] = q0[0, 0] * (0.5 + 0.25 * np.sin(2 * np.pi * x)[None, :] * np.cos(2 * np.pi * y)[:, None])
    return cfg, q0


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp64_tracers_bitwise(orc, case):
    cfg, q0 = _setup(CASES[case])
    print(f"case: tracers FP64 {CASES[case]}", flush=True)
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    t_o, p_o, _ = sim.run(max_steps=10)
    want = sim.get_q()
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        t_g, _ = ctx.run(10, use_graph=True)
        got = ctx.download()
    assert t_g == t_o
    np.testing.assert_array_equal(got, want)


def test_fp64_tracers_host_loop_equals_device_loop():
    cfg, q0 = _setup(CASES[1])
    with M.HydroContext(cfg) as a, M.HydroContext(cfg) as b:
        for c in (a, b):
            c.upload(q0)
            c.fill_bcs()
        a.run(6, use_graph=False)
        t, p = 0.0, 0.0
        for _ in range(6):
            dt = b.compute_dt(t, p)
            b.time_step(dt)
            t += dt
            if t >= p + cfg.output_delta:
                p = t
        np.testing.assert_array_equal(a.download(), b.download())


def test_fp32c_tracers_close_to_fp64():
    cfg, q0 = _setup(CASES[0], nx=128, ny=128)
    out = {}
    for form in (M.FP64, M.FP32C):
        with M.HydroContext(cfg, precision_form=form) as ctx:
            ctx.upload(q0)
            ctx.fill_bcs()
            ctx.run(20)
            out[form] = ctx.download()
    g = cfg.ng
    a, b = out[M.FP32C][:, 0, g:-g, g:-g], out[M.FP64][:, 0, g:-g, g:-g]
    rel = np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))
    assert rel[0] < 1e-5 and rel[4] < 1e-5 and rel[5] < 1e-5, rel


def test_unsupported_tracer_combinations():
    cfg, _ = _setup(CASES[0])
    cfg.bc = ("constant", "zerograd", "periodic", "periodic")
    with pytest.raises(M.MoscapError, match="constant"):
        M.HydroContext(cfg)
    c1 = M.Config(num_dim=1, nx=64, ng=3, reconstruction="weno5z", num_tracers=1, bc=("zerograd",) * 2)
    with pytest.raises(M.MoscapError, match="2-D"):
        M.HydroContext(c1)
    cfg, q0 = _setup(CASES[0])
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        with pytest.raises(M.MoscapError, match="tracer"):
            ctx.adjoint_forward(2)
