"""The gravity source term on the GPU: FP64 bitwise equal to the oracle (2-D
and 1-D, several steppers), the reference's kelvin_helmholtz problem through
the host layer, the compensated form within tolerance, and the adjoint's
gravity term against the torch reference gradient."""
import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import host, problems

pytestmark = pytest.mark.gpu

CASES = [
    dict(reconstruction="weno5z", riemann_solver="hllc", time_stepper="ssprk3", bc=("periodic",) * 4),
    dict(reconstruction="muscl", slope_limiter="minmod", riemann_solver="hll", time_stepper="rk2",
         bc=("periodic", "periodic", "wall", "wall")),
    dict(reconstruction="ppm", riemann_solver="rusanov", time_stepper="ssprk4",
         bc=("zerograd", "zerograd", "symmetric", "symmetricoutflowdiode")),
]


def _oracle_run(orc, cfg, q0, steps):
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    t, _, _ = sim.run(max_steps=steps)
    return t, sim.get_q()


@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp64_gravity_bitwise(orc, case):
    print(f"case: gravity FP64 {CASES[case]}", flush=True)
    cfg = problems.kh_config(64)
    cfg.ny, cfg.max_cfl, cfg.gravity = 48, 0.4, (-0.7, 0.3, 0.0)
    for k, v in CASES[case].items():
        setattr(cfg, k, v)
    q0 = problems.kh_band(cfg)
    t_o, want = _oracle_run(orc, cfg, q0, 8)
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        t_g, _ = ctx.run(8, use_graph=True)
        got = ctx.download()
    assert t_g == t_o
    np.testing.assert_array_equal(got, want)


def test_fp64_gravity_1d(orc):
    cfg = M.Config(num_dim=1, nx=300, ng=3, bc=("zerograd", "wall"), reconstruction="weno5z", riemann_solver="hllc",
                   time_stepper="ssprk3", max_cfl=0.4, gravity=(-2.0, 0.0, 0.0))
    x = (np.arange(cfg.xc) - cfg.ng + 0.5) * cfg.dx
    rho = 1.0 + 0.5 * (x > 0.5)
    q0 = np.zeros(cfg.shape)
    q0[0, 0, 0] = rho
    q0[2, 0, 0] = 1.0 / 0.4
    t_o, want = _oracle_run(orc, cfg, q0, 20)
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        t_g, _ = ctx.run(20)
        got = ctx.download()
    assert t_g == t_o
    np.testing.assert_array_equal(got, want)


KH_RT = ("simulation:\n  num_dim: 2\ngrid:\n  x: 32\n  y: 96\n  x_dim: 0.5\n  x_start: -0.25\n  y_start: -0.75\n"
         "boundary:\n  xs: periodic\n  xe: periodic\n  ys: wall\n  ye: wall\nscheme:\n  reconstruction: weno5z\n"
         "  riemann_solver: hllc\ntimestep:\n  max_cfl: 0.4\n  max_steps: 15\noutput:\n  name: rt\n"
         "problem:\n  name: kelvin_helmholtz\n")


def test_reference_kelvin_helmholtz_through_host(orc, tmp_path):
    """The reference's gravity-driven problem: the host's initial state and gravity, run by the
    host driver on the GPU, equal the oracle's run of the same state bitwise."""
    from scipy.io import netcdf_file
    s = host.setup_summary(KH_RT)
    q0 = host.initial_state(KH_RT)
    cfg = M.Config(num_dim=2, nx=32, ny=96, ng=3, x_dim=0.5, x_start=-0.25, y_start=-0.75,
                   bc=("periodic", "periodic", "wall", "wall"), reconstruction="weno5z", riemann_solver="hllc",
                   time_stepper="ssprk3", max_cfl=0.4, gravity=tuple(s["gravity"]))
    t_o, want = _oracle_run(orc, cfg, q0, 15)
    host.run_config(KH_RT, tmp_path)
    with netcdf_file(tmp_path / "rt.nc", "r", mmap=False) as f:
        times, Q = f.variables["time"][:].copy(), f.variables["Q"][:].copy()
    assert times[-1] == t_o
    np.testing.assert_array_equal(Q[-1], want)


def test_fp32c_gravity_close_to_fp64():
    cfg = problems.kh_config(128)
    cfg.gravity = (-0.7, 0.3, 0.0)
    q0 = problems.kh_band(cfg)
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
    assert rel[0] < 1e-5 and rel[3] < 1e-5, rel


def test_adjoint_with_gravity_matches_torch():
    torch = pytest.importorskip("torch")
    from mosscap_amd.adjoint import KHAdjointProblem
    from tests import torch_hydro as th

    grav = (-0.7, 0.3, 0.0)
    cfg = problems.kh_config(24)
    cfg.ny, cfg.max_cfl, cfg.gravity = 20, 0.4, grav
    prob = KHAdjointProblem(cfg=cfg, nsteps=6, checkpoint_every=4, riemann_solver="hllc", seed=3)
    rng = np.random.default_rng(23)
    control = 0.01 * rng.uniform(-1, 1, size=(cfg.ny, cfg.nx))
    prob.forward(control)
    target = prob.forward(0.02 * rng.uniform(-1, 1, size=control.shape))
    rho = prob.forward(control)
    diff = rho - target
    g_hip = prob.backward(diff / np.linalg.norm(diff))

    ng = cfg.ng
    c = torch.tensor(control, dtype=torch.float64, requires_grad=True)
    base = torch.from_numpy(problems.kh_band(cfg, seed=3)[:, 0].copy())
    vx, vy = torch.from_numpy(prob.vx), torch.from_numpy(prob.vy)
    rho0 = base[0, ng:-ng, ng:-ng] + c
    q = base.clone()
    q[0, ng:-ng, ng:-ng] = rho0
    q[1, ng:-ng, ng:-ng] = rho0 * vx
    q[2, ng:-ng, ng:-ng] = rho0 * vy
    q[3, ng:-ng, ng:-ng] = 2.5 / 0.4 + 0.5 * rho0 * (vx * vx + vy * vy)
    q = th.fill_periodic(q, ng)
    for dt in prob.dts:
        q = th.ssprk3_step(q, 1.4, float(dt), cfg.dx, ng, "hllc", grav=grav)
    loss = torch.linalg.norm(q[0, ng:-ng, ng:-ng] - torch.from_numpy(target))
    loss.backward()
    g_ref = c.grad.numpy()
    np.testing.assert_allclose(rho, q[0, ng:-ng, ng:-ng].detach().numpy(), rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(g_hip, g_ref, rtol=1e-8, atol=1e-8 * np.abs(g_ref).max())
