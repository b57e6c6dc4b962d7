"""3-D grids on the GPU (z sweep + per-plane x-y march + per-axis ghost
passes) against the oracle's NumDim = 3 path: FP64 bitwise for every
reconstruction, Riemann solver, time stepper and boundary type, through the
step-by-step entry points and the graph-replayed device loop; gravity with a z
component; the compensated form within tolerance; LTE EOS fields; refused
inputs."""
import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import Config, problems

pytestmark = pytest.mark.gpu

CASES = [
    dict(reconstruction="weno5z", riemann_solver="hllc", time_stepper="ssprk3",
         bc=("periodic",) * 6),
    dict(reconstruction="muscl", slope_limiter="vanleer", riemann_solver="hll", time_stepper="rk2",
         bc=("wall", "zerograd", "periodic", "periodic", "symmetric", "symmetricoutflowdiode")),
    dict(reconstruction="ppm", slope_limiter="minmod", riemann_solver="rusanov", time_stepper="ssprk4",
         bc=("zerograd", "zerograd", "wall", "wall", "wall", "zerograd")),
    dict(reconstruction="fog", riemann_solver="hllc", time_stepper="ssprk3",
         bc=("symmetricoutflowdiode", "wall", "zerograd", "symmetric", "periodic", "periodic")),
    dict(reconstruction="muscl", slope_limiter="monotonizedcentral", riemann_solver="hllc", time_stepper="ssprk3",
         bc=("periodic", "periodic", "constant", "zerograd", "zerograd", "constant")),
]
SHAPES = [(40, 36, 28), (70, 20, 9)]  # the second spans two strips in x


def _cfg(case, shape):
    nx, ny, nz = shape
    # CFL 0.3: the unsplit 3-D update adds three directions' Courant numbers
    cfg = Config(num_dim=3, nx=nx, ny=ny, nz=nz, ng=3, x_dim=nx / 32.0, max_cfl=0.3, **CASES[case])
    if "constant" in cfg.bc:
        cfg.bc_const = tuple((1.0, 0.05 * k, -0.03, 0.02, 2.6) for k in range(6))
    return cfg


def _oracle(orc, cfg, q0, steps):
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    t, p, _ = sim.run(max_steps=steps)
    return t, p, sim.get_q()


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fp64_3d_bitwise(orc, case, shape):
    print(f"case: 3-D FP64 {CASES[case]} {shape}", flush=True)
    cfg = _cfg(case, shape)
    # a 3x over-pressure: a 10x blast spanning the 9-plane periodic z extent of the second shape
    # drives WENO5-Z/HLLC to negative pressures (NaN) within a step, in the oracle as well
    q0 = problems.blast_3d(cfg, seed=case + 1, p_ratio=3.0)
    t_o, p_o, want = _oracle(orc, cfg, q0, 6)
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        t_g, p_g = ctx.run(6, use_graph=True)
        got = ctx.download()
    assert (t_g, p_g) == (t_o, p_o)
    np.testing.assert_array_equal(got, want)


def test_fp64_3d_step_entry_points(orc):
    """mh_compute_dt / mh_stage / mh_time_step, one call each, against the oracle's."""
    print("case: 3-D step entry points", flush=True)
    cfg = _cfg(1, (40, 36, 28))
    q0 = problems.blast_3d(cfg, seed=9)
    sim = orc.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        np.testing.assert_array_equal(ctx.download(), sim.get_q())
        dt = ctx.compute_dt(0.0, 0.0)
        assert dt == sim.compute_dt(0.0, 0.0)
        ctx.time_step(dt)
        sim.time_step(dt)
        np.testing.assert_array_equal(ctx.download(), sim.get_q())
        sim.begin_step()
        for s in range(2):  # rk2: two stages, each ending with the ghost fill
            ctx.stage(s, dt)
            sim.stage(s, dt)
            sim.fill_bcs()
        np.testing.assert_array_equal(ctx.download(), sim.get_q())


def test_fp64_3d_gravity(orc):
    print("case: 3-D gravity", flush=True)
    cfg = _cfg(2, (40, 36, 28))
    cfg.gravity = (-0.4, 0.25, -0.7)
    q0 = problems.blast_3d(cfg, seed=4)
    t_o, _, want = _oracle(orc, cfg, q0, 6)
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        t_g, _ = ctx.run(6)
        got = ctx.download()
    assert t_g == t_o
    np.testing.assert_array_equal(got, want)


def test_fp32c_3d_close_to_fp64():
    print("case: 3-D FP32C", flush=True)
    cfg = _cfg(0, (64, 48, 40))
    q0 = problems.blast_3d(cfg, seed=2)
    out = {}
    for form in (M.FP64, M.FP32C):
        with M.HydroContext(cfg, precision_form=form) as ctx:
            ctx.upload(q0)
            ctx.fill_bcs()
            ctx.run(10)
            out[form] = ctx.download()
    g = cfg.ng
    a = out[M.FP32C][:, g:-g, g:-g, g:-g].reshape(5, -1)
    b = out[M.FP64][:, g:-g, g:-g, g:-g].reshape(5, -1)
    rel = np.abs(a - b).sum(1) / np.abs(b).sum(1)
    assert rel[0] < 1e-5 and rel[4] < 1e-5, rel


def test_lte_fields_3d(orc):
    print("case: 3-D LTE EOS fields", flush=True)
    cfg = _cfg(0, (40, 36, 28))
    cfg.gamma = 5.0 / 3.0
    q = problems.blast_3d(cfg, seed=6)
    q[0:4] *= 1e-7  # rho ~ 1e-7 kg/m^3 at T ~ 1e4..1e5 K: partly ionised hydrogen
    q[4] *= 8.0
    with M.HydroContext(cfg) as ctx:
        ctx.upload(q)
        y, T = ctx.eos_fields(which=0)
    yo, To = orc.lte_eos(3, q, cfg.gamma)
    # the bisection's path is the same wherever T is bitwise equal; elsewhere a last-bit
    # difference of the device's pow/exp moved one decision by a step of the bisection
    T, To, y, yo = T.ravel(), To.ravel(), y.ravel(), yo.ravel()
    same = T == To
    assert same.mean() > 0.99
    # y is evaluated at the last trial temperature, which can differ where the runs stopped the
    # bisection one iteration apart with the same T
    exact = np.isclose(y[same], yo[same], rtol=1e-12, atol=1e-15)
    assert exact.mean() > 0.99
    np.testing.assert_allclose(y[same], yo[same], rtol=1e-6, atol=1e-12)
    assert np.all(np.abs(T - To) <= 0.3 * np.maximum(1.0, 1e-4 * To))


@pytest.mark.parametrize("what", ["userfn", "tracers", "adjoint"])
def test_3d_refusals(what):
    cfg = _cfg(0, (16, 16, 16))
    if what == "userfn":
        cfg.bc = ("periodic",) * 4 + ("userfn", "wall")
        with pytest.raises(M.MoscapError, match="userfn"):
            M.HydroContext(cfg)
    elif what == "tracers":
        cfg.num_tracers = 1
        with pytest.raises(M.MoscapError, match="2-D"):
            M.HydroContext(cfg)
    else:
        with M.HydroContext(cfg) as ctx:
            ctx.upload(problems.blast_3d(cfg))
            ctx.fill_bcs()
            with pytest.raises(M.MoscapError, match="2-D"):
                ctx.adjoint_forward(2)


def test_sedov_3d_through_host_driver(orc, tmp_path):
    """The host layer's 3-D sedov run by the C++ driver on the GPU equals the oracle's run of
    the same initial state, bit for bit (netCDF output read back)."""
    from scipy.io import netcdf_file
    from mosscap_amd import host
    from tests.test_host_layer import SEDOV_3D
    print("case: 3-D sedov via host driver", flush=True)
    q0 = host.initial_state(SEDOV_3D)
    cfg = Config(num_dim=3, nx=20, ny=18, nz=16, ng=3, x_dim=2.0, x_start=-1.0, y_start=-0.9, z_start=-0.8,
                 bc=("zerograd", "zerograd", "wall", "wall", "periodic", "periodic"), reconstruction="ppm",
                 riemann_solver="hll", time_stepper="ssprk3", max_cfl=0.3)
    t_o, _, want = _oracle(orc, cfg, q0, 4)
    host.run_config(SEDOV_3D, tmp_path)
    with netcdf_file(tmp_path / "sedov3d.nc", "r", mmap=False) as f:
        times, Q = f.variables["time"][:].copy(), f.variables["Q"][:].copy()
    assert times[-1] == t_o
    np.testing.assert_array_equal(Q[-1], want)


@pytest.mark.parametrize("form,tol", [(M.FP64, 1e-13), (M.FP32C, 1e-6)])
def test_conservation_256_cubed(form, tol):
    """At the 3-D bench size (256^3, periodic) every interior sum of rho, m and E is conserved: the
    z sweep's face fluxes and the march's x/y faces telescope, so only rounding remains."""
    print(f"case: 3-D conservation 256^3 form {form}", flush=True)
    cfg = Config(num_dim=3, nx=256, ny=256, nz=256, ng=3, x_dim=1.0, bc=("periodic",) * 6, reconstruction="weno5z",
                 riemann_solver="hllc", time_stepper="ssprk3", max_cfl=0.3)
    q0 = problems.blast_3d(cfg, seed=8, p_ratio=3.0)
    g = cfg.ng
    with M.HydroContext(cfg, precision_form=form) as ctx:
        ctx.upload(q0)
        ctx.fill_bcs()
        s0 = ctx.download()[:, g:-g, g:-g, g:-g].reshape(5, -1).sum(axis=1)
        ctx.run(5)
        q = ctx.download()
    assert np.isfinite(q).all()
    s1 = q[:, g:-g, g:-g, g:-g].reshape(5, -1).sum(axis=1)
    scale = np.abs(q0[:, g:-g, g:-g, g:-g]).reshape(5, -1).sum(axis=1)
    np.testing.assert_allclose(s1, s0, rtol=0, atol=tol * scale.max())
