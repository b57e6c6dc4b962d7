"""The tabulated LTE hydrogen EOS on the GPU (mh_set_tabulated_eos) against the
oracle's restatement of TabulatedLteH::update_eos run before every stage:
2-D and 3-D FP64 runs, the fields it keeps, the compensated form, the
in-place setup update, the negative-energy error and the host driver with a
table file. Device log10 and pow may differ from the host's in the last bit,
so states are compared to 1e-12 relative rather than bitwise (a 1-ulp change
of log10(eint/rho) moves y by ~1e-16 relative)."""
import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import Config, problems
from mosscap_amd.lte_table import generate, write_nc

pytestmark = pytest.mark.gpu

TABLE = generate(128, 64)


def _cfg(nd=2, **kw):
    base = dict(num_dim=nd, nx=48, ny=40, nz=24 if nd == 3 else 1, ng=3, bc=("periodic",) * (2 * nd),
                reconstruction="weno5z", riemann_solver="hllc", time_stepper="ssprk3", gamma=5.0 / 3.0, max_cfl=0.4,
                x_dim=1.0e6)
    base.update(kw)
    return Config(**base)


def _state(cfg):
    """Partly ionised hydrogen: rho ~ 1e-7 kg/m^3, eint/rho 1e8..1e10 J/kg, a 10 km/s shear flow."""
    if cfg.num_dim == 3:
        x, y, z = problems.cell_centres_3d(cfg)
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    else:
        x, y = problems.cell_centres(cfg)
        X, Y = np.meshgrid(x, y)
        Z = 0.0 * X
    # periodic over each axis's interior length, so the ghosts the fill writes equal the initial ones
    X, Y, Z = X / (cfg.nx * cfg.dx), Y / (cfg.ny * cfg.dx), Z / (max(cfg.nz, 1) * cfg.dx)
    rho = 10 ** (-7.0 + 0.4 * np.sin(2 * np.pi * X) * np.cos(2 * np.pi * (Y + Z)))
    e_rho = 10 ** (9.0 + 0.8 * np.cos(2 * np.pi * (X - Y)))
    vx = 1e4 * np.sin(2 * np.pi * Y)
    p = (cfg.gamma - 1.0) * rho * e_rho
    if cfg.num_dim == 3:
        return problems.prim_to_cons_3d(cfg, rho, vx, 0 * vx, 0 * vx, p)
    return problems.prim_to_cons(cfg, rho, vx, 0 * vx, p)


def _close(a, b, rtol=1e-12):
    scale = np.abs(b).reshape(b.shape[0], -1).max(axis=1).reshape((-1,) + (1,) * (b.ndim - 1))
    assert np.all(np.abs(a - b) <= rtol * scale), np.max(np.abs(a - b) / scale)


@pytest.mark.parametrize("nd,stepper", [(2, "ssprk3"), (2, "rk2"), (3, "ssprk4")])
def test_fp64_matches_oracle(orc, nd, stepper):
    print(f"case: tabulated EOS FP64 nd={nd} {stepper}", flush=True)
    cfg = _cfg(nd, time_stepper=stepper)
    q0 = _state(cfg)
    sim = orc.OracleSim(cfg)
    sim.set_tabulated_eos(TABLE)
    sim.set_q(q0)
    sim.update_eos()
    sim.fill_bcs()
    t_o, _, _ = sim.run(max_steps=8)
    with M.HydroContext(cfg) as ctx:
        ctx.set_tabulated_eos(TABLE)
        ctx.upload(q0)
        ctx.tabulated_eos_update()
        ctx.fill_bcs()
        t_g, _ = ctx.run(8)
        got = ctx.download()
        yg, Tg = ctx.tabulated_eos_fields()
    want = sim.get_q()
    yo, To, neg = sim.tabulated_fields()
    assert not neg
    assert abs(t_g - t_o) <= 1e-12 * t_o
    _close(got, want)
    np.testing.assert_allclose(yg, yo, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(Tg, To, rtol=1e-10)
    assert 0.01 < yg.mean() < 0.99  # the run goes through partial ionisation


def test_step_entry_points_and_graph_loop_agree():
    cfg = _cfg()
    q0 = _state(cfg)
    out = []
    for mode in ("graph", "loop", "steps"):
        with M.HydroContext(cfg) as ctx:
            ctx.set_tabulated_eos(TABLE)
            ctx.upload(q0)
            ctx.tabulated_eos_update()
            ctx.fill_bcs()
            if mode == "steps":
                t = 0.0
                for _ in range(5):
                    dt = ctx.compute_dt(t, 0.0, delta=1e30)
                    ctx.time_step(dt)
                    t += dt
            else:
                ctx.run(5, use_graph=(mode == "graph"), delta=1e30)
            out.append(ctx.download())
    np.testing.assert_array_equal(out[0], out[1])
    np.testing.assert_array_equal(out[0], out[2])


def test_with_tracers_and_gravity(orc):
    """The tabulated EOS's stage-0 copy carries the tracer planes; the gravity/tracer stage
    instance reads the updated input."""
    cfg = _cfg(num_tracers=2, gravity=(-30.0, 12.0, 0.0))
    q0 = _state(cfg)
    q0[4] = 0.3 * q0[0]
    q0[5] = q0[0] * (q0[1] > 0)
    sim = orc.OracleSim(cfg)
    sim.set_tabulated_eos(TABLE)
    sim.set_q(q0)
    sim.update_eos()
    sim.fill_bcs()
    t_o, _, _ = sim.run(max_steps=6)
    with M.HydroContext(cfg) as ctx:
        ctx.set_tabulated_eos(TABLE)
        ctx.upload(q0)
        ctx.tabulated_eos_update()
        ctx.fill_bcs()
        t_g, _ = ctx.run(6)
        got = ctx.download()
    assert abs(t_g - t_o) <= 1e-12 * t_o
    _close(got, sim.get_q())


def test_autotune_leaves_the_eos_fields():
    """mh_autotune's trial steps run update_eos; it restores the fields with the state, so a run
    after it equals a run without it."""
    cfg = _cfg()
    q0 = _state(cfg)
    out = []
    for tune in (False, True):
        with M.HydroContext(cfg) as ctx:
            ctx.set_tabulated_eos(TABLE)
            ctx.upload(q0)
            ctx.tabulated_eos_update()
            ctx.fill_bcs()
            if tune:
                ctx.autotune(steps=2, candidates=[8, 16])
            ctx.run(4, use_graph=False)
            out.append((ctx.download(), ctx.tabulated_eos_fields()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1][0], out[1][1][0])


def test_fp32c_close_to_fp64():
    cfg = _cfg()
    q0 = _state(cfg)
    res = {}
    for form in (M.FP64, M.FP32C):
        with M.HydroContext(cfg, precision_form=form) as ctx:
            ctx.set_tabulated_eos(TABLE)
            ctx.upload(q0)
            ctx.tabulated_eos_update()
            ctx.fill_bcs()
            ctx.run(10)
            res[form] = ctx.download()
    g = cfg.ng
    a, b = res[M.FP32C][:, 0, g:-g, g:-g], res[M.FP64][:, 0, g:-g, g:-g]
    rel = np.abs(a - b).sum(axis=(1, 2)) / np.abs(b).sum(axis=(1, 2))
    assert rel[0] < 1e-5 and rel[3] < 1e-5, rel


def test_negative_internal_energy_is_an_error():
    cfg = _cfg()
    q0 = _state(cfg)
    q0[3, 0, 10, 10] = -1.0
    with M.HydroContext(cfg) as ctx:
        ctx.set_tabulated_eos(TABLE)
        ctx.upload(q0)
        with pytest.raises(M.MoscapError, match="eint negative"):
            ctx.tabulated_eos_update()


def test_host_driver_with_table_file(orc, tmp_path):
    """eos.type: tabulatedlteh with a table file: the host driver's run (the C++ reader, setup's
    update_eos, the device stages, the ion_frac / T output) against the oracle's run of the same
    initial state. A Sedov blast at rho = 1 with p ~ 1e9..1e10 Pa sits inside the table."""
    from scipy.io import netcdf_file
    from mosscap_amd import host
    table = generate(96, 48, log_rho=(-3.0, 2.0))
    path = tmp_path / "lte.nc"
    write_nc(table, path)
    text = ("simulation:\n  num_dim: 2\ngrid:\n  x: 40\n  y: 40\n  x_dim: 2.0\n  x_start: -1.0\n  y_start: -1.0\n"
            "boundary:\n  xs: zerograd\n  xe: zerograd\n  ys: zerograd\n  ye: zerograd\n"
            "scheme:\n  reconstruction: ppm\n  riemann_solver: hll\ntimestep:\n  max_cfl: 0.3\n  max_steps: 6\n"
            f"eos:\n  type: tabulatedlteh\n  table_path: {path}\noutput:\n  name: tab\n"
            "problem:\n  name: sedov\n  energy: 1.0e9\n  p_ambient: 1.0e9\n")
    q0 = host.initial_state(text)
    cfg = Config(num_dim=2, nx=40, ny=40, ng=3, x_dim=2.0, x_start=-1.0, y_start=-1.0, bc=("zerograd",) * 4,
                 reconstruction="ppm", riemann_solver="hll", time_stepper="ssprk3", gamma=5.0 / 3.0, max_cfl=0.3)
    sim = orc.OracleSim(cfg)
    sim.set_tabulated_eos(table)
    sim.set_q(q0)
    sim.update_eos()
    sim.fill_bcs()
    t_o, _, _ = sim.run(max_steps=6)
    host.run_config(text, tmp_path)
    with netcdf_file(tmp_path / "tab.nc", "r", mmap=False) as f:
        times, Q = f.variables["time"][:].copy(), f.variables["Q"][:].copy()
        ion, T = f.variables["ion_frac"][:].copy(), f.variables["T"][:].copy()
    assert abs(times[-1] - t_o) <= 1e-12 * t_o
    _close(Q[-1], sim.get_q())
    yo, To, _ = sim.tabulated_fields()
    np.testing.assert_allclose(ion[-1], yo, rtol=1e-10, atol=1e-14)
    np.testing.assert_allclose(T[-1], To, rtol=1e-10)
    assert 0.0 < yo.min() and yo.max() < 1.0
