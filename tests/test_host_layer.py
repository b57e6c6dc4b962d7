"""Host layer on the CPU: YAML reading, get_or / enum semantics and error
messages (MosscapConfig.hpp:62-123), setup defaults and checks
(SetupModules.hpp:9-262, Appendix of key defaults), the problem generators
against the Python generators bitwise, and the netCDF output read back with
scipy's independent netCDF-3 reader."""
import re
from pathlib import Path

import numpy as np
import pytest

from mosscap_amd import host, problems

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def H():
    host.lib()
    return host


def test_library_exports_declared_symbols(H):
    text = re.sub(r"/\*.*?\*/", "", (ROOT / "include" / "mosscap_host.h").read_text(), flags=re.S)
    names = set(re.findall(r"\b(mhh_\w+)\s*\(", text))
    assert names and names <= set(H.SYMBOLS)
    assert all(hasattr(H.lib(), n) for n in names)


# ---- YAML + get_or -------------------------------------------------------------
DOC = """
# comment line
grid:
  x: 64          # trailing comment
  y: 48
  x_dim: 2.5e-1
  name: "quoted # not a comment"
flags:
  a: yes
  b: Off
  c: true
nested:
  deeper:
    value: -7
    empty:
list_flow: [1.0, 2, 3e-3]
list_block:
  - one
  - two
"""


@pytest.mark.parametrize("key,kind,default,want", [
    ("grid.x", "int", 0, "64"),
    ("grid.y", "int", 0, "48"),
    ("grid.z", "int", 256, "256"),          # absent leaf -> default
    ("nope.x", "int", 5, "5"),              # absent head -> default
    ("grid.x_dim", "double", 1.0, "0.25"),
    ("grid.x", "double", 0.0, "64"),        # ints read as reals
    ("grid.name", "string", "", "quoted # not a comment"),
    ("flags.a", "bool", False, "true"),
    ("flags.b", "bool", True, "false"),
    ("flags.c", "bool", False, "true"),
    ("nested.deeper.value", "int", 0, "-7"),
    ("nested.deeper.empty", "int", 11, "11"),  # key with no value -> default
])
def test_get_or(H, key, kind, default, want):
    assert H.get_or(DOC, key, kind, default) == want


def test_get_or_conversion_error_names_key_type_and_value(H):
    with pytest.raises(H.HostError, match=r'Unable to convert key "x_dim" to type "int". Has value: "2.5e-1".'):
        H.get_or(DOC, "grid.x_dim", "int", 0)
    with pytest.raises(H.HostError, match=r'to type "bool"'):
        H.get_or(DOC, "grid.x", "bool", False)


@pytest.mark.parametrize("bad", ["a:\n\tb: 1\n", "a: [1, 2\n", "a: 1\n  b: 2\n"])
def test_yaml_errors(H, bad):
    with pytest.raises(H.HostError, match="YAML"):
        H.get_or(bad, "a", "int", 0)


# ---- setup ------------------------------------------------------------------------
MIN = "simulation:\n  num_dim: 2\nproblem:\n  name: shock_tube_2d\n"


def test_setup_defaults(H):
    s = H.setup_summary(MIN)
    # SetupModules.hpp defaults: ng 3, 256 cells, x_dim 1, wall BCs, muscl/MC/hll, ssprk3, CFL 0.8
    assert (s["ng"], s["xc"], s["yc"], s["zc"]) == (3, 262, 262, 1)
    assert s["dx"] == 1.0 / 256
    assert s["bc"][:4] == [0, 0, 0, 0]
    assert (s["reconstruction"], s["slope_limit"], s["riemann_solver"], s["time_stepper"]) == (1, 1, 1, 1)
    assert s["max_cfl"] == 0.8 and s["gamma"] == 1.4
    assert s["max_time"] == 0.2  # shock_tube_2d's own default (ShockTube2d.cpp:48)
    assert s["delta"] == 0.1 and s["single_file"] and s["output_name"] == "output_shock_tube_2d"
    assert s["max_steps"] == -1 and s["precision_form"] == 0


def test_setup_reads_every_key(H):
    text = """
simulation:
  num_dim: 2
grid:
  num_ghost: 3
  x: 40
  y: 24
  x_dim: 2.0
  x_start: -1.0
  y_start: -0.5
boundary:
  xs: Constant
  xs_const: [1.0, 0.5, 0.25, 9.0, 3.0]
  xe: symmetricOutflowDiode
  ys: periodic
  ye: PERIODIC
scheme:
  reconstruction: WENO5Z
  slope_limiter: minmod
  riemann_solver: hllc
timestep:
  scheme: ssprk4
  max_cfl: 0.3
  max_time: 0.7
  max_steps: 12
eos:
  gamma: 1.6666666666666667
output:
  name: run
  single_file: false
  delta_t: 0.05
  variables:
    primitive: true
problem:
  name: kh_band
hip:
  precision_form: FP32C
  span_rows: 16
  rho_floor: 1e-10
"""
    s = H.setup_summary(text)
    assert (s["xc"], s["yc"]) == (46, 30) and s["dx"] == 2.0 / 40
    assert s["loc"][:2] == [-1.0, -0.5]
    assert s["bc"][:4] == [5, 3, 1, 1]
    # 3-D conserved order (rho, mx, my, mz, E) mapped to 2-D (rho, mx, my, E): m_z dropped
    assert s["bc_const"][0][:4] == [1.0, 0.5, 0.25, 3.0]
    assert (s["reconstruction"], s["slope_limit"], s["riemann_solver"], s["time_stepper"]) == (3, 2, 2, 2)
    assert (s["max_cfl"], s["max_time"], s["max_steps"]) == (0.3, 0.7, 12)
    assert s["gamma"] == 1.6666666666666667
    assert (s["output_name"], s["single_file"], s["delta"], s["primitive"]) == ("run", False, 0.05, True)
    assert (s["precision_form"], s["span_rows"], s["rho_floor"]) == (1, 16, 1e-10)


def test_constant_bc_mapping_1d(H):
    text = ("simulation:\n  num_dim: 1\ngrid:\n  x: 32\nboundary:\n  xs: constant\n  xs_const: [1, 2, 3, 4, 5]\n"
            "problem:\n  name: shock_tube_1d\n")
    s = H.setup_summary(text)
    # 1-D takes (rho, m_x, E) from entries 0, 1, 4; y/z faces are not read in 1-D
    assert s["bc_const"][0][:3] == [1.0, 2.0, 5.0] and (s["yc"], s["zc"]) == (1, 1)
    with pytest.raises(H.HostError, match="shock_tube_2d only handles 2d problems"):
        H.setup_summary(text.replace("shock_tube_1d", "shock_tube_2d"))


@pytest.mark.parametrize("text,msg", [
    ("grid:\n  x: 8\n", r"simulation.num_dim = 0, must be in range \[1, 3\]"),
    ("simulation:\n  num_dim: 4\n", r"simulation.num_dim = 4, must be in range \[1, 3\]"),
    (MIN + "grid:\n  num_ghost: 2\nscheme:\n  reconstruction: weno5z\n",
     r"Insufficient ghost cells for reconstruction scheme weno5z, need minimum of 3."),
    (MIN + "scheme:\n  riemann_solver: roe\n", r'Unable to find associated key in ".*RiemannSolver" for "roe".'),
    (MIN + "boundary:\n  xs: outflow\n", r'for "outflow"'),
    (MIN + "boundary:\n  xs: constant\n  xs_const: xxx\n", r"Provide constant boundary values for xs_const"),
    (MIN + "eos:\n  type: tabulatedlteh\n", r"eos.table_path mosscap_lte_h_tables.nc: cannot open the file"),
    (MIN + "eos:\n  type: dexpressure\n", r"DexRT equation of state is not supported"),
    (MIN + "output:\n  variables:\n    fluxes: true\n", r"fluxes / source"),
    ("simulation:\n  num_dim: 2\n", r"No method for circular_explosion found."),
    (MIN + "grid:\n  dx: -1.0\n", r"grid.dx cannot be negative."),
    (MIN + "dex:\n  enable: true\n", r"dex.enable"),
])
def test_setup_errors(H, text, msg):
    with pytest.raises(H.HostError, match=msg):
        H.setup_summary(text)


@pytest.mark.parametrize("value", ["problem_supplied", "anything"])
def test_constant_without_list_keeps_zero_values(H, value):
    # SetupModules.hpp:115-124: only the literal default "xxx" raises; other strings leave the values unset
    s = H.setup_summary(MIN + f"boundary:\n  xs: constant\n  xs_const: {value}\n")
    assert s["bc"][0] == 5 and s["bc_const"][0] == [0.0] * 5


# ---- problem generators: C++ == Python bitwise ------------------------------------
def _cfg_yaml(name, nx, ny, ng, x_dim=1.0, x_start=0.0, y_start=0.0, extra=""):
    return (f"simulation:\n  num_dim: 2\ngrid:\n  x: {nx}\n  y: {ny}\n  num_ghost: {ng}\n  x_dim: {x_dim}\n"
            f"  x_start: {x_start}\n  y_start: {y_start}\nproblem:\n  name: {name}\n{extra}")


def test_shock_tube_ic_matches_python(H):
    cfg = problems.sod_config()
    cfg.nx, cfg.ny = 64, 40
    q = H.initial_state(_cfg_yaml("shock_tube_2d", 64, 40, 2))
    np.testing.assert_array_equal(q, problems.shock_tube_2d(cfg))


def test_shock_tube_1d_ic(H):
    q = H.initial_state("simulation:\n  num_dim: 1\ngrid:\n  x: 32\n  num_ghost: 2\nproblem:\n  name: shock_tube_1d\n"
                        "  left: [1.0, 0.75, 1.0]\n  right: [0.125, 0.0, 0.1]\n")
    dx = 1.0 / 32
    x = np.arange(36) * dx + (-(2 - 0.5) * dx)
    left = x < 0.5
    rho = np.where(left, 1.0, 0.125)
    v = np.where(left, 0.75, 0.0)
    p = np.where(left, 1.0, 0.1)
    want = np.stack([rho, rho * v, p / (1.4 - 1.0) + 0.5 * rho * (v * v)])
    np.testing.assert_array_equal(q[:, 0, 0, :], want)


def test_kh_band_ic_matches_python(H):
    cfg = problems.kh_config(48)
    q = H.initial_state(_cfg_yaml("kh_band", 48, 48, 3))
    np.testing.assert_array_equal(q, problems.kh_band(cfg))
    q2 = H.initial_state(_cfg_yaml("kh_band", 48, 48, 3, extra="  seed: 7\n  amplitude: 0.02\n"))
    np.testing.assert_array_equal(q2, problems.kh_band(cfg, seed=7, amp=0.02))
    # tracer fields (simulation.n_extra_fields): a dye marking the band
    text = _cfg_yaml("kh_band", 48, 48, 3).replace("num_dim: 2", "num_dim: 2\n  n_extra_fields: 2")
    cfg.num_tracers = 2
    q3 = H.initial_state(text)
    assert q3.shape == (6, 1, 54, 54)
    np.testing.assert_array_equal(q3, problems.kh_band(cfg))


def test_sedov_ic_matches_python(H):
    cfg = problems.sedov_config(64)
    q = H.initial_state(_cfg_yaml("sedov", 64, 64, 3, x_dim=2.0, x_start=-1.0, y_start=-1.0))
    np.testing.assert_array_equal(q, problems.sedov(cfg))


# ---- netCDF output, read back with scipy's reader ---------------------------------
def test_netcdf_single_file_layout(H, tmp_path):
    netcdf = pytest.importorskip("scipy.io").netcdf_file
    text = _cfg_yaml("kh_band", 24, 16, 3) + "output:\n  name: kh\n  variables:\n    primitive: true\n"
    H.write_initial_output(text, tmp_path)
    q0 = H.initial_state(text)
    with netcdf(tmp_path / "kh.nc", "r", mmap=False) as f:
        assert f.version_byte == 2
        assert f.dimensions["x"] == 30 and f.dimensions["y"] == 22 and f.dimensions["z"] == 1
        assert f.dimensions["var"] == 4 and f.dimensions["time"] is None  # unlimited
        a = f._attributes
        assert a["problem"] == b"kh_band" and a["num_ghost"] == 3 and a["single_file"] == 1
        assert (a["irho"], a["imx"], a["imy"], a["iene"]) == (0, 1, 2, 3)
        assert (a["ivx"], a["ivy"], a["ipre"]) == (1, 2, 3) and a["write_primitive"] == 1
        np.testing.assert_array_equal(f.variables["time"][:], [0.0, 0.5])
        dx = 1.0 / 24
        np.testing.assert_array_equal(f.variables["x"][:], np.arange(30) * dx + (-(3 - 0.5) * dx) + 0.0)
        Q = f.variables["Q"][:]
        assert Q.shape == (2, 4, 1, 22, 30)
        np.testing.assert_array_equal(Q[0], q0)
        np.testing.assert_array_equal(Q[1], q0)
        W = f.variables["W"][:]
        rho = q0[0]
        vx, vy = q0[1] / rho, q0[2] / rho
        p = (1.4 - 1.0) * (q0[3] - 0.5 * rho * (vx * vx + vy * vy))
        np.testing.assert_array_equal(W[1], np.stack([rho, vx, vy, p]))


def test_netcdf_multi_file(H, tmp_path):
    netcdf = pytest.importorskip("scipy.io").netcdf_file
    text = _cfg_yaml("shock_tube_2d", 16, 8, 2) + "output:\n  single_file: false\n"
    H.write_initial_output(text, tmp_path)
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == ["output_shock_tube_2d_00000.nc", "output_shock_tube_2d_00001.nc"]
    with netcdf(tmp_path / names[1], "r", mmap=False) as f:
        assert f.dimensions["time"] == 1
        assert float(f.variables["time"][0]) == 0.5
        np.testing.assert_array_equal(f.variables["Q"][:], H.initial_state(text))


# ---- gravity source: the reference's kelvin_helmholtz problem ---------------------
KH_RT = ("simulation:\n  num_dim: 2\ngrid:\n  x: 16\n  y: 48\n  x_dim: 0.5\n  x_start: -0.25\n  y_start: -0.75\n"
         "boundary:\n  xs: periodic\n  xe: periodic\n  ys: wall\n  ye: wall\nproblem:\n  name: kelvin_helmholtz\n")


def test_kelvin_helmholtz_registers_gravity(H):
    s = H.setup_summary(KH_RT)
    # sources.gravity.y is forced to -0.1 (set_if_not_present assigns); x keeps setup_gravity's default -1
    assert s["gravity"] == [-1.0, -0.1, 0.0]
    s = H.setup_summary(KH_RT + "sources:\n  gravity:\n    x: 0.0\n    y: -5.0\n")
    assert s["gravity"] == [0.0, -0.1, 0.0]
    assert H.setup_summary(MIN)["gravity"] is None


def test_kelvin_helmholtz_checks_aspect_ratio(H):
    with pytest.raises(H.HostError, match=r"Aspect ratio should be 3 \(y / x\) == 3 for axis = 1"):
        H.setup_summary(KH_RT.replace("y: 48", "y: 32"))


def test_kelvin_helmholtz_initial_state(H):
    q = H.initial_state(KH_RT)
    ng, dx = 3, 0.5 / 16
    y = np.arange(54) * dx + (-(ng - 0.5) * dx) - 0.75
    rho = np.where(y > 0.0, 2.0, 1.0)
    np.testing.assert_array_equal(q[0, 0, :, 5], rho)
    p_expect = 2.5 + rho * -0.1 * y
    vy = q[2, 0, :, 5] / q[0, 0, :, 5]
    e_int = q[3, 0, :, 5] - 0.5 * q[0, 0, :, 5] * vy * vy
    np.testing.assert_allclose(e_int * 0.4, p_expect, rtol=1e-12)


def test_analytic_lte_eos_setup_and_output(H, tmp_path):
    """eos.type analyticlteh: gamma defaults to 5/3 (Eos.cpp:24) and the output carries ion_frac and T
    (zeros here: no device has run update_eos)."""
    netcdf = pytest.importorskip("scipy.io").netcdf_file
    text = MIN + "eos:\n  type: AnalyticLteH\n  avg_mass: 1.2\noutput:\n  name: lte\n"
    assert H.setup_summary(text)["gamma"] == 5.0 / 3.0
    H.write_initial_output(text, tmp_path)
    with netcdf(tmp_path / "lte.nc", "r", mmap=False) as f:
        assert f.variables["ion_frac"].shape == (2, 1, 262, 262)
        assert f.variables["T"].shape == (2, 1, 262, 262)


SEDOV_3D = ("simulation:\n  num_dim: 3\ngrid:\n  x: 20\n  y: 18\n  z: 16\n  x_dim: 2.0\n  x_start: -1.0\n"
            "  y_start: -0.9\n  z_start: -0.8\nboundary:\n  xs: zerograd\n  xe: zerograd\n  ys: wall\n  ye: wall\n"
            "  zs: periodic\n  ze: periodic\nscheme:\n  reconstruction: ppm\n  riemann_solver: hll\n"
            "timestep:\n  max_cfl: 0.3\n  max_steps: 4\noutput:\n  name: sedov3d\nproblem:\n  name: sedov\n"
            "  radius_cells: 2.5\n")


def test_sedov_3d_setup_and_energy(H):
    s = H.setup_summary(SEDOV_3D)
    assert (s["num_dim"], s["xc"], s["yc"], s["zc"]) == (3, 26, 24, 22)
    assert s["bc"][4:] == [1, 1] and s["loc"] == [-1.0, -0.9, -0.8]
    q = H.initial_state(SEDOV_3D)
    assert q.shape == (5, 22, 24, 26)
    g, dx = 3, 0.1
    assert not np.any(q[1:4])
    e = q[4, g:-g, g:-g, g:-g]
    ambient = 1e-5 / (1.4 - 1.0)
    hot = e > ambient
    assert 20 < hot.sum() < 120  # a ball of radius 2.5 cells
    assert np.isclose(((e[hot] - ambient) * dx ** 3).sum() + hot.sum() * ambient * dx ** 3, 1.0, rtol=1e-12)
    x = np.arange(26) * dx - 2.5 * dx - 1.0
    y = np.arange(24) * dx - 2.5 * dx - 0.9
    z = np.arange(22) * dx - 2.5 * dx - 0.8
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    np.testing.assert_array_equal(q[4] > ambient, X * X + Y * Y + Z * Z < (2.5 * dx) ** 2)


def test_netcdf_3d_layout(H, tmp_path):
    netcdf = pytest.importorskip("scipy.io").netcdf_file
    H.write_initial_output(SEDOV_3D, tmp_path)
    q0 = H.initial_state(SEDOV_3D)
    with netcdf(tmp_path / "sedov3d.nc", "r", mmap=False) as f:
        assert (f.dimensions["x"], f.dimensions["y"], f.dimensions["z"], f.dimensions["var"]) == (26, 24, 22, 5)
        a = f._attributes
        assert (a["irho"], a["imx"], a["imy"], a["imz"], a["iene"]) == (0, 1, 2, 3, 4)
        np.testing.assert_array_equal(f.variables["z"][:], np.arange(22) * 0.1 + (-(3 - 0.5) * 0.1) - 0.8)
        np.testing.assert_array_equal(f.variables["Q"][0], q0)


# ---- tabulated LTE EOS table files -------------------------------------------------
def test_lte_table_reader_formats(H, tmp_path):
    """The C++ reader takes CDF-1 and CDF-2 files with int or double scalars and float or double
    tables (the generator writes CDF-2), refuses HDF5 and reports missing variables."""
    from scipy.io import netcdf_file
    from mosscap_amd.lte_table import generate, write_nc
    t = generate(12, 6)
    good = tmp_path / "t2.nc"
    write_nc(t, good)
    base = MIN + "eos:\n  type: tabulatedlteh\n  table_path: {}\n"
    s = H.setup_summary(base.format(good))
    assert s["gamma"] == 5.0 / 3.0

    def check(lt, yt, tt):
        assert (lt["n_eint_rho"], lt["n_rho"]) == (t.n_eint_rho, t.n_rho)
        assert (lt["log_eint_rho_start"], lt["log_eint_rho_step"]) == (t.log_eint_rho_start, t.log_eint_rho_step)
        assert (lt["log_rho_start"], lt["log_rho_step"]) == (t.log_rho_start, t.log_rho_step)
        assert lt["y_sum"] == sum(float(v) for v in yt.ravel())
        assert lt["log_T_sum"] == sum(float(v) for v in tt.ravel())

    check(s["lte_table"], t.y_table, t.log_T_table)
    # CDF-1, float tables, double counts
    p1 = tmp_path / "t1.nc"
    with netcdf_file(str(p1), "w", version=1) as f:
        f.createDimension("a", t.n_rho)
        f.createDimension("b", t.n_eint_rho)
        for nm, val in (("log_eint_rho_start", t.log_eint_rho_start), ("log_eint_rho_step", t.log_eint_rho_step),
                        ("n_eint_rho", float(t.n_eint_rho)), ("log_rho_start", t.log_rho_start),
                        ("log_rho_step", t.log_rho_step), ("n_rho", float(t.n_rho))):
            v = f.createVariable(nm, "d", ())
            v[()] = val
        for nm, arr in (("y_table", t.y_table), ("log_T_table", t.log_T_table)):
            v = f.createVariable(nm, "f", ("a", "b"))
            v[:] = arr.astype(np.float32)
    s1 = H.setup_summary(base.format(p1))
    check(s1["lte_table"], t.y_table.astype(np.float32).astype(np.float64),
          t.log_T_table.astype(np.float32).astype(np.float64))
    h5 = tmp_path / "t.h5"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + bytes(64))
    with pytest.raises(H.HostError, match="HDF5"):
        H.setup_summary(base.format(h5))
    p3 = tmp_path / "t3.nc"
    with netcdf_file(str(p3), "w", version=2) as f:
        v = f.createVariable("log_eint_rho_start", "d", ())
        v[()] = 1.0
    with pytest.raises(H.HostError, match="no variable log_eint_rho_step"):
        H.setup_summary(base.format(p3))


def test_driver_executable_help():
    """The command-line driver links against both libraries and answers --help without a GPU."""
    import subprocess
    exe = Path(__file__).resolve().parent.parent / "mosscap_amd" / "mosscap_hip"
    out = subprocess.run([str(exe), "--help"], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0 and "usage: mosscap_hip" in out.stdout
