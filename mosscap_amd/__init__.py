"""MI355X-native finite-volume Euler hot path of Mosscap: Python bindings.

The product is the HIP library ``libmosscap_hip.so`` behind the C ABI in
``include/mosscap_hip.h``; this module binds it with ctypes for the test
harness, the benchmark and the adjoint's ``torch.autograd.Function``. There is
no CPU fallback: every entry point fails loudly when the library (or a GPU) is
missing.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# MOSSCAP_HIP_LIB selects an alternative build of the same library (tuning experiments)
LIB_PATH = Path(os.environ.get("MOSSCAP_HIP_LIB", str(_HERE / "libmosscap_hip.so")))

# Enum values, as in the reference's name tables (Reconstruct.hpp:14-31,
# Riemann.hpp:15-19, TimeStepping.hpp:15-19, State.hpp:71-79).
RECONSTRUCTION = {"fog": 0, "muscl": 1, "ppm": 2, "weno5z": 3}
SLOPE_LIMITER = {"vanleer": 0, "monotonizedcentral": 1, "minmod": 2}
RIEMANN_SOLVER = {"rusanov": 0, "hll": 1, "hllc": 2}
TIME_STEPPER = {"rk2": 0, "ssprk3": 1, "ssprk4": 2}
BOUNDARY = {"wall": 0, "periodic": 1, "symmetric": 2, "symmetricoutflowdiode": 3, "zerograd": 4,
            "constant": 5, "userfn": 6}
MIN_GHOST = {"fog": 1, "muscl": 2, "ppm": 3, "weno5z": 3}  # Reconstruct.hpp:47-61
NUM_STAGES = {"rk2": 2, "ssprk3": 3, "ssprk4": 4}
FP64, FP32C = 0, 1

MH_OK, MH_ERR_INVALID, MH_ERR_UNSUPPORTED, MH_ERR_HIP = 0, 1, 2, 3


class MhGridSize(C.Structure):
    _fields_ = [("xc", C.c_int32), ("yc", C.c_int32), ("zc", C.c_int32), ("ng", C.c_int32)]


class MhGridLoc(C.Structure):
    _fields_ = [("x", C.c_double), ("y", C.c_double), ("z", C.c_double)]


class MhBoundaries(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("xs", "xe", "ys", "ye", "zs", "ze")] + \
               [(n + "_const", C.c_double * 5) for n in ("xs", "xe", "ys", "ye", "zs", "ze")]


class MhSchemes(C.Structure):
    _fields_ = [("reconstruction", C.c_int32), ("slope_limit", C.c_int32), ("riemann_solver", C.c_int32),
                ("time_stepper", C.c_int32)]


class MhOptions(C.Structure):
    _fields_ = [("precision_form", C.c_int32), ("span_rows", C.c_int32), ("rho_floor", C.c_float),
                ("p_floor", C.c_float)]


class MhStats(C.Structure):
    _fields_ = [("floor_hits", C.c_uint64), ("steps", C.c_int64), ("last_dt", C.c_double), ("time", C.c_double),
                ("prev_output_time", C.c_double)]


class MhLteTable(C.Structure):
    _fields_ = [("log_eint_rho_start", C.c_double), ("log_eint_rho_step", C.c_double), ("n_eint_rho", C.c_int32),
                ("log_rho_start", C.c_double), ("log_rho_step", C.c_double), ("n_rho", C.c_int32),
                ("y_table", C.c_void_p), ("log_T_table", C.c_void_p)]


class MhDeviceView(C.Structure):
    _fields_ = [("precision_form", C.c_int32), ("q", C.c_void_p), ("q_hi", C.c_void_p), ("q_lo", C.c_void_p),
                ("num_values", C.c_int64)]


USER_BC_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)


class MhTiming(C.Structure):
    _fields_ = [("stage_ms", C.c_double), ("stage_launches", C.c_int64), ("ghost_ms", C.c_double),
                ("ghost_launches", C.c_int64), ("clock_ms", C.c_double), ("clock_launches", C.c_int64),
                ("step_ms", C.c_double)]


class MoscapError(RuntimeError):
    """A non-zero status from the C ABI (the reference throws std::runtime_error)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every symbol include/mosscap_hip.h declares
_SIGNATURES = [
    ("mh_create", C.c_int, [C.POINTER(MhGridSize), C.c_double, C.POINTER(MhGridLoc), C.POINTER(MhBoundaries),
                            C.c_int32, C.c_int32, C.POINTER(MhSchemes), C.c_double, C.POINTER(MhOptions),
                            C.POINTER(C.c_void_p)]),
    ("mh_destroy", None, [C.c_void_p]),
    ("mh_last_error", C.c_char_p, [C.c_void_p]),
    ("mh_num_values", C.c_int64, [C.c_void_p]),
    ("mh_upload", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_download", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_upload_device", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_download_device", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_fill_bcs", C.c_int, [C.c_void_p]),
    ("mh_compute_dt", C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                C.POINTER(C.c_double)]),
    ("mh_time_step", C.c_int, [C.c_void_p, C.c_double]),
    ("mh_stage", C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    ("mh_run", C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_double), C.c_double,
                         C.POINTER(C.c_double), C.c_double, C.c_int32]),
    ("mh_run_timed", C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.POINTER(C.c_double), C.c_double,
                               C.POINTER(C.c_double), C.c_double, C.POINTER(MhTiming)]),
    ("mh_autotune", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]),
    ("mh_get_stats", C.c_int, [C.c_void_p, C.POINTER(MhStats)]),
    ("mh_stream", C.c_void_p, [C.c_void_p]),
    ("mh_set_user_bc", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_set_gravity", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_double, C.c_double]),
    ("mh_eos_fields", C.c_int, [C.c_void_p, C.c_int32, C.c_double, C.c_int32, C.c_void_p, C.c_void_p]),
    ("mh_set_tabulated_eos", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_tabulated_eos_fields", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_tabulated_eos_update", C.c_int, [C.c_void_p]),
    ("mh_device_state", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_adjoint_forward", C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_double, C.c_void_p, C.c_void_p]),
    ("mh_adjoint_backward", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_adjoint_backward_ms", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_adjoint_tangent", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_step_noalloc", C.c_int, [C.c_void_p, C.c_double, C.c_void_p, C.c_int32, C.c_void_p, C.c_double,
                                  C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_probe_vjp", C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double]),
    ("mh_batch_create", C.c_int, [C.c_int32, C.c_void_p, C.c_int32, C.c_double, C.POINTER(MhBoundaries),
                                  C.POINTER(MhSchemes), C.c_double, C.POINTER(MhOptions), C.POINTER(C.c_void_p)]),
    ("mh_batch_destroy", None, [C.c_void_p]),
    ("mh_batch_last_error", C.c_char_p, [C.c_void_p]),
    ("mh_batch_num_values", C.c_int64, [C.c_void_p]),
    ("mh_batch_offsets", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_batch_upload", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_batch_download", C.c_int, [C.c_void_p, C.c_void_p]),
    ("mh_batch_fill_bcs", C.c_int, [C.c_void_p]),
    ("mh_batch_run", C.c_int, [C.c_void_p, C.c_int64, C.c_double, C.c_double, C.c_int32, C.c_int32]),
    ("mh_batch_times", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p]),
    ("mh_batch_set_waves", C.c_int, [C.c_void_p, C.c_int32]),
    ("mh_batch_autotune", C.c_int, [C.c_void_p, C.c_int32, C.c_void_p, C.POINTER(C.c_int32), C.c_void_p]),
    ("mh_batch_stream", C.c_void_p, [C.c_void_p]),
    ("mh_batch_stats", C.c_int, [C.c_void_p, C.POINTER(MhStats)]),
    ("mh_probe_reconstruct", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int64, C.c_void_p, C.c_void_p,
                                       C.c_void_p]),
    ("mh_probe_riemann", C.c_int, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_double, C.c_int64, C.c_void_p,
                                   C.c_void_p, C.c_void_p]),
    ("mh_probe_math", C.c_int, [C.c_int32, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p]),
]
SYMBOLS = [s[0] for s in _SIGNATURES]


def lib() -> C.CDLL:
    """Load the HIP library (built in-tree by ``python -m mosscap_amd.build``)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m mosscap_amd.build` "
                              "(there is no CPU fallback)")
        h = C.CDLL(str(LIB_PATH))
        for name, res, args in _SIGNATURES:
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _check(code: int, ctx=None) -> None:
    if code != MH_OK:
        msg = lib().mh_last_error(ctx)
        raise MoscapError(code, msg.decode() if msg else "unknown error")


def _ptr(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


@dataclass
class Config:
    """Problem/scheme configuration (the YAML keys of SetupModules.hpp:9-177)."""
    num_dim: int = 2
    nx: int = 256
    ny: int = 256
    nz: int = 1  # interior planes of a 3-D grid (grid.z)
    ng: int = 3
    x_dim: float = 1.0
    x_start: float = 0.0
    y_start: float = 0.0
    z_start: float = 0.0
    bc: tuple = ("wall", "wall", "wall", "wall")  # xs, xe, ys, ye[, zs, ze]
    bc_const: tuple = ((0.0,) * 5,) * 4  # conserved values per face, same order as bc
    reconstruction: str = "muscl"
    slope_limiter: str = "monotonizedcentral"
    riemann_solver: str = "hll"
    time_stepper: str = "ssprk3"
    gamma: float = 1.4
    max_cfl: float = 0.8
    max_time: float = 1.0
    output_delta: float = 0.1
    num_tracers: int = 0  # simulation.n_extra_fields
    gravity: tuple | None = None  # (gx, gy, gz) of a gravity source term (Gravity.cpp), or None

    @property
    def xc(self) -> int:
        return self.nx + 2 * self.ng

    @property
    def yc(self) -> int:
        return self.ny + 2 * self.ng if self.num_dim > 1 else 1

    @property
    def zc(self) -> int:
        return self.nz + 2 * self.ng if self.num_dim > 2 else 1

    @property
    def nvar(self) -> int:
        return 2 + self.num_dim + self.num_tracers

    @property
    def dx(self) -> float:
        return self.x_dim / self.nx  # setup_grid, SetupModules.hpp:29-31

    @property
    def shape(self) -> tuple:
        return (self.nvar, self.zc, self.yc, self.xc)


def _abi_structs(cfg: Config):
    """The grid size, boundaries and schemes of `cfg` as C-ABI structs."""
    sz = MhGridSize(cfg.xc, cfg.yc, cfg.zc, cfg.ng)
    b = MhBoundaries()
    names = ("xs", "xe", "ys", "ye", "zs", "ze")
    for k, (n, kind) in enumerate(zip(names, cfg.bc)):
        setattr(b, n, BOUNDARY[kind.lower()])
        arr = getattr(b, n + "_const")
        for i, v in enumerate(cfg.bc_const[k] if k < len(cfg.bc_const) else ()):
            arr[i] = v
    sch = MhSchemes(RECONSTRUCTION[cfg.reconstruction], SLOPE_LIMITER[cfg.slope_limiter],
                    RIEMANN_SOLVER[cfg.riemann_solver], TIME_STEPPER[cfg.time_stepper])
    return sz, b, sch


def step_noalloc(cfg: Config, q_in_dev: int, q_out_dev: int, dt: float = 0.0, max_cfl: float | None = None,
                 stream: int = 0) -> float:
    """One FP64 step of a caller-owned device state without device allocations (mh_step_noalloc):
    q_in_dev / q_out_dev are device addresses of ghost-filled [var][k][j][i] float64 arrays. Returns
    the dt used (the CFL dt of q_in when dt <= 0)."""
    sz, b, sch = _abi_structs(cfg)
    used = C.c_double()
    _check(lib().mh_step_noalloc(C.byref(sz), cfg.dx, C.byref(b), cfg.num_dim, C.byref(sch), cfg.gamma,
                                 cfg.max_cfl if max_cfl is None else max_cfl, dt, C.c_void_p(q_in_dev),
                                 C.c_void_p(q_out_dev), C.byref(used), C.c_void_p(stream or None)))
    return used.value


class HydroContext:
    """One problem on the GPU: owns the C-ABI context."""

    def __init__(self, cfg: Config, precision_form: int = FP64, span_rows: int = 0,
                 rho_floor: float = 0.0, p_floor: float = 0.0):
        L = lib()
        self.cfg = cfg
        sz, b, sch = _abi_structs(cfg)
        loc = MhGridLoc(cfg.x_start, cfg.y_start, cfg.z_start)
        opt = MhOptions(precision_form, span_rows, rho_floor, p_floor)
        h = C.c_void_p()
        _check(L.mh_create(C.byref(sz), cfg.dx, C.byref(loc), C.byref(b), cfg.num_dim, cfg.num_tracers, C.byref(sch),
                           cfg.gamma,
                           C.byref(opt), C.byref(h)))
        self._h = h
        self.n_values = L.mh_num_values(h)
        if cfg.gravity is not None:
            self.set_gravity(cfg.gravity)

    def eos_fields(self, which: int = 1, avg_mass: float = 1.0, ionisation_e: bool = False):
        """Analytic LTE hydrogen (y, T) of every cell, shaped [z][y][x]: which=1 the most recent stage's
        input (what update_eos saw), which=0 the current state."""
        shape = (self.cfg.zc, self.cfg.yc, self.cfg.xc)
        n = int(np.prod(shape))
        y, T = np.empty(n), np.empty(n)
        self._c(lib().mh_eos_fields(self._h, which, avg_mass, int(ionisation_e), _ptr(y), _ptr(T)))
        return y.reshape(shape), T.reshape(shape)

    def set_tabulated_eos(self, table) -> None:
        """eos.type = tabulatedlteh with `table` (a mosscap_amd.lte_table.LteTable), or None to remove it."""
        if table is None:
            self._c(lib().mh_set_tabulated_eos(self._h, None))
            return
        y = np.ascontiguousarray(table.y_table, dtype=np.float64)
        lt = np.ascontiguousarray(table.log_T_table, dtype=np.float64)
        assert y.shape == lt.shape == (table.n_rho, table.n_eint_rho)
        t = MhLteTable(table.log_eint_rho_start, table.log_eint_rho_step, table.n_eint_rho, table.log_rho_start,
                       table.log_rho_step, table.n_rho, y.ctypes.data, lt.ctypes.data)
        self._c(lib().mh_set_tabulated_eos(self._h, C.byref(t)))

    def tabulated_eos_update(self) -> None:
        """update_eos on the current state in place (setup_sim's call before its first ghost fill)."""
        self._c(lib().mh_tabulated_eos_update(self._h))

    def tabulated_eos_fields(self):
        """(y, T) kept by the tabulated EOS's last update_eos, shaped [z][y][x]."""
        shape = (self.cfg.zc, self.cfg.yc, self.cfg.xc)
        y, T = np.empty(shape), np.empty(shape)
        self._c(lib().mh_tabulated_eos_fields(self._h, _ptr(y), _ptr(T)))
        return y, T

    def set_gravity(self, g) -> None:
        """Register (gx, gy, gz) as a gravity source term, or remove it with None."""
        if g is None:
            self._c(lib().mh_set_gravity(self._h, 0, 0.0, 0.0, 0.0))
        else:
            self._c(lib().mh_set_gravity(self._h, 1, float(g[0]), float(g[1]), float(g[2])))

    def close(self) -> None:
        if getattr(self, "_h", None):
            lib().mh_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def _c(self, code):
        _check(code, self._h)

    def upload(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.n_values, (q.shape, self.n_values)
        self._c(lib().mh_upload(self._h, _ptr(q)))

    def download(self) -> np.ndarray:
        q = np.empty(self.cfg.shape, dtype=np.float64)
        self._c(lib().mh_download(self._h, _ptr(q)))
        return q

    def upload_device(self, ptr: int) -> None:
        self._c(lib().mh_upload_device(self._h, C.c_void_p(ptr)))

    def download_device(self, ptr: int) -> None:
        self._c(lib().mh_download_device(self._h, C.c_void_p(ptr)))

    def fill_bcs(self) -> None:
        self._c(lib().mh_fill_bcs(self._h))

    def compute_dt(self, time: float, prev_output_time: float, max_cfl: float | None = None,
                   max_time: float | None = None, delta: float | None = None) -> float:
        c = self.cfg
        dt = C.c_double()
        self._c(lib().mh_compute_dt(self._h, c.max_cfl if max_cfl is None else max_cfl, time,
                                    c.max_time if max_time is None else max_time, prev_output_time,
                                    c.output_delta if delta is None else delta, C.byref(dt)))
        return dt.value

    def time_step(self, dt: float) -> None:
        self._c(lib().mh_time_step(self._h, dt))

    def stage(self, stage: int, dt: float) -> None:
        self._c(lib().mh_stage(self._h, stage, dt))

    def run(self, nsteps: int, time: float = 0.0, prev_output_time: float = 0.0, use_graph: bool = True,
            max_cfl: float | None = None, max_time: float | None = None, delta: float | None = None):
        c = self.cfg
        t = C.c_double(time)
        p = C.c_double(prev_output_time)
        self._c(lib().mh_run(self._h, nsteps, c.max_cfl if max_cfl is None else max_cfl, C.byref(t),
                             c.max_time if max_time is None else max_time, C.byref(p),
                             c.output_delta if delta is None else delta, 1 if use_graph else 0))
        return t.value, p.value

    def run_timed(self, nsteps: int, time: float = 0.0, prev_output_time: float = 0.0) -> tuple:
        c = self.cfg
        t = C.c_double(time)
        p = C.c_double(prev_output_time)
        tm = MhTiming()
        self._c(lib().mh_run_timed(self._h, nsteps, c.max_cfl, C.byref(t), c.max_time, C.byref(p),
                                   c.output_delta, C.byref(tm)))
        return t.value, p.value, tm

    def autotune(self, steps: int = 3, candidates=None):
        """Pick the span height (rows per wavefront; -1 the default layout) by timing; returns
        (best, {span: ms/step})."""
        cand = np.array(list(candidates or [4, 8, 12, 16, 18, 20, 24, 28, 32, 48, 64, -1]) + [0], dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int32()
        self._c(lib().mh_autotune(self._h, steps, _ptr(cand), C.byref(best), _ptr(ms)))
        return best.value, {int(c): float(m) for c, m in zip(cand[:-1], ms[:-1]) if m > 0}

    def stats(self) -> MhStats:
        s = MhStats()
        self._c(lib().mh_get_stats(self._h, C.byref(s)))
        return s

    def adjoint_forward(self, nsteps: int, checkpoint_every: int = 20, dts=None, max_cfl: float | None = None):
        """Forward steps recording dt (frozen for the backward pass) and checkpoints; returns the dts."""
        out = np.empty(nsteps)
        din = None if dts is None else np.ascontiguousarray(dts, dtype=np.float64)
        self._c(lib().mh_adjoint_forward(self._h, nsteps, checkpoint_every,
                                         self.cfg.max_cfl if max_cfl is None else max_cfl,
                                         None if din is None else _ptr(din), _ptr(out)))
        return out

    def adjoint_backward(self, lambda_final: np.ndarray) -> np.ndarray:
        """dJ/dQ^0 from dJ/dQ^N (arrays shaped like Q; ghost entries ignored / zero)."""
        lf = np.ascontiguousarray(lambda_final, dtype=np.float64).reshape(self.cfg.shape)
        out = np.empty(self.cfg.shape)
        self._c(lib().mh_adjoint_backward(self._h, _ptr(lf), _ptr(out)))
        return out

    def adjoint_tangent(self, dq0: np.ndarray):
        """Forward-mode derivative of the recorded run along dq0 (shaped like Q; ghost entries
        ignored): returns (dQ^N, Q^N), both ghost-filled; Q^N repeats the forward result."""
        d = np.ascontiguousarray(dq0, dtype=np.float64).reshape(self.cfg.shape)
        dq, q = np.empty(self.cfg.shape), np.empty(self.cfg.shape)
        self._c(lib().mh_adjoint_tangent(self._h, _ptr(d), _ptr(dq), _ptr(q)))
        return dq, q

    def adjoint_backward_ms(self) -> float:
        """Device milliseconds of the last adjoint_backward (recompute + reverse, transfers excluded)."""
        ms = C.c_double()
        self._c(lib().mh_adjoint_backward_ms(self._h, C.byref(ms)))
        return ms.value

    @property
    def stream(self) -> int:
        return lib().mh_stream(self._h) or 0

    def set_user_bc(self, fn) -> None:
        """Install fn(view: MhDeviceView) -> None as the user_bc hook (None removes it). It runs after
        every ghost fill with the device pointers of the buffer just filled; an exception in fn
        fails the library call that ran it."""
        if fn is None:
            self._c(lib().mh_set_user_bc(self._h, None, None))
            self._user_bc = None
            return

        def hook(h, _user):
            try:
                v = MhDeviceView()
                if lib().mh_device_state(h, C.byref(v)) != 0:
                    return 1
                fn(v)
                return 0
            except Exception:  # reported through the library's return code
                import traceback
                traceback.print_exc()
                return 1

        self._user_bc = USER_BC_FN(hook)  # keep the trampoline alive
        self._c(lib().mh_set_user_bc(self._h, C.cast(self._user_bc, C.c_void_p), None))


class BatchContext:
    """An ensemble of independent 1-D problems of uneven length on the GPU (mh_batch_*)."""

    def __init__(self, lengths, ng=3, x_dim=1.0, bc=("zerograd", "zerograd"), bc_const=((0.0,) * 5,) * 2,
                 reconstruction="weno5z", slope_limiter="monotonizedcentral", riemann_solver="hllc",
                 time_stepper="ssprk3", gamma=1.4, precision_form=FP64, rho_floor=0.0, p_floor=0.0):
        self.lengths = np.ascontiguousarray(lengths, dtype=np.int32)
        self.ng = ng
        b = MhBoundaries()
        for n, kind, cv in zip(("xs", "xe"), bc, bc_const):
            setattr(b, n, BOUNDARY[kind.lower()])
            arr = getattr(b, n + "_const")
            for i, v in enumerate(cv):
                arr[i] = v
        sch = MhSchemes(RECONSTRUCTION[reconstruction], SLOPE_LIMITER[slope_limiter], RIEMANN_SOLVER[riemann_solver],
                        TIME_STEPPER[time_stepper])
        opt = MhOptions(precision_form, 0, rho_floor, p_floor)
        h = C.c_void_p()
        code = lib().mh_batch_create(len(self.lengths), _ptr(self.lengths), ng, x_dim, C.byref(b), C.byref(sch),
                                     gamma, C.byref(opt), C.byref(h))
        if code != MH_OK:
            raise MoscapError(code, lib().mh_batch_last_error(None).decode())
        self._h = h
        self.n_values = lib().mh_batch_num_values(h)
        off = np.zeros(len(self.lengths), dtype=np.int64)
        self._c(lib().mh_batch_offsets(h, _ptr(off)))
        self.offsets = off

    def _c(self, code):
        if code != MH_OK:
            raise MoscapError(code, lib().mh_batch_last_error(self._h).decode())

    def close(self):
        if getattr(self, "_h", None):
            lib().mh_batch_destroy(self._h)
            self._h = None

    __del__ = close

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def block(self, q: np.ndarray, p: int) -> np.ndarray:
        """View of problem p's [3][n_p + 2 ng] block inside the packed array q."""
        xc = int(self.lengths[p]) + 2 * self.ng
        o = int(self.offsets[p])
        return q[o:o + 3 * xc].reshape(3, xc)

    def upload(self, q):
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == self.n_values
        self._c(lib().mh_batch_upload(self._h, _ptr(q)))

    def download(self):
        q = np.empty(self.n_values)
        self._c(lib().mh_batch_download(self._h, _ptr(q)))
        return q

    def fill_bcs(self):
        self._c(lib().mh_batch_fill_bcs(self._h))

    def run(self, nsteps, max_cfl=0.4, max_time=1e30, persistent=True, use_graph=True):
        self._c(lib().mh_batch_run(self._h, nsteps, max_cfl, max_time, 0 if persistent else 1, 1 if use_graph else 0))

    def times(self):
        t = np.empty(len(self.lengths))
        dt = np.empty(len(self.lengths))
        self._c(lib().mh_batch_times(self._h, _ptr(t), _ptr(dt)))
        return t, dt

    def set_waves(self, waves: int):
        self._c(lib().mh_batch_set_waves(self._h, waves))

    def autotune(self, steps=3, waves_per_cu=None):
        cand = np.array(list(waves_per_cu or [4, 8, 12, 16, 20, 24, 28, 32]) + [0], dtype=np.int32)
        ms = np.zeros(len(cand))
        best = C.c_int32()
        self._c(lib().mh_batch_autotune(self._h, steps, _ptr(cand), C.byref(best), _ptr(ms)))
        return best.value, {int(c): float(m) for c, m in zip(cand[:-1], ms[:-1]) if m > 0}

    def stats(self) -> MhStats:
        s = MhStats()
        self._c(lib().mh_batch_stats(self._h, C.byref(s)))
        return s

    @property
    def stream(self) -> int:
        return lib().mh_batch_stream(self._h) or 0


def run_to_time(ctx: HydroContext, time: float = 0.0, prev_output_time: float = 0.0, max_steps: int = -1):
    """main.cpp:54-71 on the host: while time < max_time { dt; step; output bookkeeping }."""
    c = ctx.cfg
    steps = 0
    dts = []
    while time < c.max_time and (max_steps < 0 or steps < max_steps):
        dt = ctx.compute_dt(time, prev_output_time)
        ctx.time_step(dt)
        time += dt
        dts.append(dt)
        steps += 1
        if time >= prev_output_time + c.output_delta:
            prev_output_time = time
    return time, prev_output_time, np.array(dts)


def probe_reconstruct(recon: str, limiter: str, stencils: np.ndarray, prec: int = FP64):
    s = np.ascontiguousarray(stencils, dtype=np.float64).reshape(-1, 5)
    n = s.shape[0]
    wl = np.empty(n)
    wr = np.empty(n)
    _check(lib().mh_probe_reconstruct(RECONSTRUCTION[recon], SLOPE_LIMITER[limiter], prec, n, _ptr(s), _ptr(wl),
                                      _ptr(wr)))
    return wl, wr


def probe_riemann(solver: str, axis: int, nd: int, gamma: float, wl: np.ndarray, wr: np.ndarray,
                  prec: int = FP64) -> np.ndarray:
    a = np.ascontiguousarray(wl, dtype=np.float64).reshape(-1, 2 + nd)
    b = np.ascontiguousarray(wr, dtype=np.float64).reshape(-1, 2 + nd)
    f = np.empty_like(a)
    _check(lib().mh_probe_riemann(RIEMANN_SOLVER[solver], axis, nd, prec, gamma, a.shape[0], _ptr(a), _ptr(b),
                                  _ptr(f)))
    return f


def probe_vjp(what: int, inputs: np.ndarray, cot: np.ndarray, gamma: float = 1.4) -> np.ndarray:
    nin = 5 if what < 2 else 8
    a = np.ascontiguousarray(inputs, dtype=np.float64).reshape(-1, nin)
    c = np.ascontiguousarray(cot, dtype=np.float64)
    out = np.empty_like(a)
    _check(lib().mh_probe_vjp(what, a.shape[0], _ptr(a), _ptr(c), _ptr(out), gamma))
    return out


def probe_math(op: int, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    a = np.ascontiguousarray(a, dtype=np.float64)
    bb = np.ascontiguousarray(b if b is not None else a, dtype=np.float64)
    out = np.empty_like(a)
    _check(lib().mh_probe_math(op, a.size, _ptr(a), _ptr(bb), _ptr(out)))
    return out
