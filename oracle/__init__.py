"""TEST INFRASTRUCTURE: ctypes wrapper of the CPU oracle (mh_oracle.cpp).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg may use
this module; the product never does. The oracle restates Mosscap's unfused
hydro sequence op for op (file:line citations in mh_oracle.cpp).

Functions whose transcription is pinned only by hand-worked examples and
invariants (the reference has no golden vectors): riemann_flux<Hllc>,
riemann_flux<Hll> (its sR quirk), reconstruct<Weno5Z>, reconstruct<Ppm>'s
monotonicity branch, fill_one_bc_impl's index maps and the dt clamp. See
tests/test_oracle_*.py for the pins.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
# MH_ORACLE_LIB points the checks at another build of the oracle (a build with a
# planted mistake shows which pins catch it)
LIB_PATH = Path(os.environ.get("MH_ORACLE_LIB", _HERE / "_build" / "libmh_oracle.so"))

RECON = {"fog": 0, "muscl": 1, "ppm": 2, "weno5z": 3}
LIMITER = {"vanleer": 0, "monotonizedcentral": 1, "minmod": 2}
SOLVER = {"rusanov": 0, "hll": 1, "hllc": 2}
STEPPER = {"rk2": 0, "ssprk3": 1, "ssprk4": 2}
BC = {"wall": 0, "periodic": 1, "symmetric": 2, "symmetricoutflowdiode": 3, "zerograd": 4, "constant": 5,
      "userfn": 6}


class OrcConfig(C.Structure):
    _fields_ = [("num_dim", C.c_int32), ("num_tracers", C.c_int32), ("xc", C.c_int32), ("yc", C.c_int32),
                ("zc", C.c_int32), ("ng", C.c_int32), ("dx", C.c_double), ("loc", C.c_double * 3),
                ("bc", C.c_int32 * 6), ("bc_const", (C.c_double * 5) * 6), ("recon", C.c_int32),
                ("limiter", C.c_int32), ("solver", C.c_int32), ("stepper", C.c_int32), ("gamma", C.c_double),
                ("gravity", C.c_int32), ("grav", C.c_double * 3)]


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} missing: run `python -m mosscap_amd.build --oracle`")
        h = C.CDLL(str(LIB_PATH))
        vp, d, i32, i64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64
        sig = {
            "orc_create": (vp, [C.POINTER(OrcConfig)]),
            "orc_destroy": (None, [vp]),
            "orc_num_values": (i64, [vp]),
            "orc_set_q": (None, [vp, vp]),
            "orc_get_q": (None, [vp, vp]),
            "orc_get_w": (None, [vp, vp]),
            "orc_fill_bcs": (None, [vp]),
            "orc_cons_to_prim_all": (None, [vp]),
            "orc_compute_dt": (d, [vp, d, d, d, d, d]),
            "orc_time_step": (None, [vp, d]),
            "orc_begin_step": (None, [vp]),
            "orc_stage": (None, [vp, C.c_int, d]),
            "orc_run": (i64, [vp, d, C.POINTER(d), d, C.POINTER(d), d, i64, vp]),
            "orc_cons_to_prim": (None, [C.c_int, d, vp, vp]),
            "orc_prim_to_cons": (None, [C.c_int, d, vp, vp]),
            "orc_prim_to_flux": (None, [C.c_int, C.c_int, d, vp, vp]),
            "orc_sound_speed": (d, [C.c_int, d, vp]),
            "orc_slope_limiter": (d, [C.c_int, d, d]),
            "orc_reconstruct": (None, [C.c_int, C.c_int, i64, vp, vp, vp]),
            "orc_riemann": (None, [C.c_int, C.c_int, C.c_int, d, i64, vp, vp, vp]),
            "orc_lte_eos": (C.c_int, [C.c_int, i64, vp, d, d, C.c_int, vp, vp]),
            "orc_set_tabulated_eos": (None, [vp, d, d, C.c_int32, d, d, C.c_int32, vp, vp]),
            "orc_tabulated_fields": (C.c_int, [vp, vp, vp]),
            "orc_set_y_space": (None, [vp, vp]),
            "orc_update_eos": (None, [vp]),
            "orc_tab_sample": (None, [d, d, C.c_int32, d, d, C.c_int32, vp, vp, i64, vp, vp, vp, vp]),
            "orc_profile_count": (C.c_int, []),
            "orc_profile_name": (C.c_char_p, [C.c_int]),
            "orc_profile_ms": (d, [C.c_int]),
            "orc_profile_reset": (None, []),
        }
        for name, (res, args) in sig.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def _p(a: np.ndarray) -> C.c_void_p:
    return C.c_void_p(a.ctypes.data)


class OracleSim:
    """The reference's Simulation state for one problem, on the CPU."""

    def __init__(self, cfg):
        """cfg: mosscap_amd.Config (or anything with the same attributes)."""
        self.cfg = cfg
        c = OrcConfig()
        c.num_dim = cfg.num_dim
        c.num_tracers = getattr(cfg, "num_tracers", 0)
        c.xc, c.yc, c.zc, c.ng = cfg.xc, cfg.yc, getattr(cfg, "zc", 1), cfg.ng
        c.dx = cfg.dx
        c.loc[0], c.loc[1], c.loc[2] = cfg.x_start, cfg.y_start, getattr(cfg, "z_start", 0.0)
        for k in range(6):
            c.bc[k] = BC[cfg.bc[k].lower()] if k < len(cfg.bc) else BC["wall"]
            for v in range(5):
                c.bc_const[k][v] = cfg.bc_const[k][v] if k < len(cfg.bc_const) else 0.0
        c.recon = RECON[cfg.reconstruction]
        c.limiter = LIMITER[cfg.slope_limiter]
        c.solver = SOLVER[cfg.riemann_solver]
        c.stepper = STEPPER[cfg.time_stepper]
        c.gamma = cfg.gamma
        g = getattr(cfg, "gravity", None)  # (gx, gy, gz) of a registered gravity source, or None
        if g is not None:
            c.gravity = 1
            for k in range(3):
                c.grav[k] = g[k]
        self._c = c
        self._h = lib().orc_create(C.byref(c))

    def __del__(self):
        if getattr(self, "_h", None):
            lib().orc_destroy(self._h)
            self._h = None

    @property
    def shape(self):
        return self.cfg.shape

    def set_q(self, q: np.ndarray) -> None:
        q = np.ascontiguousarray(q, dtype=np.float64)
        assert q.size == lib().orc_num_values(self._h)
        lib().orc_set_q(self._h, _p(q))

    def get_q(self) -> np.ndarray:
        q = np.empty(self.shape)
        lib().orc_get_q(self._h, _p(q))
        return q

    def get_w(self) -> np.ndarray:
        w = np.empty(self.shape)
        lib().orc_get_w(self._h, _p(w))
        return w

    def set_tabulated_eos(self, t) -> None:
        """eos.type = tabulatedlteh with table t (attributes log_eint_rho_start/step, n_eint_rho,
        log_rho_start/step, n_rho, y_table and log_T_table [n_rho][n_eint_rho])."""
        self._tab = (np.ascontiguousarray(t.y_table, dtype=np.float64),
                     np.ascontiguousarray(t.log_T_table, dtype=np.float64))
        lib().orc_set_tabulated_eos(self._h, t.log_eint_rho_start, t.log_eint_rho_step, t.n_eint_rho,
                                    t.log_rho_start, t.log_rho_step, t.n_rho, _p(self._tab[0]), _p(self._tab[1]))

    def tabulated_fields(self):
        """(y, T, eint_negative) of the last update_eos, shaped [z][y][x]."""
        shape = self.shape[1:]
        y, T = np.empty(shape), np.empty(shape)
        neg = lib().orc_tabulated_fields(self._h, _p(y), _p(T))
        return y, T, bool(neg)

    def set_y_space(self, y: np.ndarray) -> None:
        y = np.ascontiguousarray(y, dtype=np.float64)
        lib().orc_set_y_space(self._h, _p(y))

    def update_eos(self) -> None:
        lib().orc_update_eos(self._h)

    def fill_bcs(self) -> None:
        lib().orc_fill_bcs(self._h)

    def cons_to_prim_all(self) -> None:
        lib().orc_cons_to_prim_all(self._h)

    def compute_dt(self, time, prev_output_time, max_cfl=None, max_time=None, delta=None) -> float:
        c = self.cfg
        return lib().orc_compute_dt(self._h, c.max_cfl if max_cfl is None else max_cfl, time,
                                    c.max_time if max_time is None else max_time, prev_output_time,
                                    c.output_delta if delta is None else delta)

    def time_step(self, dt: float) -> None:
        lib().orc_time_step(self._h, dt)

    def begin_step(self) -> None:
        lib().orc_begin_step(self._h)

    def stage(self, s: int, dt: float) -> None:
        lib().orc_stage(self._h, s, dt)

    def run(self, time=0.0, prev_output_time=0.0, max_steps=-1, max_cfl=None, max_time=None, delta=None):
        c = self.cfg
        t = C.c_double(time)
        p = C.c_double(prev_output_time)
        cap = max_steps if max_steps >= 0 else 1_000_000
        dts = np.zeros(cap)
        n = lib().orc_run(self._h, c.max_cfl if max_cfl is None else max_cfl, C.byref(t),
                          c.max_time if max_time is None else max_time, C.byref(p),
                          c.output_delta if delta is None else delta, max_steps, _p(dts))
        return t.value, p.value, dts[:n]


# ---- single functors ----------------------------------------------------------
def cons_to_prim(nd, gamma, q):
    q = np.ascontiguousarray(q, dtype=np.float64)
    w = np.empty(2 + nd)
    lib().orc_cons_to_prim(nd, gamma, _p(q), _p(w))
    return w


def prim_to_cons(nd, gamma, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    q = np.empty(2 + nd)
    lib().orc_prim_to_cons(nd, gamma, _p(w), _p(q))
    return q


def prim_to_flux(axis, nd, gamma, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    f = np.empty(2 + nd)
    lib().orc_prim_to_flux(axis, nd, gamma, _p(w), _p(f))
    return f


def sound_speed(nd, gamma, w):
    w = np.ascontiguousarray(w, dtype=np.float64)
    return lib().orc_sound_speed(nd, gamma, _p(w))


def slope_limiter(name, um, up):
    return lib().orc_slope_limiter(LIMITER[name], um, up)


def reconstruct(recon, limiter, stencils):
    s = np.ascontiguousarray(stencils, dtype=np.float64).reshape(-1, 5)
    n = s.shape[0]
    wl = np.empty(n)
    wr = np.empty(n)
    lib().orc_reconstruct(RECON[recon], LIMITER[limiter], n, _p(s), _p(wl), _p(wr))
    return wl, wr


def riemann(solver, axis, nd, gamma, wl, wr):
    a = np.ascontiguousarray(wl, dtype=np.float64).reshape(-1, 2 + nd)
    b = np.ascontiguousarray(wr, dtype=np.float64).reshape(-1, 2 + nd)
    f = np.empty_like(a)
    lib().orc_riemann(SOLVER[solver], axis, nd, gamma, a.shape[0], _p(a), _p(b), _p(f))
    return f


def profile(reset: bool = False) -> dict:
    """Accumulated wall milliseconds per reference kernel name since the last reset."""
    L = lib()
    out = {L.orc_profile_name(k).decode(): L.orc_profile_ms(k) for k in range(L.orc_profile_count())}
    if reset:
        L.orc_profile_reset()
    return {k: v for k, v in out.items() if v > 0.0}


def tab_sample(t, log_eint_rho, log_rho):
    """TabulatedLteH::sample on arrays of points: (y, log10 T)."""
    a = np.ascontiguousarray(log_eint_rho, dtype=np.float64).ravel()
    b = np.ascontiguousarray(log_rho, dtype=np.float64).ravel()
    yt = np.ascontiguousarray(t.y_table, dtype=np.float64)
    lt = np.ascontiguousarray(t.log_T_table, dtype=np.float64)
    y, lT = np.empty(a.size), np.empty(a.size)
    lib().orc_tab_sample(t.log_eint_rho_start, t.log_eint_rho_step, t.n_eint_rho, t.log_rho_start, t.log_rho_step,
                         t.n_rho, _p(yt), _p(lt), a.size, _p(a), _p(b), _p(y), _p(lT))
    return y, lT


def lte_eos(nd, q, gamma, avg_mass=1.0, ionisation_e=False):
    """Analytic LTE hydrogen EOS fields (ionisation fraction y, temperature T) of a state
    q[var][...] (every cell), AnalyticLteH.hpp:51-125. Raises on an inverted bracket."""
    q = np.ascontiguousarray(q, dtype=np.float64)
    n = q[0].size
    y, T = np.empty(n), np.empty(n)
    rc = lib().orc_lte_eos(nd, n, _p(q), gamma, avg_mass, int(ionisation_e), _p(y), _p(T))
    if rc:
        raise RuntimeError("Temperature bounds flipped!")
    return y.reshape(q[0].shape), T.reshape(q[0].shape)

