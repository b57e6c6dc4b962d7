"""ctypes bindings of the host layer (libmosscap_host.so, include/mosscap_host.h):
YAML configuration with the reference's get_or semantics, the problem
registry, netCDF output and the driver loop, all in C++ above the HIP C ABI.
The same layer is the command-line driver ``mosscap_amd/mosscap_hip``.
"""
from __future__ import annotations

import ctypes as C
import json
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
HOST_LIB = _HERE / "libmosscap_host.so"
DRIVER = _HERE / "mosscap_hip"
_lib = None


class HostError(RuntimeError):
    pass


class RunResult(C.Structure):
    _fields_ = [("steps", C.c_int64), ("time", C.c_double), ("outputs", C.c_int32), ("last_dt", C.c_double)]


SYMBOLS = {
    "mhh_setup_summary": [C.c_char_p, C.c_char_p, C.c_int64, C.c_char_p, C.c_int64],
    "mhh_initial_state": [C.c_char_p, C.c_void_p, C.c_int64, C.c_char_p, C.c_int64],
    "mhh_get": [C.c_char_p, C.c_char_p, C.c_int32, C.c_char_p, C.c_char_p, C.c_int64, C.c_char_p, C.c_int64],
    "mhh_write_initial_output": [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int64],
    "mhh_run_config": [C.c_char_p, C.c_char_p, C.c_int32, C.c_void_p, C.c_char_p, C.c_int64],
}


def lib():
    global _lib
    if _lib is None:
        if not HOST_LIB.exists():
            raise HostError(f"{HOST_LIB} is not built (python -m mosscap_amd.build)")
        L = C.CDLL(str(HOST_LIB))
        for name, args in SYMBOLS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_int
        _lib = L
    return _lib


def _call(fn, *args):
    err = C.create_string_buffer(4096)
    rc = fn(*args, err, len(err))
    if rc != 0:
        raise HostError(err.value.decode() or f"error {rc}")


def setup_summary(yaml_text: str) -> dict:
    """Host setup only: the resolved configuration (grid, boundaries, schemes, times, output)."""
    out = C.create_string_buffer(1 << 16)
    _call(lib().mhh_setup_summary, yaml_text.encode(), out, len(out))
    return json.loads(out.value.decode())


def initial_state(yaml_text: str) -> np.ndarray:
    """The problem generator's Q, shaped [var][z][y][x], ghosts included."""
    s = setup_summary(yaml_text)
    nvar = s["num_dim"] + 2 + s["num_tracers"]
    q = np.empty((nvar, s["zc"], s["yc"], s["xc"]))
    _call(lib().mhh_initial_state, yaml_text.encode(), q.ctypes.data, q.size)
    return q


def get_or(yaml_text: str, key: str, kind: str, default) -> str:
    """get_or<T>(config, key, default) with T in int/double/bool/string; the value as text."""
    t = {"int": 0, "double": 1, "bool": 2, "string": 3}[kind]
    out = C.create_string_buffer(4096)
    d = (("true" if default else "false") if kind == "bool" else str(default)).encode()
    _call(lib().mhh_get, yaml_text.encode(), key.encode(), t, d, out, len(out))
    return out.value.decode()


def write_initial_output(yaml_text: str, directory) -> None:
    """Writes the initial state as two output records (t = 0 and t = 0.5) without a GPU."""
    _call(lib().mhh_write_initial_output, yaml_text.encode(), str(directory).encode())


def run_config(yaml_text: str, directory, verbose: bool = False) -> RunResult:
    """setup_sim + the main loop on the GPU, output files into `directory`."""
    r = RunResult()
    _call(lib().mhh_run_config, yaml_text.encode(), str(directory).encode(), int(verbose), C.byref(r))
    return r
