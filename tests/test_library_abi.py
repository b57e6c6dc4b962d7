"""The C-ABI library loads and exports every symbol include/*.h declares (CPU;
no compute calls), and host-side argument checking refuses bad setups the way
the reference's setup does."""
import ctypes as C
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent


def declared_functions():
    names = []
    for h in sorted((ROOT / "include").glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", h.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(mh_\w+)\s*\(", text, flags=re.M)
    return sorted(set(names))


def test_header_parsed():
    names = declared_functions()
    assert "mh_create" in names and "mh_run" in names and len(names) >= 15


def test_library_exports_every_declared_symbol(mh):
    lib = mh.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python bindings cover the header
    assert set(declared_functions()) <= set(mh.SYMBOLS)


def test_create_rejects_bad_setups(mh):
    # these fail on the host before any device call
    cfg = mh.Config(num_dim=2, nx=16, ny=16, ng=2, reconstruction="weno5z")
    with pytest.raises(mh.MoscapError, match="Insufficient ghost cells"):
        mh.HydroContext(cfg)
    cfg = mh.Config(num_dim=2, nx=16, ny=16, ng=3, bc=("periodic", "wall", "wall", "wall"))
    with pytest.raises(mh.MoscapError, match="periodic"):
        mh.HydroContext(cfg)
    cfg = mh.Config(num_dim=3, nx=16, ny=16, ng=3)
    with pytest.raises(mh.MoscapError):
        mh.HydroContext(cfg)
