"""Build recipe for the HIP library (gfx950) and the CPU oracle.

    python -m mosscap_amd.build            # both
    python -m mosscap_amd.build --lib      # HIP library only

Outputs (git-ignored; built in-tree so they travel with the source tree):
    mosscap_amd/libmosscap_hip.so   the product: kernels + C ABI (include/mosscap_hip.h)
    mosscap_amd/libmosscap_host.so  host layer: YAML setup, problems, output, driver (include/mosscap_host.h)
    mosscap_amd/mosscap_hip         command-line driver (the reference's main.cpp)
    oracle/_build/libmh_oracle.so   the CPU oracle (test infrastructure)

Every HIP translation unit is compiled with packed-FP32 instructions
disabled. The FP64 drop-in units use -ffp-contract=off so no multiply-add is
fused (the reference CPU build has none); the compensated-FP32 unit allows
contraction (-ffp-contract=fast-honor-pragmas) except where a pragma pins it
off: the error-free transforms of the stage combine and the CFL signal speed
(which the stage kernels and the separate CFL kernel must compute to the same
bits). Plain -ffp-contract=fast lets the backend fuse across the pragmas.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "mosscap_amd"
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libmosscap_hip.so"
HOST_LIB = PKG / "libmosscap_host.so"
HOST_EXE = PKG / "mosscap_hip"
HOST_UNITS = ["yaml_lite.cpp", "setup.cpp", "problems.cpp", "output_nc.cpp", "table_nc.cpp", "host_capi.cpp"]
ORACLE_SRC = ROOT / "oracle" / "mh_oracle.cpp"
ORACLE_LIB = ROOT / "oracle" / "_build" / "libmh_oracle.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "-O3", f"--offload-arch={ARCH}", "-fPIC", "-std=c++20", "-x", "hip",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]

# (source, contraction mode)
UNITS = [
    ("kernels/stage_f64.hip", "off"),
    ("kernels/stage_f32c.hip", "fast-honor-pragmas"),
    ("kernels/stage3d_f64.hip", "off"),
    ("kernels/stage3d_f32c.hip", "fast-honor-pragmas"),
    ("kernels/common.hip", "off"),
    ("kernels/batch_f64.hip", "off"),
    ("kernels/batch_f32c.hip", "fast-honor-pragmas"),
    ("kernels/adjoint.hip", "off"),
    ("kernels/tangent.hip", "off"),
    ("capi.cpp", "off"),
    ("capi_batch.cpp", "off"),
    ("capi_adjoint.cpp", "off"),
    ("capi_noalloc.cpp", "off"),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(line for line in proc.stdout.splitlines()
                    if "packed-fp32-ops" not in line and "is not a recognized feature" not in line)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")


def build_lib(jobs: int = 4, defines: list[str] | None = None, out: Path | None = None) -> Path:
    """Compile the HIP units and link the shared library (`defines`: extra -D
    knobs, e.g. MH_STAGE_WAVES_PER_EU=2, for experiments built to `out`)."""
    lib = Path(out) if out else LIB
    build = BUILD if out is None else lib.parent / (lib.stem + "_obj")
    build.mkdir(parents=True, exist_ok=True)
    extra = [f"-D{d}" for d in (defines or [])]
    objs = []
    cmds = []
    for src, contract in UNITS:
        obj = build / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([HIPCC, *COMMON, *extra, f"-ffp-contract={contract}", "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = lib.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp)])
    os.replace(tmp, lib)
    return lib


def _git_hash() -> str:
    try:
        r = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "--short", "HEAD"], stdout=subprocess.PIPE,
                           stderr=subprocess.DEVNULL, text=True)
        return r.stdout.strip() or "unknown"
    except OSError:
        return "unknown"


def build_host() -> Path:
    """The host layer (plain C++, no device code) linked against the HIP library,
    and the command-line driver. -ffp-contract=off keeps the initial-condition
    and output arithmetic identical to the Python generators."""
    cxx = shutil.which("g++") or "g++"
    flags = ["-O2", "-std=c++17", "-fPIC", "-ffp-contract=off", "-Wall", f"-I{ROOT / 'include'}",
             f'-DMOSSCAP_GIT_HASH="{_git_hash()}"']
    tmp = HOST_LIB.with_suffix(".so.tmp")
    _run([cxx, *flags, "-shared", *[str(CSRC / "host" / u) for u in HOST_UNITS], "-o", str(tmp),
          f"-L{PKG}", "-lmosscap_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_LIB)
    tmp = HOST_EXE.with_suffix(".tmp")
    _run([cxx, *flags, str(CSRC / "host" / "main.cpp"), "-o", str(tmp), f"-L{PKG}", "-lmosscap_host",
          "-lmosscap_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, HOST_EXE)
    return HOST_LIB


def build_oracle() -> Path:
    ORACLE_LIB.parent.mkdir(parents=True, exist_ok=True)
    cxx = shutil.which("g++") or "g++"
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cxx, "-O3", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
          str(ORACLE_SRC), "-o", str(tmp)])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="HIP library only")
    ap.add_argument("--oracle", action="store_true", help="oracle only")
    ap.add_argument("-j", type=int, default=4)
    ap.add_argument("-D", dest="defines", action="append", default=[], help="extra preprocessor knob")
    ap.add_argument("--out", default=None, help="library path (experiments)")
    a = ap.parse_args(argv)
    if not a.oracle:
        print("built", build_lib(a.j, a.defines, a.out), flush=True)
        if a.out is None:
            print("built", build_host(), flush=True)
    if not a.lib:
        print("built", build_oracle(), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
