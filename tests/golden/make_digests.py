"""Writes digests of the oracle's FP64 end states for the full-length workloads.

* kh2048_fp64_200.json: KH shear band 2048 x 2048 (WENO5-Z + HLLC + SSPRK3,
  CFL 0.4, periodic, ng 3, the seeded state of `problems.kh_band`), 200 steps.
* sedov256_fp64_t005.json: Sedov blast 256 x 256 on [-1, 1]^2 (PPM-MC + HLL +
  SSPRK3, CFL 0.3, zerograd, ambient p 1e-5), the driver loop to t = 0.05.

A full state is too large to commit (135 MB for KH 2048^2), so each fixture
holds the SHA-256 of its bytes (every cell, ghosts included, float64
little-endian, the layout `download()` returns), the number of steps, the
simulated time, and per-variable sums of the interior as a coarse
cross-check. The GPU tests compare their FP64 results' digests with these
(bit-exact) and run the oracle live only when they differ, to report the rel
L1 per variable.

    python tests/golden/make_digests.py [kh2048] [sedov256]    # KH: ~3 min on 8 cores
"""
from __future__ import annotations

import hashlib
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

HERE = Path(__file__).resolve().parent


def digest(q: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(q, dtype="<f8").tobytes()).hexdigest()


def interior_sums(q: np.ndarray, g: int) -> list:
    a = q[:, 0, g:-g, g:-g]
    return [float(a[v].sum()) for v in range(a.shape[0])]


def record(name: str, workload: str, cfg, q0, max_steps: int) -> dict:
    import oracle
    sim = oracle.OracleSim(cfg)
    sim.set_q(q0)
    sim.fill_bcs()
    t0 = time.perf_counter()
    t, p, dts = sim.run(max_steps=max_steps)
    el = time.perf_counter() - t0
    q = sim.get_q()
    rec = {"workload": workload, "steps": int(len(dts)), "time": t, "time_hex": float(t).hex(),
           "prev_output_time": p, "sha256": digest(q), "input_sha256": digest(q0),
           "interior_sums": interior_sums(q, cfg.ng), "oracle_seconds": round(el, 1)}
    (HERE / name).write_text(json.dumps(rec, indent=1) + "\n")
    print(name, json.dumps(rec, indent=1))
    return rec


def main(argv) -> int:
    from mosscap_amd import problems
    which = set(argv) or {"kh2048", "sedov256"}
    if "kh2048" in which:
        cfg = problems.kh_config(2048)
        record("kh2048_fp64_200.json",
               "KH 2048x2048 periodic, ng 3, WENO5-Z + HLLC + SSPRK3, CFL 0.4, seed 20261202, 200 steps",
               cfg, problems.kh_band(cfg), 200)
    if "sedov256" in which:
        cfg = problems.sedov_config(256)
        record("sedov256_fp64_t005.json",
               "Sedov 256x256 on [-1,1]^2, ng 3, PPM(MC) + HLL + SSPRK3, CFL 0.3, zerograd, to t = 0.05",
               cfg, problems.sedov(cfg), -1)
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
