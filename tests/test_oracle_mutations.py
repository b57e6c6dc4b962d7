"""Planted mistakes in the oracle are caught by its pins (CPU only).

Each case copies oracle/mh_oracle.cpp with one plausible transcription error
planted, builds it, and runs the pins against that build (MH_ORACLE_LIB): at
least one pin must fail. The cases are the branches no invariant fixes on its
own: HLLC's shock factor (gamma+1)/(2 gamma) (Riemann.hpp:105-110), the floor
of its contact pressure (130), the sign of the PVRS velocity term (103), the
right wave speed's sound speed (111), HLL's -1e-20 clamp (59), WENO5-Z's
squared Z-indicators (Reconstruct.hpp:256-258, 268-270) and its indicator
weight 1/4 (224), and the MC limiter's 1/4 (Reconstruct.hpp:139).
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "oracle" / "mh_oracle.cpp"
PINS = ["tests/test_oracle_pins_exact.py", "tests/test_oracle_functors.py"]

MUTATIONS = {
    "hllc_shock_factor_gm1": [
        ("(gamma + 1.0) / (2.0 * gamma) * (p_mid / wL[ie]", "(gamma - 1.0) / (2.0 * gamma) * (p_mid / wL[ie]"),
        ("(gamma + 1.0) / (2.0 * gamma) * (p_mid / wR[ie]", "(gamma - 1.0) / (2.0 * gamma) * (p_mid / wR[ie]")],
    "hllc_no_contact_pressure_floor": [
        ("const double cp = std::max((mL * tR + mR * tL) / (mL + mR), 0.0);",
         "const double cp = (mL * tR + mR * tL) / (mL + mR);")],
    "hllc_pvrs_velocity_sign": [
        ("(pL + pR + (wL[iv] - wR[iv]) * rho_avg * cs_avg)", "(pL + pR + (wR[iv] - wL[iv]) * rho_avg * cs_avg)")],
    "hllc_right_speed_uses_csL": [
        ("const double aR = wR[iv] + csR * cRstar;", "const double aR = wR[iv] + csL * cRstar;")],
    "hll_no_tiny_clamp": [
        ("const double sL = std::min(-tiny, std::min(wL[iv] - csL, wR[iv] - csR));",
         "const double sL = std::min(wL[iv] - csL, wR[iv] - csR);")],
    "weno_exponent_1": [
        ("double a0 = 3.0 * (1.0 + sq(ind0));", "double a0 = 3.0 * (1.0 + (ind0));"),
        ("double a1 = 6.0 * (1.0 + sq(ind1));", "double a1 = 6.0 * (1.0 + (ind1));"),
        ("double a2 = 1.0 * (1.0 + sq(ind2));", "double a2 = 1.0 * (1.0 + (ind2));"),
        ("a0 = 3.0 * (1.0 + sq(ind2));", "a0 = 3.0 * (1.0 + (ind2));"),
        ("a1 = 6.0 * (1.0 + sq(ind1));\n        a2 = 1.0 * (1.0 + sq(ind0));",
         "a1 = 6.0 * (1.0 + (ind1));\n        a2 = 1.0 * (1.0 + (ind0));")],
    "weno_indicator_weight": [
        ("const double c0 = 13.0 / 12.0, c1 = 0.25;", "const double c0 = 13.0 / 12.0, c1 = 0.75;")],
    "mc_limiter_half": [
        ("std::min(0.25 * std::abs(um + up), std::abs(up))", "std::min(0.5 * std::abs(um + up), std::abs(up))")],
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not found")
    src = SRC.read_text()
    base = tmp_path_factory.mktemp("oracle_mutants")
    jobs = {}
    for name, edits in MUTATIONS.items():
        text = src
        for old, new in edits:
            assert text.count(old) == 1, (name, old)  # the planted edit still applies to the oracle
            text = text.replace(old, new)
        d = base / name
        d.mkdir()
        (d / "mh_oracle.cpp").write_text(text)
        jobs[name] = [cxx, "-O1", "-std=c++17", "-fopenmp", "-ffp-contract=off", "-fPIC", "-shared",
                      str(d / "mh_oracle.cpp"), "-o", str(d / "lib.so")]
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        futs = {n: ex.submit(subprocess.run, c, capture_output=True, text=True) for n, c in jobs.items()}
        for n, f in futs.items():
            r = f.result()
            assert r.returncode == 0, r.stderr
    return {n: base / n / "lib.so" for n in MUTATIONS}


@pytest.mark.parametrize("name", sorted(MUTATIONS))
def test_planted_mistake_is_caught(mutant_libs, name):
    env = dict(os.environ, MH_ORACLE_LIB=str(mutant_libs[name]))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", *PINS], cwd=ROOT,
                       env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 1, f"no pin failed for {name}:\n{r.stdout[-2000:]}"
    assert "failed" in r.stdout
