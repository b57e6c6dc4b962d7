"""Tables for the tabulated LTE hydrogen EOS (eos.type = tabulatedlteh).

The reference reads `eos.table_path` (default mosscap_lte_h_tables.nc) with
the scalars log_eint_rho_start/step, n_eint_rho, log_rho_start/step, n_rho and
the 2-D variables y_table and log_T_table, [n_rho][n_eint_rho]
(TabulatedLteH.hpp:30-55). This module holds that table in memory, reads and
writes it as a netCDF classic file, and builds one from the LTE equilibrium
of pure hydrogen:

    eint = p / (gamma - 1) + rho (chi_H / m_H) y,   p = rho (k_B / m_H) (1 + y) T,
    y    = Saha ionisation fraction at n_H = rho / m_H and T (AnalyticLteH.hpp:9-19),

solved for T at every (rho, eint / rho) node by bisection in log T. That
internal energy is the one update_eos keeps consistent (its new_eint,
TabulatedLteH.hpp:150-152), so a state built from the table's (y, T) reads
back the same node.

    python -m mosscap_amd.lte_table out.nc [--n-eint-rho 256 --n-rho 128 ...]
"""
from __future__ import annotations

import argparse
from dataclasses import dataclass

import numpy as np

H_MASS = 1.6737830080950003e-27  # kg
K_B = 1.380649e-23  # J/K
CHI_H = 2.178710282685096e-18  # J


@dataclass
class LteTable:
    log_eint_rho_start: float
    log_eint_rho_step: float
    n_eint_rho: int
    log_rho_start: float
    log_rho_step: float
    n_rho: int
    y_table: np.ndarray  # [n_rho][n_eint_rho]
    log_T_table: np.ndarray

    def axes(self):
        """(log10 eint/rho, log10 rho) of the nodes."""
        return (self.log_eint_rho_start + self.log_eint_rho_step * np.arange(self.n_eint_rho),
                self.log_rho_start + self.log_rho_step * np.arange(self.n_rho))


def saha_y(nh_tot, temp):
    """Ionisation fraction of pure hydrogen in LTE: the root of y^2 n / (1 - y) = X(T) of the Saha
    relation in AnalyticLteH.hpp:9-19, written as 2X / (X + sqrt(X^2 + 4 n X)) so that it keeps its
    precision when X >> n (fully ionised)."""
    x = 2.4146830395719654e21 * temp ** 1.5 * np.exp(-157763.42386247337 / temp)
    with np.errstate(invalid="ignore"):
        y = 2.0 * x / (x + np.sqrt(x * x + 4.0 * nh_tot * x))
    return np.where(x > 0.0, y, 0.0)


def lte_eint(rho, temp, gamma):
    """Internal energy density [J m^-3] and y of LTE hydrogen at (rho, T)."""
    y = saha_y(rho / H_MASS, temp)
    p = rho * (K_B / H_MASS) * (1.0 + y) * temp
    return p / (gamma - 1.0) + rho * (CHI_H / H_MASS) * y, y


def generate(n_eint_rho: int = 256, n_rho: int = 128, log_eint_rho=(7.0, 13.0), log_rho=(-12.0, -2.0),
             gamma: float = 5.0 / 3.0, t_min: float = 100.0, t_max: float = 1e9) -> LteTable:
    """Table on log-spaced axes: log10(eint/rho) in [log_eint_rho], log10(rho) in [log_rho]
    (end points included). Nodes below t_min or above t_max are clamped there."""
    ex_step = (log_eint_rho[1] - log_eint_rho[0]) / (n_eint_rho - 1)
    r_step = (log_rho[1] - log_rho[0]) / (n_rho - 1)
    ex = log_eint_rho[0] + ex_step * np.arange(n_eint_rho)
    lr = log_rho[0] + r_step * np.arange(n_rho)
    rho = 10.0 ** lr[:, None] * np.ones((1, n_eint_rho))
    target = rho * 10.0 ** ex[None, :]
    lo = np.full(rho.shape, np.log10(t_min))
    hi = np.full(rho.shape, np.log10(t_max))
    for _ in range(80):  # eint(T) increases with T
        mid = 0.5 * (lo + hi)
        e, _ = lte_eint(rho, 10.0 ** mid, gamma)
        below = e < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    log_t = 0.5 * (lo + hi)
    _, y = lte_eint(rho, 10.0 ** log_t, gamma)
    return LteTable(float(log_eint_rho[0]), float(ex_step), int(n_eint_rho), float(log_rho[0]), float(r_step),
                    int(n_rho), np.ascontiguousarray(y), np.ascontiguousarray(log_t))


def write_nc(t: LteTable, path) -> None:
    """netCDF classic (64-bit offset) file with the reference's variable names."""
    from scipy.io import netcdf_file

    with netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("n_rho", t.n_rho)
        f.createDimension("n_eint_rho", t.n_eint_rho)
        for name, value, kind in (("log_eint_rho_start", t.log_eint_rho_start, "d"),
                                  ("log_eint_rho_step", t.log_eint_rho_step, "d"),
                                  ("n_eint_rho", t.n_eint_rho, "i"), ("log_rho_start", t.log_rho_start, "d"),
                                  ("log_rho_step", t.log_rho_step, "d"), ("n_rho", t.n_rho, "i")):
            v = f.createVariable(name, kind, ())
            v[()] = value
        for name, a in (("y_table", t.y_table), ("log_T_table", t.log_T_table)):
            v = f.createVariable(name, "d", ("n_rho", "n_eint_rho"))
            v[:] = a


def read_nc(path) -> LteTable:
    from scipy.io import netcdf_file

    with netcdf_file(str(path), "r", mmap=False) as f:
        v = f.variables
        return LteTable(float(v["log_eint_rho_start"][()]), float(v["log_eint_rho_step"][()]),
                        int(v["n_eint_rho"][()]), float(v["log_rho_start"][()]),
                        float(v["log_rho_step"][()]), int(v["n_rho"][()]),
                        np.array(v["y_table"][:], dtype=np.float64), np.array(v["log_T_table"][:], dtype=np.float64))


def _main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="Write a tabulated LTE hydrogen EOS table (netCDF classic).")
    ap.add_argument("path")
    ap.add_argument("--n-eint-rho", type=int, default=256)
    ap.add_argument("--n-rho", type=int, default=128)
    ap.add_argument("--log-eint-rho", type=float, nargs=2, default=(7.0, 13.0))
    ap.add_argument("--log-rho", type=float, nargs=2, default=(-12.0, -2.0))
    ap.add_argument("--gamma", type=float, default=5.0 / 3.0)
    a = ap.parse_args(argv)
    write_nc(generate(a.n_eint_rho, a.n_rho, a.log_eint_rho, a.log_rho, a.gamma), a.path)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
