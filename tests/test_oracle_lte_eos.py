"""Oracle pins for the analytic LTE hydrogen EOS (AnalyticLteH.hpp:51-125): the
ionisation fraction solves the Saha equation n_H y^2 / (1 - y) = X(T) at the
temperature found, that temperature reproduces the cell's internal energy to
the bisection's 0.1 K tolerance (with and without the ionisation energy),
the limits (cold: y -> 0; hot and tenuous: y -> 1) come out, and an
inverted temperature bracket is reported."""
import numpy as np
import pytest

H_MASS, K_B, CHI_H = 1.6737830080950003e-27, 1.380649e-23, 2.178710282685096e-18


def _state(rho, T_guess, y_guess, gamma, vx=0.3):
    """Conserved state whose internal energy is that of (rho, y_guess, T_guess) without ionisation energy."""
    eint = rho * (K_B / H_MASS) * (1 + y_guess) * T_guess / (gamma - 1)
    q = np.empty((4,) + np.shape(rho))
    q[0] = rho
    q[1] = rho * vx
    q[2] = 0.0
    q[3] = eint + 0.5 * rho * vx * vx
    return q, eint


def _saha_X(T):
    return 2.4146830395719654e21 * T ** 1.5 * np.exp(-157763.42386247337 / T)


def _temperature_of_y(y, nh):
    """Invert the Saha solution y(T) (increasing in T) by bisection: the trial temperature the
    returned y belongs to."""
    lo, hi = np.full_like(y, 10.0), np.full_like(y, 1e8)
    target = nh * y * y / np.maximum(1 - y, 1e-300)
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        up = _saha_X(mid) < target
        lo, hi = np.where(up, mid, lo), np.where(up, hi, mid)
    return 0.5 * (lo + hi)


@pytest.mark.parametrize("ionisation", [False, True])
def test_saha_and_energy(orc, ionisation):
    """y solves Saha at a trial temperature T_y, T_y satisfies the energy equation
    eint = rho (k/m)(1+y) T/(gamma-1) [+ y rho chi/m] to the bisection tolerance, and the
    returned temperature is the bracket's lower end (T <= T_y)."""
    gamma = 5.0 / 3.0
    rng = np.random.default_rng(7)
    rho = 10 ** rng.uniform(-10, -5, 400)
    T0 = 10 ** rng.uniform(3.6, 4.6, 400)
    q, eint = _state(rho, T0, 0.5, gamma)
    y, T = orc.lte_eos(2, q, gamma, ionisation_e=ionisation)
    nh = rho / H_MASS
    mid = (y > 1e-6) & (y < 1 - 1e-6)  # where the inversion is well conditioned
    T_y = _temperature_of_y(y[mid], nh[mid])
    # the root of T = T_energy(y_saha(T)), solved independently to 1e-9 K
    e_to_T = (gamma - 1) / (rho[mid] * (K_B / H_MASS))
    ion = 1.0 if ionisation else 0.0

    def f(Tt):
        X = _saha_X(Tt)
        ys = 0.5 * (-X + np.sqrt(X * X + 4 * nh[mid] * X)) / nh[mid]
        return Tt - e_to_T / (1 + ys) * (eint[mid] - ion * ys * rho[mid] * (CHI_H / H_MASS))

    lo, hi = np.full(mid.sum(), 100.0), e_to_T * eint[mid]
    for _ in range(200):
        m = 0.5 * (lo + hi)
        neg = f(m) <= 0
        lo, hi = np.where(neg, m, lo), np.where(neg, hi, m)
    root = 0.5 * (lo + hi)
    # y belongs to the last trial temperature, within the final bracket (<= 0.2 K) of the root
    assert np.all(np.abs(T_y - root) <= 0.25 + 1e-9 * root)
    assert np.all(T[mid] <= T_y + 1e-3 + 1e-7 * T_y)
    assert mid.sum() > 100
    assert np.all((y >= 0) & (y <= 1)) and np.all(T >= 100.0)


def test_limits(orc):
    gamma = 5.0 / 3.0
    # cold and dense: neutral; hot and tenuous (where the reference's closed form is still well
    # conditioned, 4 n_H / X >> 1e-16): fully ionised
    q, _ = _state(np.array([1e-3, 1e-8]), np.array([2000.0, 3e4]), np.array([0.0, 1.0]), gamma, vx=0.0)
    y, T = orc.lte_eos(2, q, gamma)
    assert y[0] < 1e-10 and abs(T[0] - 2000.0) < 0.2
    assert y[1] > 0.999 and abs(T[1] - 3e4) / 3e4 < 1e-3


def test_inverted_bracket_is_reported(orc):
    gamma = 5.0 / 3.0
    q, _ = _state(np.array([1e-6]), np.array([50.0]), np.array([0.0]), gamma, vx=0.0)  # below the 100 K floor
    with pytest.raises(RuntimeError, match="flipped"):
        orc.lte_eos(2, q, gamma)
