"""Pins for the oracle's single functors (CPU only).

The reference ships no golden vectors, so each functor is pinned by closed
forms on exactly representable states, textbook properties (consistency,
reproduction of linear data, ENO behaviour at a jump, order of accuracy) and
hand-worked examples chosen so that a plausible transcription mistake (a
swapped index, a dropped term, an else-if instead of sequential ifs, the HLL
wave-speed quirk "fixed") changes the answer.
"""
import math

import numpy as np
import pytest

G = 1.4


# ---- EOS ------------------------------------------------------------------
def test_cons_prim_closed_form(orc):
    # rho=2, v=(3,4), p=5: E = 5/0.4 + 0.5*2*25 = 37.5 (all exact in binary)
    w = orc.cons_to_prim(2, 1.5, [2.0, 6.0, 8.0, 37.5 - 12.5 + 10.0])  # gamma 1.5: E = 5/0.5 + 25 = 35
    assert list(w) == [2.0, 3.0, 4.0, 5.0]
    q = orc.prim_to_cons(2, 1.5, [2.0, 3.0, 4.0, 5.0])
    assert list(q) == [2.0, 6.0, 8.0, 35.0]
    # 1-D index map: Pres = 2
    assert list(orc.cons_to_prim(1, 1.5, [2.0, 6.0, 19.0])) == [2.0, 3.0, 5.0]


def test_prim_to_flux_closed_form(orc):
    w = [2.0, 3.0, 4.0, 5.0]
    # x: (rho vx, rho vx vx + p, rho vx vy, (E + p) vx) with E = 35 (gamma 1.5)
    assert list(orc.prim_to_flux(0, 2, 1.5, w)) == [6.0, 23.0, 24.0, 120.0]
    # y: (rho vy, rho vy vx, rho vy vy + p, (E + p) vy)
    assert list(orc.prim_to_flux(1, 2, 1.5, w)) == [8.0, 24.0, 37.0, 160.0]


def test_round_trip_random(orc):
    rng = np.random.default_rng(1)
    for _ in range(200):
        w = np.array([rng.uniform(0.1, 10), rng.normal(), rng.normal(), rng.uniform(0.01, 100)])
        w2 = orc.cons_to_prim(2, G, orc.prim_to_cons(2, G, w))
        np.testing.assert_allclose(w2, w, rtol=1e-12, atol=1e-12)


def test_sound_speed(orc):
    assert orc.sound_speed(2, G, [1.4, 0.0, 0.0, 1.0]) == 1.0
    assert orc.sound_speed(1, 2.0, [2.0, 7.0, 8.0]) == math.sqrt(8.0)


# ---- limiters ---------------------------------------------------------------
@pytest.mark.parametrize("um,up,mm,mc,vl", [
    (1.0, 2.0, 1.0, 1.5, 4.0 / 3.0),
    (1.0, 10.0, 1.0, 2.0, 20.0 / 11.0),
    (-1.0, 2.0, 0.0, 0.0, 0.0),
    (-1.0, -3.0, -1.0, -2.0, -1.5),
    (0.0, 1.0, 0.0, 0.0, 0.0),
    (2.0, 2.0, 2.0, 2.0, 2.0),
])
def test_limiters_closed_form(orc, um, up, mm, mc, vl):
    assert orc.slope_limiter("minmod", um, up) == mm
    assert orc.slope_limiter("monotonizedcentral", um, up) == pytest.approx(mc, rel=1e-15)
    assert orc.slope_limiter("vanleer", um, up) == pytest.approx(vl, rel=1e-15)


def test_limiter_signed_zero(orc):
    # copysign sees the sign of zero: MC(-0, +x) -> (-1 + 1) * ... = 0
    assert orc.slope_limiter("monotonizedcentral", -0.0, 1.0) == 0.0
    assert orc.slope_limiter("minmod", 0.0, 1.0) == 0.0


# ---- reconstruction ----------------------------------------------------------
ALL_RECON = [("fog", "minmod"), ("muscl", "vanleer"), ("muscl", "monotonizedcentral"), ("muscl", "minmod"),
             ("ppm", "vanleer"), ("ppm", "monotonizedcentral"), ("ppm", "minmod"), ("weno5z", "minmod")]


@pytest.mark.parametrize("recon,lim", ALL_RECON)
def test_uniform_data_exact(orc, recon, lim):
    for c in (1.0, 2.5, -3.0):
        wl, wr = orc.reconstruct(recon, lim, [[c] * 5])
        assert wl[0] == c and wr[0] == c


@pytest.mark.parametrize("recon,lim", ALL_RECON[1:])
def test_linear_data_reproduced(orc, recon, lim):
    # s = a + k h: face values a -+ h/2 (exact binary fractions)
    a, h = 1.0, 0.25
    wl, wr = orc.reconstruct(recon, lim, [[a - 2 * h, a - h, a, a + h, a + 2 * h]])
    assert wl[0] == a - h / 2 and wr[0] == a + h / 2


def test_fog(orc):
    wl, wr = orc.reconstruct("fog", "minmod", [[9.0, 8.0, 3.0, 7.0, 6.0]])
    assert wl[0] == 3.0 and wr[0] == 3.0


def test_muscl_extremum_flattens(orc):
    for lim in ("vanleer", "monotonizedcentral", "minmod"):
        wl, wr = orc.reconstruct("muscl", lim, [[0.0, 1.0, 3.0, 1.0, 0.0]])
        assert wl[0] == 3.0 and wr[0] == 3.0


def test_ppm_sequential_fixes_worked_example(orc):
    # s = (0, 0, 1, 4, 5), minmod slopes (0, 1, 1): unlimited edges (1/3, 5/2);
    # test (ii) fires (|5/2 - 1| >= 2|1/3 - 1|): wR = 3 - 2/3 = 7/3; (iii) then fails.
    wl, wr = orc.reconstruct("ppm", "minmod", [[0.0, 0.0, 1.0, 4.0, 5.0]])
    assert wl[0] == pytest.approx(1.0 / 3.0, rel=1e-15)
    assert wr[0] == pytest.approx(7.0 / 3.0, rel=1e-15)
    # mirrored data exercises (iii) with the *updated* wR
    wl, wr = orc.reconstruct("ppm", "minmod", [[5.0, 4.0, 1.0, 0.0, 0.0]])
    assert wr[0] == pytest.approx(1.0 / 3.0, rel=1e-15)
    assert wl[0] == pytest.approx(7.0 / 3.0, rel=1e-15)
    # local extremum: flatten to the cell value
    wl, wr = orc.reconstruct("ppm", "monotonizedcentral", [[0.0, 1.0, 3.0, 1.0, 0.0]])
    assert wl[0] == 3.0 and wr[0] == 3.0


def test_weno_eno_at_jump(orc):
    # Step between cells 0 and 1: the smooth left stencil gets all the weight on
    # both faces (the right face uses the flipped indicator order).
    wl, wr = orc.reconstruct("weno5z", "minmod", [[0.0, 0.0, 0.0, 1.0, 1.0]])
    assert abs(wl[0]) < 1e-30 and abs(wr[0]) < 1e-30
    wl, wr = orc.reconstruct("weno5z", "minmod", [[0.0, 0.0, 1.0, 1.0, 1.0]])
    assert abs(wl[0] - 1.0) < 1e-30 and abs(wr[0] - 1.0) < 1e-30


def test_weno_mirror_symmetry(orc):
    rng = np.random.default_rng(3)
    s = rng.normal(size=(500, 5))
    wl, wr = orc.reconstruct("weno5z", "minmod", s)
    wl2, wr2 = orc.reconstruct("weno5z", "minmod", s[:, ::-1])
    np.testing.assert_allclose(wl, wr2, rtol=1e-12, atol=1e-13)
    np.testing.assert_allclose(wr, wl2, rtol=1e-12, atol=1e-13)


def test_weno_fifth_order(orc):
    # face values of sin from exact cell averages: error ~ h^5 on smooth data
    errs = []
    for n in (32, 64, 128):
        h = 2 * np.pi / n
        x = (np.arange(-2, 3) + 0.5) * h + 0.3  # cell edges at 0.3 + k h
        avg = (np.cos(x - h / 2) - np.cos(x + h / 2)) / h
        wl, wr = orc.reconstruct("weno5z", "minmod", [avg])
        errs.append(abs(wl[0] - np.sin(0.3)) + abs(wr[0] - np.sin(0.3 + h)))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > 4.5, (errs, rates)


# ---- Riemann solvers ------------------------------------------------------------
def _state(rho, v, p, nd=2, axis=0):
    w = np.zeros(2 + nd)
    w[0] = rho
    w[1 + axis] = v
    w[1 + nd] = p
    return w


@pytest.mark.parametrize("solver", ["rusanov", "hllc"])
@pytest.mark.parametrize("axis", [0, 1])
def test_consistency(orc, solver, axis):
    rng = np.random.default_rng(7)
    for _ in range(50):
        w = np.array([rng.uniform(0.1, 5), rng.normal(), rng.normal(), rng.uniform(0.1, 5)])
        f = orc.riemann(solver, axis, 2, G, w, w)[0]
        np.testing.assert_allclose(f, orc.prim_to_flux(axis, 2, G, w), rtol=1e-12, atol=1e-12)


def test_hllc_stationary_contact_exact(orc):
    wl = _state(1.0, 0.0, 1.0)
    wr = _state(0.125, 0.0, 1.0)
    assert list(orc.riemann("hllc", 0, 2, G, wl, wr)[0]) == [0.0, 1.0, 0.0, 0.0]
    wl = _state(1.0, 0.0, 1.0, axis=1)
    wr = _state(0.125, 0.0, 1.0, axis=1)
    assert list(orc.riemann("hllc", 1, 2, G, wl, wr)[0]) == [0.0, 0.0, 1.0, 0.0]


def _hll_textbook(wl, wr, sl, sr, axis=0):
    ql, qr = orc_q(wl), orc_q(wr)
    import oracle
    fl = oracle.prim_to_flux(axis, 2, G, wl)
    fr = oracle.prim_to_flux(axis, 2, G, wr)
    return (sr * fl - sl * fr + sr * sl * (qr - ql)) / (sr - sl)


def orc_q(w):
    import oracle
    return oracle.prim_to_cons(2, G, w)


def test_hll_wave_speed_quirk(orc):
    # vL = 0.5, vR = -0.5, csL = 1, csR = 2. Riemann.hpp:60 takes
    # sR = max(vL + csL, vL + csR) = 2.5 (the textbook bound is 1.5); sL = -2.5.
    wl = _state(1.4, 0.5, 1.0)           # cs = 1
    wr = _state(1.4, -0.5, 4.0)          # cs = 2
    f = orc.riemann("hll", 0, 2, G, wl, wr)[0]
    np.testing.assert_allclose(f, _hll_textbook(wl, wr, -2.5, 2.5), rtol=1e-14)
    assert not np.allclose(f, _hll_textbook(wl, wr, -2.5, 1.5), rtol=1e-6)


def test_rusanov_mean_speed(orc):
    # Riemann.hpp:40 uses the mean 0.5*(csL+|vL|+csR+|vR|), not the max
    wl = _state(1.4, 0.5, 1.0)
    wr = _state(1.4, -0.5, 4.0)
    a = 0.5 * (1.0 + 0.5 + 2.0 + 0.5)
    ql, qr = orc_q(wl), orc_q(wr)
    fl, fr = orc.prim_to_flux(0, 2, G, wl), orc.prim_to_flux(0, 2, G, wr)
    np.testing.assert_allclose(orc.riemann("rusanov", 0, 2, G, wl, wr)[0], 0.5 * (fl + fr - a * (qr - ql)),
                               rtol=1e-14)


def test_hllc_supersonic_upwind(orc):
    # all waves moving right: the flux is the left physical flux
    wl = _state(1.0, 5.0, 1.0)
    wr = _state(0.5, 5.0, 0.8)
    np.testing.assert_allclose(orc.riemann("hllc", 0, 2, G, wl, wr)[0], orc.prim_to_flux(0, 2, G, wl),
                               rtol=1e-13)
    wl = _state(1.0, -5.0, 1.0)
    wr = _state(0.5, -5.0, 0.8)
    np.testing.assert_allclose(orc.riemann("hllc", 0, 2, G, wl, wr)[0], orc.prim_to_flux(0, 2, G, wr),
                               rtol=1e-13)


def test_hllc_moving_contact_exact_speed(orc):
    # equal pressure and velocity: the contact moves at v; mass flux = rho_upwind * v
    wl = _state(1.0, 0.3, 1.0)
    wr = _state(0.25, 0.3, 1.0)
    f = orc.riemann("hllc", 0, 2, G, wl, wr)[0]
    assert f[0] == pytest.approx(1.0 * 0.3, rel=1e-14)
    assert f[1] == pytest.approx(1.0 * 0.3 * 0.3 + 1.0, rel=1e-14)
