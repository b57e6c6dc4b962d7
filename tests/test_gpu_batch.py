"""Batched ensemble of uneven 1-D problems (one persistent launch per stage):
every problem against its own oracle run, bitwise in FP64; the persistent and
one-launch-per-problem modes, graphs on and off, give identical states."""
import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import BatchContext, Config, problems

pytestmark = pytest.mark.gpu


def _small(nprob=24, seed=7, n_max=400):
    lengths, q0, _ = problems.riemann_batch(nprob=nprob, seed=seed, n_min=40, n_max=n_max)
    return lengths, q0


@pytest.mark.parametrize("recon,solver,stepper", [("weno5z", "hllc", "ssprk3"), ("muscl", "hll", "rk2"),
                                                  ("ppm", "rusanov", "ssprk4")])
def test_batch_fp64_matches_oracle_per_problem(orc, recon, solver, stepper):
    lengths, q0 = _small()
    nsteps = 15
    kw = dict(reconstruction=recon, slope_limiter="minmod", riemann_solver=solver, time_stepper=stepper)
    with BatchContext(lengths, **kw) as b:
        b.upload(q0)
        b.fill_bcs()
        b.run(nsteps, max_cfl=0.4)
        q = b.download()
        times, _ = b.times()
        for p, n in enumerate(lengths):
            cfg = Config(num_dim=1, nx=int(n), ng=3, bc=("zerograd", "zerograd", "wall", "wall"), max_cfl=0.4,
                         max_time=1e30, output_delta=1e30, **kw)
            sim = orc.OracleSim(cfg)
            sim.set_q(b.block(q0, p).reshape(cfg.shape))
            sim.fill_bcs()
            t, _, _ = sim.run(max_steps=nsteps)
            assert times[p] == t, p
            np.testing.assert_array_equal(b.block(q, p), sim.get_q().reshape(3, -1), err_msg=f"problem {p}")


@pytest.mark.parametrize("form,bc", [(M.FP64, "zerograd"), (M.FP32C, "zerograd"), (M.FP32C, "periodic")])
def test_batch_modes_identical(form, bc):
    """FP32C persistent stages read ghosts from their sources (one fill per run); the
    one-launch-per-problem mode fills after every stage: both give the same bits. A few
    near-vacuum members of this ensemble go non-finite (NaN in the same cells in every mode,
    which the comparison accepts); their clocks are not compared, since a NaN signal speed
    folds into the CFL max as whatever bit pattern the kernel that meets it produced."""
    lengths, q0 = _small(nprob=40, seed=3, n_max=900)
    kw = dict(rho_floor=1e-10, p_floor=1e-12) if form == M.FP32C else {}
    out = []
    finite = None
    for persistent, graph in [(True, True), (True, False), (False, False), (False, True)]:
        with BatchContext(lengths, precision_form=form, bc=(bc, bc), **kw) as b:
            b.upload(q0)
            b.fill_bcs()
            if persistent:
                b.set_waves(7)  # a grid much smaller than the work list
            b.run(6, persistent=persistent, use_graph=graph)
            q = b.download()
            out.append((q, b.times()[0]))
            if finite is None:
                finite = np.array([np.isfinite(b.block(q, p)).all() for p in range(len(lengths))])
    assert finite.sum() >= len(lengths) - 4, finite
    for q, t in out[1:]:
        np.testing.assert_array_equal(q, out[0][0])
        if form == M.FP64:
            np.testing.assert_array_equal(t[finite], out[0][1][finite])
        else:
            # the compensated form's clocks may differ in the last bit between the two modes'
            # kernels (its stage uses dt rounded to FP32, so the states above still agree)
            np.testing.assert_allclose(t[finite], out[0][1][finite], rtol=1e-15, atol=0)


def test_batch_fp32c_close_to_fp64():
    """FP32C with positivity floors stays finite on every member; on the members the
    plain FP64 method (which, like the reference, has no floors) keeps finite, the
    two agree to rel L1 <= 1e-5. Seed 11 contains one near-vacuum double
    rarefaction (problem 11) on which FP64 goes non-finite."""
    lengths, q0 = _small(nprob=64, seed=11, n_max=2000)
    res = []
    for form in (M.FP64, M.FP32C):
        kw = dict(rho_floor=1e-10, p_floor=1e-12) if form == M.FP32C else {}
        with BatchContext(lengths, precision_form=form, **kw) as b:
            b.upload(q0)
            b.fill_bcs()
            b.run(100)
            res.append((b, b.download()))
    (b64, r), (b32, a) = res
    bad = [p for p in range(len(lengths)) if not np.isfinite(b32.block(a, p)).all()]
    assert not bad, f"non-finite FP32C members {bad}"
    finite = [p for p in range(len(lengths)) if np.isfinite(b64.block(r, p)).all()]
    assert len(finite) >= len(lengths) - 2
    num = sum(np.abs(b32.block(a, p) - b64.block(r, p)).sum() for p in finite)
    den = sum(np.abs(b64.block(r, p)).sum() for p in finite)
    assert num / den < 1e-5, num / den


@pytest.mark.parametrize("form", [M.FP64, M.FP32C])
def test_batch_autotune_restores_state(form):
    """The trial runs of the autotuner leave no trace: state, per-problem time and last dt, and the
    floor-hit count are what they were before it (FP32C: seed 11 holds near-vacuum members, and
    p_floor 1e-2 sits at the bottom of the ensemble's pressure range, so the floors fire)."""
    lengths, q0 = _small(nprob=32, seed=11 if form == M.FP32C else 5, n_max=600)
    kw = dict(rho_floor=1e-10, p_floor=1e-2) if form == M.FP32C else {}
    with BatchContext(lengths, precision_form=form, **kw) as a, BatchContext(lengths, precision_form=form, **kw) as b:
        for c in (a, b):
            c.upload(q0)
            c.fill_bcs()
            c.run(2)
        best, table = a.autotune(steps=2, waves_per_cu=[1, 4])
        assert len(table) == 2
        for x, y in zip(a.times(), b.times()):
            np.testing.assert_array_equal(x, y)
        assert a.stats().floor_hits == b.stats().floor_hits
        if form == M.FP32C:
            assert a.stats().floor_hits > 0
        np.testing.assert_array_equal(a.download(), b.download())
        a.run(5)
        b.run(5)
        np.testing.assert_array_equal(a.download(), b.download())


@pytest.mark.parametrize("stepper", ["rk2", "ssprk3", "ssprk4"])
def test_batch_graph_follows_run_arguments(stepper):
    """Graph-replayed persistent steps (for three or more stages the per-problem dt and clock
    update run inside the first two stages) equal the launch-by-launch steps bit for bit, also
    when max_cfl and max_time change between calls (the captured steps are rebuilt); a
    problem that reaches max_time stops there (its later steps have dt = 0)."""
    lengths, q0 = _small(nprob=30, seed=11, n_max=500)
    calls = [(6, 0.4, 1e30), (5, 0.25, 1e30), (40, 0.4, 0.05), (3, 0.3, 0.05)]
    out = []
    for graph in (False, True):
        with BatchContext(lengths, time_stepper=stepper) as b:
            b.upload(q0)
            b.fill_bcs()
            for n, cfl, tmax in calls:
                b.run(n, max_cfl=cfl, max_time=tmax, use_graph=graph)
            times, dts = b.times()
            out.append((b.download(), times, dts))
    np.testing.assert_array_equal(out[1][0], out[0][0])
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])
    assert np.any(out[0][1] == 0.05)  # some problems ended exactly at max_time
