"""Discrete adjoint of the full SSPRK3 step (dt frozen), through the C ABI and
the torch.autograd.Function.

* the hand-written VJPs equal the dual-number Jacobians of the same functors;
* on a small periodic grid the gradient equals torch autograd of an independent
  float64 restatement of the step (tests/torch_hydro.py), rtol 1e-8;
* on the full workload (KH 512^2, 200 steps, checkpoint every 20), along 16
  seeded directions, directional derivatives agree with the exact
  tangent-linear run of the same FP64 map to 1e-10, and with FP64 central
  finite differences to 1e-5 at eps = 1e-5 and 1e-3 at eps = 1e-3 (relative to
  the largest |FD|; the larger eps crosses kinks of the map, see below).
"""
import time

import numpy as np
import pytest

import mosscap_amd as M
from mosscap_amd import problems
from mosscap_amd.adjoint import KHAdjointProblem, kh_adjoint_function

pytestmark = pytest.mark.gpu


def test_weno_vjp_matches_dual():
    rng = np.random.default_rng(0)
    s = rng.normal(size=(20000, 5))
    s[:2000] = np.round(s[:2000])  # ties and flat stencils
    cot = rng.normal(size=(20000, 2))
    np.testing.assert_allclose(M.probe_vjp(0, s, cot), M.probe_vjp(1, s, cot), rtol=1e-9, atol=1e-9)


@pytest.mark.parametrize("solver,axis", [("rusanov", 0), ("rusanov", 1), ("hllc", 0), ("hllc", 1)])
def test_riemann_vjp_matches_dual(solver, axis):
    """Hand-written Riemann VJPs against the dual-number Jacobian of the same functor, over
    sub- and supersonic states (every HLLC branch: p_mid above/below each side, aM sign,
    the bM/bP floors)."""
    rng = np.random.default_rng(1 + axis + 10 * (solver == "hllc"))
    n = 40000
    w = np.empty((n, 8))
    for o in (0, 4):
        w[:, o] = 10 ** rng.uniform(-1, 1, n)
        w[:, o + 1:o + 3] = rng.normal(size=(n, 2)) * np.where(rng.uniform(size=(n, 1)) < 0.3, 5.0, 1.0)
        w[:, o + 3] = 10 ** rng.uniform(-2, 2, n)
    cot = rng.normal(size=(n, 4))
    a = {"rusanov": 2, "hllc": 6}[solver] + 2 * axis
    hand, dual = M.probe_vjp(a, w, cot), M.probe_vjp(a + 1, w, cot)
    scale = np.abs(dual).max(axis=1, keepdims=True)
    np.testing.assert_allclose(hand / scale, dual / scale, rtol=0, atol=1e-11)


def _small_problem(solver, nsteps=6, every=4):
    cfg = problems.kh_config(24)
    cfg.ny = 20
    cfg.max_cfl = 0.4
    return KHAdjointProblem(cfg=cfg, nsteps=nsteps, checkpoint_every=every, riemann_solver=solver, seed=3)


@pytest.mark.parametrize("solver", ["rusanov", "hllc"])
def test_gradient_matches_torch_reference(solver):
    torch = pytest.importorskip("torch")
    from tests import torch_hydro as th

    prob = _small_problem(solver)
    rng = np.random.default_rng(17)
    control = 0.01 * rng.uniform(-1, 1, size=(prob.cfg.ny, prob.cfg.nx))
    prob.forward(control)  # records the frozen dts
    target = prob.forward(0.02 * rng.uniform(-1, 1, size=control.shape))
    rho = prob.forward(control)
    diff = rho - target
    grad_rho = diff / np.linalg.norm(diff)
    g_hip = prob.backward(grad_rho)

    # torch reference with the same dts
    cfg, ng = prob.cfg, prob.cfg.ng
    c = torch.tensor(control, dtype=torch.float64, requires_grad=True)
    base = torch.from_numpy(problems.kh_band(cfg, seed=3)[:, 0].copy())
    vx = torch.from_numpy(prob.vx)
    vy = torch.from_numpy(prob.vy)
    rho0 = base[0, ng:-ng, ng:-ng] + c
    q = base.clone()
    q[0, ng:-ng, ng:-ng] = rho0
    q[1, ng:-ng, ng:-ng] = rho0 * vx
    q[2, ng:-ng, ng:-ng] = rho0 * vy
    q[3, ng:-ng, ng:-ng] = 2.5 / 0.4 + 0.5 * rho0 * (vx * vx + vy * vy)
    q = th.fill_periodic(q, ng)
    for dt in prob.dts:
        q = th.ssprk3_step(q, 1.4, float(dt), cfg.dx, ng, solver)
    loss = torch.linalg.norm(q[0, ng:-ng, ng:-ng] - torch.from_numpy(target))
    loss.backward()
    g_ref = c.grad.numpy()
    np.testing.assert_allclose(rho, q[0, ng:-ng, ng:-ng].detach().numpy(), rtol=1e-12, atol=1e-13)
    scale = np.abs(g_ref).max()
    np.testing.assert_allclose(g_hip, g_ref, rtol=1e-8, atol=1e-8 * scale)


def test_autograd_function_roundtrip():
    torch = pytest.importorskip("torch")
    prob = _small_problem("rusanov", nsteps=4, every=2)
    fn = kh_adjoint_function()
    c = torch.zeros((prob.cfg.ny, prob.cfg.nx), dtype=torch.float32, requires_grad=True)
    target = torch.from_numpy(prob.forward(np.zeros((prob.cfg.ny, prob.cfg.nx)) + 0.003))
    rho = fn.apply(c, prob)
    loss = torch.linalg.norm(rho - target)
    loss.backward()
    assert c.grad is not None and c.grad.dtype == torch.float32 and torch.isfinite(c.grad).all()


@pytest.mark.parametrize("solver", ["rusanov", "hllc"])
def test_full_workload_fd_agreement(solver):
    """KH 512^2, 200 steps, dt frozen, checkpoint every 20; 16 seeded directions.

    Central differences at eps = 1e-3 carry an O(eps) error of their own: the discrete
    map has kinks (|beta0 - beta2| in WENO5-Z, the HLLC and min/max switches) that a
    perturbation of that size crosses in some cells. The deviation from the adjoint
    shrinks with eps (1e-3: <= 5e-4 of the largest |FD|; 1e-5: <= 3e-6), which is the
    check that the adjoint is the derivative of the map it differentiates; the exact check
    is test_full_workload_tangent_agreement."""
    prob = KHAdjointProblem(n=512, nsteps=200, checkpoint_every=20, riemann_solver=solver)
    n = prob.cfg.nx
    control = 0.01 * (2 * problems.uniform01(17, np.arange(n * n)).reshape(n, n) - 1)
    prob.forward(control)  # records the frozen dt sequence
    target = prob.forward(0.01 * (2 * problems.uniform01(18, np.arange(n * n)).reshape(n, n) - 1))

    def loss(c):
        return float(np.linalg.norm(prob.forward(c) - target))

    t0 = time.perf_counter()
    rho = prob.forward(control)
    t_fwd = time.perf_counter() - t0
    diff = rho - target
    t0 = time.perf_counter()
    grad = prob.backward(diff / np.linalg.norm(diff))
    t_bwd = time.perf_counter() - t0
    rel = {}
    for eps in (1e-3, 1e-5):
        dev, fds = [], []
        for k in range(16):
            d = 2 * problems.uniform01(19 + 1000 * k, np.arange(n * n)).reshape(n, n) - 1
            d /= np.linalg.norm(d)
            fd = (loss(control + eps * d) - loss(control - eps * d)) / (2 * eps)
            fds.append(fd)
            dev.append(abs(float((grad * d).sum()) - fd))
        rel[eps] = np.array(dev) / max(abs(f) for f in fds)
    print(f"{solver}: max rel FD deviation eps=1e-3 {rel[1e-3].max():.2e}, eps=1e-5 {rel[1e-5].max():.2e}; "
          f"backward/forward wall {t_bwd / t_fwd:.2f}")
    assert rel[1e-5].max() <= 1e-5, rel[1e-5]
    assert rel[1e-3].max() <= 1e-3, rel[1e-3]


@pytest.mark.parametrize("solver", ["rusanov", "hllc"])
def test_full_workload_tangent_agreement(solver):
    """KH 512^2, 200 steps, dt frozen, checkpoint every 20: the adjoint gradient against the exact
    directional derivatives of the same FP64 map, from the tangent-linear (forward-mode, dual
    number) run of the recorded steps, along the 16 seeded directions of the FD check:
    <dJ/dc, d> = <dJ/drho_N, drho_N/dc . d>, both exact up to rounding, to 1e-10 of the largest
    derivative. The tangent run's value part must reproduce the forward's final density bit for bit
    (it is the same arithmetic on the same branches)."""
    prob = KHAdjointProblem(n=512, nsteps=200, checkpoint_every=20, riemann_solver=solver)
    n = prob.cfg.nx
    control = 0.01 * (2 * problems.uniform01(17, np.arange(n * n)).reshape(n, n) - 1)
    prob.forward(control)  # records the frozen dt sequence
    target = prob.forward(0.01 * (2 * problems.uniform01(18, np.arange(n * n)).reshape(n, n) - 1))
    rho = prob.forward(control)
    diff = rho - target
    lam = diff / np.linalg.norm(diff)
    grad = prob.backward(lam)
    adj, tan = [], []
    for k in range(16):
        d = 2 * problems.uniform01(19 + 1000 * k, np.arange(n * n)).reshape(n, n) - 1
        d /= np.linalg.norm(d)
        drho, rho_t = prob.tangent(d)
        if k == 0:
            np.testing.assert_array_equal(rho_t, rho)
        adj.append(float((grad * d).sum()))
        tan.append(float((lam * drho).sum()))
    adj, tan = np.array(adj), np.array(tan)
    dev = np.abs(adj - tan) / np.abs(tan).max()
    print(f"{solver}: adjoint vs tangent, max deviation {dev.max():.2e} of max |derivative| {np.abs(tan).max():.3e}")
    assert dev.max() <= 1e-10, dev


@pytest.mark.parametrize("bc", [
    ("periodic", "periodic", "periodic", "periodic"),
    ("periodic", "periodic", "wall", "wall"),
    ("zerograd", "zerograd", "periodic", "periodic"),
    ("zerograd", "wall", "symmetric", "zerograd"),
])
@pytest.mark.parametrize("solver", ["rusanov", "hllc"])
def test_tangent_agreement_per_boundary(bc, solver):
    """Small KH grid (40 x 36, 9 steps, checkpoint every 4) on each kind of face: the adjoint
    gradient against the tangent-linear run along 4 directions, to 1e-10 of the largest
    derivative. A periodic axis takes its wrap-around terms inside the line sweeps; the other
    faces fold their ghosts' cotangents into their sources (with the normal momentum negated at
    walls)."""
    cfg = problems.kh_config(40)
    cfg.ny, cfg.max_cfl, cfg.bc = 36, 0.4, bc
    prob = KHAdjointProblem(cfg=cfg, nsteps=9, checkpoint_every=4, riemann_solver=solver, seed=5)
    rng = np.random.default_rng(31)
    control = 0.01 * rng.uniform(-1, 1, size=(cfg.ny, cfg.nx))
    prob.forward(control)
    target = prob.forward(0.02 * rng.uniform(-1, 1, size=control.shape))
    rho = prob.forward(control)
    lam = (rho - target) / np.linalg.norm(rho - target)
    grad = prob.backward(lam)
    adj, tan = [], []
    for k in range(4):
        d = rng.uniform(-1, 1, size=control.shape)
        drho, _ = prob.tangent(d)
        adj.append(float((grad * d).sum()))
        tan.append(float((lam * drho).sum()))
    adj, tan = np.array(adj), np.array(tan)
    dev = np.abs(adj - tan) / np.abs(tan).max()
    assert dev.max() <= 1e-10, (dev, tan)
