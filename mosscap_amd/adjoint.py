"""Discrete adjoint of the KH workload behind a torch.autograd.Function.

The control is a density perturbation delta_rho on the interior (float32, as
the workload states); the map is delta_rho -> Q^0 (the KH band with the
perturbation added to rho before prim->cons) -> N frozen-dt steps of the
FP64 GPU path -> final interior density. The backward pass runs the library's
hand-written adjoint (mh_adjoint_backward) and maps dJ/dQ^0 back through the
initial-state construction: d(rho, m_x, m_y, E)/d(delta_rho) = (1, v_x, v_y,
|v|^2/2).

torch is plumbing here: tensors in, tensors out; every computation runs in
the HIP library.
"""
from __future__ import annotations

import numpy as np

from . import FP64, Config, HydroContext, problems


class KHAdjointProblem:
    """Owns the FP64 context, the base initial state and the frozen dt sequence."""

    def __init__(self, n: int = 512, nsteps: int = 200, checkpoint_every: int = 20, riemann_solver: str = "rusanov",
                 seed: int = 20261202, cfg: Config | None = None, autotune: bool = False):
        self.cfg = cfg or problems.kh_config(n)
        self.cfg.riemann_solver = riemann_solver
        self.nsteps = nsteps
        self.every = checkpoint_every
        self.seed = seed
        self.ctx = HydroContext(self.cfg, precision_form=FP64)
        self.dts = None
        g = self.cfg.ng
        base = problems.kh_band(self.cfg, seed=seed)
        rho = base[0, 0, g:-g, g:-g]
        self.vx = base[1, 0, g:-g, g:-g] / rho
        self.vy = base[2, 0, g:-g, g:-g] / rho
        self.span_rows = 0
        if autotune:  # span height of the forward (and recomputed) stages for this grid
            self.ctx.upload(base)
            self.ctx.fill_bcs()
            self.span_rows, _ = self.ctx.autotune(steps=3)

    def initial_state(self, control: np.ndarray) -> np.ndarray:
        return problems.kh_band(self.cfg, seed=self.seed, rho_perturb=np.asarray(control, dtype=np.float64))

    def forward(self, control: np.ndarray, freeze_dt: bool = True) -> np.ndarray:
        """Final interior density. The first call records the dt sequence; later calls reuse it."""
        q0 = self.initial_state(control)
        self.ctx.upload(q0)
        self.ctx.fill_bcs()
        dts = self.ctx.adjoint_forward(self.nsteps, self.every, dts=self.dts if freeze_dt else None)
        if self.dts is None:
            self.dts = dts
        g = self.cfg.ng
        return self.ctx.download()[0, 0, g:-g, g:-g].copy()

    def backward(self, grad_rho_final: np.ndarray) -> np.ndarray:
        """d(loss)/d(control) from d(loss)/d(final interior density)."""
        g = self.cfg.ng
        lam = np.zeros(self.cfg.shape)
        lam[0, 0, g:-g, g:-g] = grad_rho_final
        l0 = self.ctx.adjoint_backward(lam)[:, 0, g:-g, g:-g]
        return l0[0] + self.vx * l0[1] + self.vy * l0[2] + 0.5 * (self.vx ** 2 + self.vy ** 2) * l0[3]


    def tangent(self, direction: np.ndarray):
        """Exact directional derivative of the final interior density along a control direction
        (forward mode over the recorded, frozen-dt steps), with the final density itself."""
        g = self.cfg.ng
        dq = np.zeros(self.cfg.shape)
        d = np.asarray(direction, dtype=np.float64)
        dq[0, 0, g:-g, g:-g] = d
        dq[1, 0, g:-g, g:-g] = self.vx * d
        dq[2, 0, g:-g, g:-g] = self.vy * d
        dq[3, 0, g:-g, g:-g] = 0.5 * (self.vx ** 2 + self.vy ** 2) * d
        dqn, qn = self.ctx.adjoint_tangent(dq)
        return dqn[0, 0, g:-g, g:-g].copy(), qn[0, 0, g:-g, g:-g].copy()


def kh_adjoint_function():
    """Build the torch.autograd.Function (imports torch lazily)."""
    import torch

    class KHAdjoint(torch.autograd.Function):
        @staticmethod
        def forward(ctx, control, problem):
            rho = problem.forward(control.detach().cpu().numpy())
            ctx.problem = problem
            ctx.dtype, ctx.device = control.dtype, control.device
            return torch.from_numpy(rho).to(control.device)

        @staticmethod
        def backward(ctx, grad_out):
            g = ctx.problem.backward(grad_out.detach().cpu().numpy().astype(np.float64))
            return torch.from_numpy(g).to(device=ctx.device, dtype=ctx.dtype), None

    return KHAdjoint


def _main(argv=None) -> int:
    """Forward + reverse sweep of the KH adjoint workload; prints the reverse sweep's device time
    (the command the adjoint kernel-trace summaries are taken with)."""
    import argparse

    ap = argparse.ArgumentParser(description=_main.__doc__)
    ap.add_argument("--n", type=int, default=512, help="grid cells per side")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--every", type=int, default=20, help="checkpoint interval (steps)")
    ap.add_argument("--solver", nargs="+", default=["rusanov", "hllc"])
    a = ap.parse_args(argv)
    for solver in a.solver:
        prob = KHAdjointProblem(n=a.n, nsteps=a.steps, checkpoint_every=a.every, riemann_solver=solver)
        c = 0.01 * (2 * problems.uniform01(17, np.arange(a.n * a.n)).reshape(a.n, a.n) - 1)
        rho = prob.forward(c)
        prob.backward(rho / np.linalg.norm(rho))
        print(f"{solver}: {a.n}^2, {a.steps} steps, checkpoint every {a.every}: "
              f"reverse sweep {prob.ctx.adjoint_backward_ms():.3f} ms", flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(_main())
