"""Differentiable rollouts of converged VBD steps (implicit differentiation on the device).

Forward: `steps` implicit-Euler steps, each iterated on the GPU until
|grad G| <= max(abs_tol, rel_tol |M g|) (vbd_step_converged; by default |grad G| < 1e-6) by VBD
sweeps, or (forward_method="newton", precision 64) by Newton steps with conjugate-gradient solves, from
x_0 = rest with initial velocity v_0 and per-tet materials (mu, lambda).  The reference has no
differentiation; the step itself follows VBDsolver::step (VBD.cpp:60-101, TetMesh.h:305-348).

Precision: 64 (default) runs the differentiation entry points on the handle's FP64 state and
checkpoints it in double, so every step converges below 1e-6; 32 runs the FP32 kernels on the
FP32 state, whose converged steps stall near H ulp(x) per vertex (about 5e-3 |M g| on a 16^3
unit cube) and end on the stagnation test.

Backward, per step n (x_{n+1} = argmin G, v_{n+1} = c (x_{n+1} - x_n)/h, c = 0.999; with
precision 64 the vectors cross the ABI in double, since the update of a cancels most of its terms):
    a_hat = a + (c/h) b
    A(x_{n+1}) w = a_hat           on the device: block-Jacobi preconditioned conjugate gradients
                                   with precision 64 (vbd_adjoint_solve_cg_f64), colored block
                                   Gauss-Seidel with Chebyshev otherwise (vbd_adjoint_solve)
    mu_bar  += -V (F - C) : dF(w),  lambda_bar += -V (J - 1) C : dF(w)   (vbd_parameter_gradients)
    a <- (M/h^2) w - (c/h) b,  b <- (M/h) w
and v0_bar = b at the end.  Positions and velocities are checkpointed every `checkpoint_every`
steps; each segment's states are recomputed from its checkpoint during the backward pass (the
device step is bit-deterministic, so the recomputed states equal the forward ones).
"""
from __future__ import annotations

import time
import warnings
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _capi as A
from .solver import DeviceMesh, VBDParams, prepare_host


@dataclass
class RolloutStats:
    forward_iterations: List[int] = field(default_factory=list)
    forward_residuals: List[float] = field(default_factory=list)
    adjoint_sweeps: List[int] = field(default_factory=list)
    adjoint_residuals: List[float] = field(default_factory=list)
    adjoint_fallbacks: int = 0  # solves rerun with the next method after breaking down or diverging
    unconverged_steps: int = 0  # precision 64: forward steps that ended above abs_tol
    forward_inner_iterations: List[int] = field(default_factory=list)  # Newton: CG iterations per step
    forward_seconds: float = 0.0
    backward_seconds: float = 0.0


class DiffSim:
    """Device rollout + adjoint for one mesh with per-tet materials."""

    def __init__(self, mesh, steps: int = 30, h: float = 1.0 / 60.0, rho: float = 0.95, max_iterations: int = 4000,
                 check_every: int = 10, rel_tol: float = 0.0, stall_checks: int = 10, adj_max_iterations: int = 8000,
                 adj_rel_tol: Optional[float] = None, adj_rho: float = 0.95, adj_stall_checks: int = 5,
                 checkpoint_every: int = 5, device: int = 0, damping: float = 0.999, precision: int = 64,
                 abs_tol: float = 1e-6, adj_method: str = "auto", forward_method: str = "vbd"):
        if mesh.tet_mu is None:
            raise ValueError("the differentiable rollout needs per-tet materials (mesh.tet_mu / tet_lambda)")
        t, off, verts, dminv, vol = prepare_host(mesh)
        self.dev = DeviceMesh(device).init(t.mVertPos, t.vertexMass, t.isFixed, t.tets, dminv, vol, off, verts,
                                           tet_mu=mesh.tet_mu, tet_lambda=mesh.tet_lambda)
        self.params = VBDParams(dt=h, numIterations=max_iterations, substeps=1, gravity=mesh.gravity,
                                useAcceleration=rho > 0, accelerationRho=rho, velocityDamping=damping)
        self.dev.set_params(self.params)
        self.precision = int(precision)
        self.dev.set_differentiation_precision(self.precision)
        if adj_rel_tol is None:  # the FP32 adjoint residual stalls near 1e-6
            adj_rel_tol = 1e-10 if self.precision == 64 else 1e-5
        # the step's h and damping as the device holds them (float32): the backward recursion
        # (a <- M/h^2 w - (c/h) b) cancels most of its terms, so it must use the forward's values
        self.steps, self.h, self.c = steps, float(np.float32(h)), float(np.float32(damping))
        self.max_iterations, self.check_every, self.rel_tol, self.abs_tol = max_iterations, check_every, rel_tol, abs_tol
        self.stall_checks, self.adj_stall_checks = stall_checks, adj_stall_checks
        self.adj_max_iterations, self.adj_rel_tol, self.adj_rho = adj_max_iterations, adj_rel_tol, adj_rho
        # "cg": block-Jacobi preconditioned conjugate gradients (precision 64; about a third of the
        # Chebyshev Gauss-Seidel solve's launches on the 16^3 workload); "gs": colored block
        # Gauss-Seidel with Chebyshev; "auto": cg with precision 64, else gs
        if adj_method not in ("auto", "cg", "gs"):
            raise ValueError("adj_method must be 'auto', 'cg' or 'gs'")
        if adj_method == "cg" and self.precision != 64:
            raise ValueError("the conjugate-gradient adjoint solve needs precision 64")
        self.adj_method = ("cg" if self.precision == 64 else "gs") if adj_method == "auto" else adj_method
        # converged steps by VBD sweeps ("vbd") or by Newton steps with conjugate-gradient solves
        # ("newton", precision 64: the same minimiser of G, about a third of the launches on the
        # 16^3 workload)
        if forward_method not in ("vbd", "newton"):
            raise ValueError("forward_method must be 'vbd' or 'newton'")
        if forward_method == "newton" and self.precision != 64:
            raise ValueError("the Newton forward needs precision 64")
        self.forward_method = forward_method
        self.dev.set_converged_method(A.VBD_CONVERGED_NEWTON if forward_method == "newton" else A.VBD_CONVERGED_VBD)
        self.k = max(1, checkpoint_every)
        self.x0 = t.mVertPos.copy()
        self.mass = t.vertexMass.astype(np.float64)
        self.free = t.isFixed == 0
        self.num_tets = t.numTets
        self.stats = RolloutStats()

    # -- forward
    def _step(self):
        it, res = self.dev.step_converged(self.max_iterations, self.check_every, self.abs_tol, self.rel_tol,
                                          self.stall_checks)
        return it, res

    def _state(self):
        """(positions, velocities) of the device state, in double with precision 64."""
        if self.precision == 64:
            return self.dev.state_f64()
        return self.dev.positions(), self.dev.velocities()

    def _set_state(self, x, v=None):
        if self.precision == 64:
            self.dev.set_state_f64(x, v)
        else:
            self.dev.set_positions(x)
            if v is not None:
                self.dev.set_velocities(v)

    def rollout(self, mu, lam, v0) -> Tuple[np.ndarray, Dict[int, Tuple[np.ndarray, np.ndarray]]]:
        """Returns (x_N, checkpoints {n: (x_n, v_n)})."""
        t0 = time.perf_counter()
        self.dev.set_materials(mu, lam)
        self.dev.set_positions(self.x0)
        self.dev.set_velocities(np.ascontiguousarray(v0, np.float32).reshape(-1, 3))
        ck = {}
        self.stats = RolloutStats()
        for n in range(self.steps):
            if n % self.k == 0:
                ck[n] = self._state()
            it, res = self._step()
            self.stats.forward_iterations.append(it)
            self.stats.forward_residuals.append(res)
            if self.forward_method == "newton":
                self.stats.forward_inner_iterations.append(self.dev.last_step_stats()[1])
            if self.precision == 64 and self.abs_tol > 0 and not res <= self.abs_tol:
                self.stats.unconverged_steps += 1
        if self.stats.unconverged_steps:
            # the adjoint assumes grad G = 0 at every step: its gradients are then those of the
            # states reached, not of converged steps (raise max_iterations, or lower rho)
            warnings.warn(f"{self.stats.unconverged_steps} of {self.steps} forward steps ended with |grad G| above "
                          f"{self.abs_tol:g} (largest {max(self.stats.forward_residuals):.3g})", RuntimeWarning)
        x = self._state()[0]
        self.stats.forward_seconds = time.perf_counter() - t0
        return x, ck

    def _segment_states(self, start: int, end: int, ck) -> List[np.ndarray]:
        """Positions x_start .. x_end recomputed from the checkpoint at `start`."""
        x, v = ck[start]
        self._set_state(x, v)
        xs = [x]
        for _ in range(start, end):
            self._step()
            xs.append(self._state()[0])
        return xs

    # -- backward
    def backward(self, ck, dL_dxN, dL_dvN=None):
        """Gradients (mu_bar (T,), lambda_bar (T,), v0_bar (N,3)) in float64."""
        t0 = time.perf_counter()
        h, c = self.h, self.c
        free = self.free
        m = self.mass[:, None]
        a = np.asarray(dL_dxN, np.float64).reshape(-1, 3).copy()
        b = np.zeros_like(a) if dL_dvN is None else np.asarray(dL_dvN, np.float64).reshape(-1, 3).copy()
        a[~free] = 0.0
        b[~free] = 0.0
        gmu = np.zeros(self.num_tets)
        glam = np.zeros(self.num_tets)
        starts = sorted(ck.keys())
        w = None
        for si in range(len(starts) - 1, -1, -1):
            start = starts[si]
            end = starts[si + 1] if si + 1 < len(starts) else self.steps
            xs = self._segment_states(start, end, ck)
            for n in range(end - 1, start - 1, -1):
                ahat = a + (c / h) * b
                self._set_state(xs[n + 1 - start])
                w, sweeps, rr = self._adjoint_solve(ahat, w)
                self.stats.adjoint_sweeps.append(sweeps)
                self.stats.adjoint_residuals.append(rr)
                dm, dl = (self.dev.parameter_gradients_f64(w) if self.precision == 64 else
                          self.dev.parameter_gradients(w))
                gmu += dm
                glam += dl
                w64 = w.astype(np.float64)
                a_new = m / h ** 2 * w64 - (c / h) * b
                b_new = m / h * w64
                a_new[~free] = 0.0
                b_new[~free] = 0.0
                a, b = a_new, b_new
        self.stats.backward_seconds = time.perf_counter() - t0
        return gmu, glam, b

    def _adjoint_solve(self, rhs, w0):
        """Conjugate gradients (adj_method "cg"), falling back to the Chebyshev-accelerated Gauss-
        Seidel solve if they break down; that solve falls back to plain Gauss-Seidel (rho = 0) if it
        diverges."""
        if self.adj_method == "cg":
            try:
                return self.dev.adjoint_solve_cg_f64(rhs, w0=w0, max_iterations=self.adj_max_iterations,
                                                     check_every=self.check_every, rel_tol=self.adj_rel_tol)
            except A.VBDError as e:
                if e.code != A.VBD_ERR_NUMERIC:
                    raise
                self.stats.adjoint_fallbacks += 1
        kw = dict(max_iterations=self.adj_max_iterations, check_every=self.check_every, rel_tol=self.adj_rel_tol,
                  stall_checks=self.adj_stall_checks)
        solve = self.dev.adjoint_solve_f64 if self.precision == 64 else self.dev.adjoint_solve
        if self.adj_rho > 0:
            try:
                return solve(rhs, w0=w0, rho=self.adj_rho, **kw)
            except A.VBDError as e:
                if e.code != A.VBD_ERR_NUMERIC:
                    raise
                self.stats.adjoint_fallbacks += 1
        return solve(rhs, w0=w0, rho=0.0, **kw)

    def loss_and_grads(self, mu, lam, v0, x_target):
        x, ck = self.rollout(mu, lam, v0)
        d = x.astype(np.float64) - x_target
        L = float(np.sum(d * d))
        gmu, glam, gv0 = self.backward(ck, 2.0 * d)
        return L, gmu, glam, gv0


def make_autograd_function():
    """torch.autograd.Function over DiffSim: x_N = rollout(mu, lambda, v0)."""
    import torch

    class VBDRollout(torch.autograd.Function):
        @staticmethod
        def forward(ctx, mu, lam, v0, sim: DiffSim):
            x, ck = sim.rollout(mu.detach().cpu().numpy(), lam.detach().cpu().numpy(), v0.detach().cpu().numpy())
            ctx.sim, ctx.ck = sim, ck
            ctx.dev = mu.device
            return torch.from_numpy(np.ascontiguousarray(x)).to(device=mu.device, dtype=torch.float32)

        @staticmethod
        def backward(ctx, gx):
            gmu, glam, gv0 = ctx.sim.backward(ctx.ck, gx.detach().cpu().numpy())
            dev = ctx.dev
            return (torch.from_numpy(gmu.astype(np.float32)).to(dev), torch.from_numpy(glam.astype(np.float32)).to(dev),
                    torch.from_numpy(gv0.astype(np.float32)).to(dev), None)

    return VBDRollout
