"""Closed forms the tests pin the oracle and the device against (test helpers, float64).

chebyshev_free_particles: VBDsolver::step (VBD.cpp:60-101) with mu = lambda = 0 and nothing
pinned.  Each vertex's block is then H = m/h^2 I with f = m/h^2 (y - x), so every sweep maps x
to the inertia target y exactly.  With omega_k from VBD.cpp:302-311 (omega_1 = 1,
omega_2 = 2/(2 - rho^2), omega_k = 4/(4 - rho^2 omega_{k-1})) and the bookkeeping of
VBD.cpp:75-90 (x_pp <- the positions before the sweep, taken after the blend), the iterate
after iteration n (0-based) of a substep is

    x_0 = y,    x_n = omega_{n+1} (y - x_{n-2}) + x_{n-2}  (n >= 1),

with x_{-1} the start: x_t (VBDsolver::step, VBD.cpp:72-81) or y (VBD_GPU_Solver::solveVBD's
x <- y, VBD_GPU_Wrapper.cu:334).  forwardStep (VBD.cpp:103-112) and updateVelocity
(TetMesh.h:305-311, damping 0.999) wrap it.

`slip` plants one plausible mistake at a time, so that a test can show that its tolerance
separates the right bookkeeping from these:
    "pp_after_blend": x_pp taken after the blend instead of before the sweep
    "omega_n":        omega_n used where omega_{n+1} belongs
    "blend_vs_prev":  the blend against x_{n-1} (the last iterate) instead of x_{n-2}
"""
from __future__ import annotations

import numpy as np


def omegas(n: int, rho: float) -> list:
    """omega_1 .. omega_n of VBD.cpp:302-311, in float64."""
    w = []
    for k in range(1, n + 1):
        if k == 1:
            w.append(1.0)
        elif k == 2:
            w.append(2.0 / (2.0 - rho * rho))
        else:
            w.append(4.0 / (4.0 - rho * rho * w[-1]))
    return w


def chebyshev_iterations(x_start, y, iterations, rho, accel=True, slip=None):
    """The iterates of one substep with fixed inertia target y: every sweep returns y, then the
    Chebyshev blend and bookkeeping of VBD.cpp:75-90.  Returns x after `iterations` iterations."""
    y = np.asarray(y, np.float64)
    w = omegas(max(iterations, 1) + 1, rho)
    cur = np.asarray(x_start, np.float64).copy()  # x_{-1}
    pp = np.zeros_like(y)
    for n in range(iterations):
        xn = y.copy()  # the sweep
        if accel:
            # w[n] is omega_{n+1} (w[0] = omega_1); the slip takes omega_n
            om = w[n - 1] if (slip == "omega_n" and n >= 1) else w[n]
            if om > 1.0:
                base = cur if slip == "blend_vs_prev" else pp
                xn = om * (xn - base) + base
            pp = xn.copy() if slip == "pp_after_blend" else cur
        cur = xn
    return cur


def chebyshev_free_particles(x0, v0, gravity, dt, iterations, rho, accel=True, start_from_inertia=False,
                             frames=1, substeps=1, damping=0.999, slip=None):
    """Positions after each frame, (frames, N, 3) float64."""
    x = np.asarray(x0, np.float64).copy()
    v = np.asarray(v0, np.float64).copy()
    g = np.asarray(gravity, np.float64)
    h = dt / substeps
    out = []
    for _ in range(frames):
        for _ in range(substeps):
            xprev = x.copy()
            v = v + h * g
            y = x + h * v
            x = chebyshev_iterations(y if start_from_inertia else x, y, iterations, rho, accel, slip)
            v = (x - xprev) / h * damping
        out.append(x.copy())
    return np.stack(out)
