"""FP64 CPU forward + adjoint of the implicit-Euler VBD time step -- TEST INFRASTRUCTURE ONLY.

The reference has no differentiation (the adjoint is a build-defined extension); this
module states the mathematics the device adjoint must reproduce, in float64:

  step n:  v' = v_n + h g,  y = x_n + h v'                         (TetMesh.h:319-341)
           x_{n+1} = argmin_x  G(x) = sum_i m_i/(2h^2) |x_i - y_i|^2 + E(x; mu, lambda)
           v_{n+1} = c (x_{n+1} - x_n) / h, c = 0.999             (TetMesh.h:305-311)
  pinned vertices keep x and v.
  E = sum_t V_t [mu_t/2 (|F|^2 - 3) - mu_t (J - 1) + lambda_t/2 (J - 1)^2]  (its gradient is
      Energy.cpp:93-97), F = Ds DmInv.

x_{n+1} is found by Newton's method on G (conjugate gradients preconditioned by one sparse
LU of the rest-state system, backtracking line search) to |grad G| <= tol |M g|.  The
adjoint of a loss L(x_N) runs backwards:
  a_hat = a + (c/h) b;  (M/h^2 + K(x_{n+1})) w = a_hat   (free DOFs)
  mu_bar_t  -= V_t (F - C) : dF(w),   lambda_bar_t -= V_t (J - 1) C : dF(w),
               dF(w) = sum_k w_k q_k^T (q_k = shape vectors of the tet's corners)
  a <- (M/h^2) w - (c/h) b,  b <- (M/h) w
  v0_bar = b after the first step.
Pinned by central finite differences of the same forward (tests/test_diff_oracle.py).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

_EPS = np.zeros((3, 3, 3))
for _i, _j, _k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    _EPS[_i, _j, _k] = 1.0
    _EPS[_i, _k, _j] = -1.0


@dataclass
class DiffProblem:
    rest: np.ndarray      # (N,3) float64
    tets: np.ndarray      # (T,4) int64
    fixed: np.ndarray     # (N,) bool
    mass: np.ndarray      # (N,) float64
    dminv: np.ndarray     # (T,3,3) float64
    vol: np.ndarray       # (T,) float64
    h: float
    gravity: np.ndarray   # (3,)
    damping: float = 0.999
    cache: dict = field(default_factory=dict)

    @property
    def n(self):
        return self.rest.shape[0]


def make_problem(mesh, h=1.0 / 60.0) -> DiffProblem:
    rest = mesh.rest.astype(np.float64)
    tets = mesh.tets.astype(np.int64)
    Dm = np.transpose(rest[tets[:, 1:]] - rest[tets[:, [0]]], (0, 2, 1))
    dminv = np.linalg.inv(Dm)
    vol = np.abs(np.linalg.det(Dm)) / 6.0
    mass = np.zeros(rest.shape[0])
    np.add.at(mass, tets.ravel(), np.repeat(vol * mesh.density / 4.0, 4))
    P = DiffProblem(rest, tets, mesh.fixed.astype(bool), mass, dminv, vol, h, np.asarray(mesh.gravity, np.float64))
    _prepare(P)
    return P


def _prepare(P: DiffProblem):
    T = P.tets.shape[0]
    Q = np.empty((T, 4, 3))
    Q[:, 1:, :] = P.dminv
    Q[:, 0, :] = -P.dminv.sum(axis=1)
    # dvec(F)/dx: Gm[t, 3c + j, 3n + c'] = delta(c, c') Q[t, n, j]
    Gm = np.zeros((T, 9, 12))
    for c in range(3):
        for j in range(3):
            Gm[:, 3 * c + j, c::3] = Q[:, :, j]
    dof = (3 * P.tets[:, :, None] + np.arange(3)[None, None, :]).reshape(T, 12)
    rows = np.repeat(dof, 12, axis=1).ravel()
    cols = np.tile(dof, (1, 12)).ravel()
    n3 = 3 * P.n
    key = rows * n3 + cols
    ukey, inv = np.unique(key, return_inverse=True)
    urow, ucol = ukey // n3, ukey % n3
    indptr = np.zeros(n3 + 1, np.int64)
    np.add.at(indptr, urow + 1, 1)
    indptr = np.cumsum(indptr)
    pinned = np.repeat(P.fixed, 3)
    P.cache.update(Q=Q, Gm=Gm, inv=inv, nnz=ukey.size, indptr=indptr, indices=ucol, urow=urow,
                   dir_off=pinned[urow] | pinned[ucol], dir_diag=(urow == ucol) & pinned[urow], pinned=pinned)


def kinematics(P: DiffProblem, x):
    Q = P.cache["Q"]
    F = np.einsum("tnc,tnj->tcj", x[P.tets], Q)
    C = np.empty_like(F)
    C[:, 0, 0] = F[:, 1, 1] * F[:, 2, 2] - F[:, 1, 2] * F[:, 2, 1]
    C[:, 0, 1] = F[:, 1, 2] * F[:, 2, 0] - F[:, 1, 0] * F[:, 2, 2]
    C[:, 0, 2] = F[:, 1, 0] * F[:, 2, 1] - F[:, 1, 1] * F[:, 2, 0]
    C[:, 1, 0] = F[:, 2, 1] * F[:, 0, 2] - F[:, 2, 2] * F[:, 0, 1]
    C[:, 1, 1] = F[:, 2, 2] * F[:, 0, 0] - F[:, 2, 0] * F[:, 0, 2]
    C[:, 1, 2] = F[:, 2, 0] * F[:, 0, 1] - F[:, 2, 1] * F[:, 0, 0]
    C[:, 2, 0] = F[:, 0, 1] * F[:, 1, 2] - F[:, 0, 2] * F[:, 1, 1]
    C[:, 2, 1] = F[:, 0, 2] * F[:, 1, 0] - F[:, 0, 0] * F[:, 1, 2]
    C[:, 2, 2] = F[:, 0, 0] * F[:, 1, 1] - F[:, 0, 1] * F[:, 1, 0]
    J = np.einsum("tj,tj->t", F[:, 0, :], C[:, 0, :])
    return Q, F, C, J


def energy(P: DiffProblem, x, mu, lam):
    _, F, C, J = kinematics(P, x)
    return float(np.sum(P.vol * (0.5 * mu * (np.einsum("tij,tij->t", F, F) - 3) - mu * (J - 1) + 0.5 * lam * (J - 1) ** 2)))


def gradient(P: DiffProblem, x, mu, lam):
    Q, F, C, J = kinematics(P, x)
    Pk = mu[:, None, None] * (F - C) + (lam * (J - 1))[:, None, None] * C
    g_e = P.vol[:, None, None] * np.einsum("tcj,tnj->tnc", Pk, Q)
    g = np.zeros_like(x)
    np.add.at(g, P.tets.ravel(), g_e.reshape(-1, 3))
    return g


def element_hessians(P: DiffProblem, x, mu, lam):
    """(T,12,12) exact element Hessians V G^T HF G, HF = d^2 Psi / dF^2 (9x9)."""
    _, F, C, J = kinematics(P, x)
    T = F.shape[0]
    HJ = np.einsum("ikn,jlq,tnq->tijkl", _EPS, _EPS, F, optimize=True).reshape(T, 9, 9)
    c = C.reshape(T, 9)
    HF = (mu[:, None, None] * np.eye(9)[None] + lam[:, None, None] * c[:, :, None] * c[:, None, :]
          + (lam * (J - 1) - mu)[:, None, None] * HJ)
    Gm = P.cache["Gm"]
    return P.vol[:, None, None] * (np.transpose(Gm, (0, 2, 1)) @ HF @ Gm)


def system_matrix(P: DiffProblem, x, mu, lam):
    """M/h^2 + K(x) with pinned DOFs turned into identity rows/columns (CSR, SPD)."""
    ca = P.cache
    Ke = element_hessians(P, x, mu, lam)
    data = np.bincount(ca["inv"], weights=Ke.ravel(), minlength=ca["nnz"])
    data[ca["urow"] == ca["indices"]] += np.repeat(P.mass, 3)[ca["urow"][ca["urow"] == ca["indices"]]] / P.h ** 2
    data[ca["dir_off"]] = 0.0
    data[ca["dir_diag"]] = 1.0
    n3 = 3 * P.n
    return sp.csr_matrix((data, ca["indices"], ca["indptr"]), shape=(n3, n3))


def _precond(P: DiffProblem, mu, lam):
    """LU of the rest-state system with the first material seen; reused as the CG
    preconditioner for every later system of this problem (it only affects CG speed)."""
    if "pc" not in P.cache:
        A0 = system_matrix(P, P.rest, mu, lam).tocsc()
        lu = spla.splu(A0)
        P.cache["pc"] = spla.LinearOperator(A0.shape, matvec=lu.solve)
    return P.cache["pc"]


def solve_linear(P: DiffProblem, A, b, mu, lam, rtol=1e-13):
    M = _precond(P, mu, lam)
    x, info = spla.cg(A, b, rtol=rtol, atol=0.0, maxiter=500, M=M)
    if info != 0:
        x = spla.spsolve(A.tocsc(), b)
    return x


def solve_step(P: DiffProblem, x_n, v_n, mu, lam, tol=1e-10, max_newton=50, guess=None):
    """One implicit step by Newton on G; returns (x_{n+1}, v_{n+1}, newton iterations).
    Stops at |grad G| <= tol |M g| (the float64 rounding floor of grad G is near 1e-12 |M g|
    on these meshes) or when Newton stops making progress below 1e-8 |M g|."""
    h = P.h
    free = ~P.fixed
    pinned3 = P.cache["pinned"]
    vp = v_n.copy()
    vp[free] += h * P.gravity
    y = x_n.copy()
    y[free] = x_n[free] + h * vp[free]
    scale = np.linalg.norm((P.mass[free, None] * P.gravity[None, :]).ravel()) + 1e-300
    x = x_n.copy() if guess is None else guess.copy()

    def G(z):
        return 0.5 / h ** 2 * np.sum(P.mass[:, None] * (z - y) ** 2) + energy(P, z, mu, lam)

    it = 0
    prev = np.inf
    for it in range(1, max_newton + 1):
        r = (P.mass[:, None] / h ** 2 * (x - y) + gradient(P, x, mu, lam)).ravel()
        r[pinned3] = 0.0
        rn = np.linalg.norm(r)
        if rn <= tol * scale or (rn <= 1e-8 * scale and rn > 0.5 * prev):
            break
        prev = rn
        A = system_matrix(P, x, mu, lam)
        dx = solve_linear(P, A, -r, mu, lam)
        if rn <= 1e-4 * scale:
            # local quadratic regime: full Newton step (an energy line search here would
            # compare G values that differ below float64 resolution)
            x = (x.ravel() + dx).reshape(-1, 3)
            continue
        g0 = G(x)
        slope = float(r @ dx)
        alpha = 1.0
        while True:
            xt = (x.ravel() + alpha * dx).reshape(-1, 3)
            if slope >= 0 or alpha < 1e-8 or G(xt) <= g0 + 1e-4 * alpha * slope:
                break
            alpha *= 0.5
        x = xt
    v = v_n.copy()
    v[free] = (x[free] - x_n[free]) / h * P.damping
    return x, v, it


def forward(P: DiffProblem, x0, v0, mu, lam, steps, tol=1e-10, guesses=None):
    """guesses: optional list of Newton starting points per step (e.g. a nearby trajectory)."""
    xs, vs = [x0.copy()], [v0.copy()]
    for n in range(steps):
        x, v, _ = solve_step(P, xs[-1], vs[-1], mu, lam, tol, guess=None if guesses is None else guesses[n + 1])
        xs.append(x)
        vs.append(v)
    return xs, vs


def param_grads(P: DiffProblem, x, w, mu, lam):
    """(mu_bar, lambda_bar) contributions -(d^2E/dx dtheta)^T w at x."""
    Q, F, C, J = kinematics(P, x)
    dF = np.einsum("tnc,tnj->tcj", w[P.tets], Q)
    gmu = -P.vol * np.einsum("tij,tij->t", F - C, dF)
    glam = -P.vol * (J - 1) * np.einsum("tij,tij->t", C, dF)
    return gmu, glam


def hessian_vector(P: DiffProblem, x, w, mu, lam):
    """(M/h^2 + K(x)) w on all DOFs (no pinning), for checking the device HVP."""
    Ke = element_hessians(P, x, mu, lam)
    we = w[P.tets].reshape(-1, 12)
    out = np.zeros_like(w)
    np.add.at(out, P.tets.ravel(), np.einsum("tij,tj->ti", Ke, we).reshape(-1, 3))
    return out + P.mass[:, None] / P.h ** 2 * w


def adjoint(P: DiffProblem, xs: List[np.ndarray], mu, lam, dL_dxN, dL_dvN: Optional[np.ndarray] = None):
    """Gradients of L(x_N, v_N) w.r.t. per-tet mu, lambda and v_0."""
    h, c = P.h, P.damping
    free = ~P.fixed
    pinned3 = P.cache["pinned"]
    a = dL_dxN.copy()
    b = np.zeros_like(a) if dL_dvN is None else dL_dvN.copy()
    gmu = np.zeros(P.tets.shape[0])
    glam = np.zeros(P.tets.shape[0])
    for n in range(len(xs) - 2, -1, -1):
        ahat = a.copy()
        ahat[free] += (c / h) * b[free]
        rhs = ahat.ravel()
        rhs[pinned3] = 0.0
        A = system_matrix(P, xs[n + 1], mu, lam)
        w = solve_linear(P, A, rhs, mu, lam).reshape(-1, 3)
        w[P.fixed] = 0.0
        dm, dl = param_grads(P, xs[n + 1], w, mu, lam)
        gmu += dm
        glam += dl
        a_new = np.zeros_like(a)
        b_new = np.zeros_like(b)
        a_new[free] = P.mass[free, None] / h ** 2 * w[free] - (c / h) * b[free]
        b_new[free] = P.mass[free, None] / h * w[free]
        a, b = a_new, b_new
    return gmu, glam, b


def loss_and_grads(P: DiffProblem, x0, v0, mu, lam, steps, x_target, tol=1e-10):
    xs, vs = forward(P, x0, v0, mu, lam, steps, tol)
    d = xs[-1] - x_target
    L = float(np.sum(d * d))
    gmu, glam, gv0 = adjoint(P, xs, mu, lam, 2.0 * d)
    return L, gmu, glam, gv0, xs
