"""GPU: the Chebyshev bookkeeping of the device step against the free-particle closed form.

With mu = lambda = 0 and nothing pinned every sweep maps x to the inertia target y, so the
iterates follow x_n = omega_{n+1} (y - x_{n-2}) + x_{n-2} exactly (VBD.cpp:75-90, :302-311),
with x_{-1} = x_t (VBDsolver::step) or y (solveVBD's x <- y, VBD_GPU_Wrapper.cu:334); see
tests/closed_forms.py.  The same closed form pins the oracle to 1e-12 in FP64
(tests/test_oracle.py).

Tolerance: 2e-6 of max|x| per frame (FP32 rounding of the step, the blend and the velocity
update, about 16 ulps at the largest coordinate).  Where a bookkeeping slip would be visible
(an even iteration count of at most 4, see the closed form), each slip is shown to differ from
the closed form by more than 20 times that tolerance.
"""
import numpy as np
import pytest

from closed_forms import chebyshev_free_particles, chebyshev_iterations
from vbd_tetmesh_amd import VBDParams, meshgen
from vbd_tetmesh_amd.solver import DeviceMesh, prepare_host

pytestmark = pytest.mark.gpu

H = 1.0 / 60.0
G = (0.0, -9.8, 0.0)
TOL = 2e-6
SLIPS = ("pp_after_blend", "omega_n", "blend_vs_prev")


def f32(a):
    return np.float64(np.float32(a))


def free_particles(seed=0):
    m = meshgen.beam(6, 2, 2)
    m.fixed[:] = 0
    rng = np.random.default_rng(seed)
    v0 = rng.normal(scale=2.0, size=m.x0.shape).astype(np.float32)
    x0 = (m.x0 + rng.normal(scale=0.01, size=m.x0.shape)).astype(np.float32)
    t, off, verts, dminv, vol = prepare_host(m)
    return m, x0, v0, dict(positions=x0, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                           rest_volume=vol, color_off=off, color_verts=verts, velocity=v0)


def check_slips(ref, make_bad, iterations):
    if iterations % 2 == 1 or iterations > 4:
        return
    scale = np.abs(ref).max()
    for slip in SLIPS:
        gap = np.abs(make_bad(slip) - ref).max() / scale
        assert gap > 20 * TOL, (slip, gap)


@pytest.mark.parametrize("lanes", [0, 8])
@pytest.mark.parametrize("start_from_inertia", [False, True])
@pytest.mark.parametrize("rho", [0.75, 0.95])
@pytest.mark.parametrize("iterations", [1, 2, 3, 4, 5, 20])
def test_step_follows_the_chebyshev_closed_form(iterations, rho, start_from_inertia, lanes):
    m, x0, v0, kw = free_particles()
    print(f"[chebyshev step] iters={iterations} rho={rho} x<-y={start_from_inertia} lanes={lanes}", flush=True)
    d = DeviceMesh().init(**kw)
    d.set_params(VBDParams(dt=H, numIterations=iterations, gravity=G, useAcceleration=True, accelerationRho=rho,
                           mu=0.0, lambda_=0.0, startFromInertia=start_from_inertia))
    if lanes:
        d.set_launch_config(lanes, 256, 1)
    got = []
    for _ in range(3):
        d.step(1)
        got.append(d.positions())
    got = np.stack(got)
    args = (x0, v0, [f32(c) for c in G], f32(H), iterations, f32(rho))
    ref = chebyshev_free_particles(*args, start_from_inertia=start_from_inertia, frames=3)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  vs FP64 closed form {err:.2e}", flush=True)
    assert err <= TOL
    if not start_from_inertia:
        check_slips(ref, lambda s: chebyshev_free_particles(*args, frames=3, slip=s), iterations)


@pytest.mark.parametrize("start_from_inertia", [False, True])
@pytest.mark.parametrize("iterations", [2, 4, 5, 20])
def test_solve_vbd_follows_the_chebyshev_closed_form(iterations, start_from_inertia):
    """VBD_GPU_Solver::solveVBD (vbd_solve) with the caller's inertia target: the iterations
    alone, from x_t or from y."""
    m, x0, v0, kw = free_particles(1)
    rng = np.random.default_rng(2)
    y = (x0 + rng.normal(scale=0.05, size=x0.shape)).astype(np.float32)
    print(f"[chebyshev solveVBD] iters={iterations} x<-y={start_from_inertia}", flush=True)
    d = DeviceMesh().init(**kw)
    d.set_inertia(y)
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    d.solve(iterations, sdt, 0.0, 0.0, True, 0.95, start_from_inertia)
    d.synchronize()
    got = d.positions()
    start = y if start_from_inertia else x0
    ref = chebyshev_iterations(start, y, iterations, f32(0.95))
    err = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"  vs FP64 closed form {err:.2e}", flush=True)
    assert err <= TOL
    if not start_from_inertia:
        check_slips(ref, lambda s: chebyshev_iterations(start, y, iterations, f32(0.95), slip=s), iterations)
