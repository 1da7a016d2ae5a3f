"""GPU differentiation: HVP, parameter gradients, adjoint solve, converged steps and whole
rollout gradients against the FP64 CPU adjoint oracle (oracle/diff_oracle.py), which is
itself pinned by central finite differences (tests/test_diff_oracle.py)."""
import os

import numpy as np
import pytest

from oracle import diff_oracle as D
from vbd_tetmesh_amd import meshgen
from vbd_tetmesh_amd.adjoint import DiffSim

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "diffcube16_fp64.npz")


def relmax(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-300))


@pytest.fixture(scope="module")
def small():
    m, mu_t, lam_t, v0 = meshgen.diff_workload(4, seed=0)
    P = D.make_problem(m)
    sim = DiffSim(m, steps=10)
    rng = np.random.default_rng(0)
    x = (m.rest + rng.normal(scale=0.01, size=m.rest.shape)).astype(np.float32)
    x[m.fixed == 1] = m.rest[m.fixed == 1]
    return m, mu_t, lam_t, v0, P, sim, x


def test_hessian_vector_product_matches_fp64(small):
    """(M/h^2 + K(x)) w on the device vs the assembled FP64 Hessian (tolerance 1e-5 rel)."""
    m, _, _, _, P, sim, x = small
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    sim.dev.set_positions(x)
    w = np.random.default_rng(1).normal(size=x.shape).astype(np.float32)
    w[m.fixed == 1] = 0
    hv = sim.dev.hessian_vector(w)
    ref = D.hessian_vector(P, x.astype(np.float64), w.astype(np.float64), m.tet_mu.astype(np.float64),
                           m.tet_lambda.astype(np.float64))
    fr = m.fixed == 0
    err = relmax(hv[fr], ref[fr])
    print(f"hvp rel {err:.2e}", flush=True)
    assert err <= 1e-5


def test_parameter_gradients_match_fp64(small):
    m, _, _, _, P, sim, x = small
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    sim.dev.set_positions(x)
    w = np.random.default_rng(2).normal(size=x.shape).astype(np.float32)
    w[m.fixed == 1] = 0
    dm, dl = sim.dev.parameter_gradients(w)
    rm, rl = D.param_grads(P, x.astype(np.float64), w.astype(np.float64), m.tet_mu.astype(np.float64),
                           m.tet_lambda.astype(np.float64))
    print(f"param grads rel {relmax(dm, rm):.2e} {relmax(dl, rl):.2e}", flush=True)
    assert relmax(dm, rm) <= 1e-5 and relmax(dl, rl) <= 1e-5


@pytest.mark.parametrize("rho", [0.0, 0.9])
def test_adjoint_solve_matches_direct_solve(small, rho):
    """Colored block Gauss-Seidel (with / without Chebyshev) vs a float64 direct solve."""
    m, _, _, _, P, sim, x = small
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    sim.dev.set_positions(x)
    rhs = np.random.default_rng(3).normal(size=x.shape).astype(np.float32)
    rhs[m.fixed == 1] = 0
    w, sweeps, rr = sim.dev.adjoint_solve(rhs, rel_tol=1e-6, rho=rho)
    mu, lam = m.tet_mu.astype(np.float64), m.tet_lambda.astype(np.float64)
    A = D.system_matrix(P, x.astype(np.float64), mu, lam)
    wd = D.solve_linear(P, A, rhs.astype(np.float64).ravel(), mu, lam).reshape(-1, 3)
    fr = m.fixed == 0
    err = relmax(w[fr], wd[fr])
    print(f"adjoint rho={rho}: {sweeps} sweeps, residual {rr:.1e}, vs direct {err:.2e}", flush=True)
    assert rr <= 1e-6 and err <= 1e-5
    assert np.all(w[m.fixed == 1] == 0)


def test_adjoint_solve_rejects_bad_rho_and_reports_divergence(small):
    """rho outside [0, 1) is an argument error; a diverging Chebyshev solve returns
    VBD_ERR_NUMERIC and leaves w unchanged; DiffSim then reruns plain Gauss-Seidel."""
    from vbd_tetmesh_amd import _capi as A
    m, _, _, _, _, sim, x = small
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    sim.dev.set_positions(x)
    rhs = np.random.default_rng(5).normal(size=x.shape).astype(np.float32)
    rhs[m.fixed == 1] = 0
    for bad in (1.0, -0.1):
        with pytest.raises(A.VBDError) as ei:
            sim.dev.adjoint_solve(rhs, rel_tol=1e-6, rho=bad)
        assert ei.value.code == A.VBD_ERR_ARG
    outcomes = []
    for rho in (0.99, 0.999, 0.9999):
        try:
            _, sweeps, rr = sim.dev.adjoint_solve(rhs, rel_tol=1e-6, rho=rho, max_iterations=4000)
            outcomes.append(("converged", rho, sweeps, rr))
            assert rr <= 1e-6 or sweeps == 4000
        except A.VBDError as e:
            assert e.code == A.VBD_ERR_NUMERIC
            outcomes.append(("diverged", rho))
    print("chebyshev rho outcomes:", outcomes, flush=True)


def test_diverging_chebyshev_falls_back_to_gauss_seidel(small):
    """A rollout whose Chebyshev rho diverges (slowly, so the stall test must not accept it)
    still returns the plain Gauss-Seidel gradients."""
    m, mu_t, lam_t, v0, _, _, _ = small
    x_t = m.rest.copy()
    ref = DiffSim(m, steps=4, adj_rho=0.0, adj_method="gs").loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
    sim = DiffSim(m, steps=4, adj_rho=0.9999, adj_method="gs")
    got = sim.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
    print(f"fallbacks {sim.stats.adjoint_fallbacks} of {len(sim.stats.adjoint_sweeps)} solves", flush=True)
    assert sim.stats.adjoint_fallbacks > 0
    for a, b in zip(got[1:], ref[1:]):
        assert np.linalg.norm(a - b) <= 1e-3 * np.linalg.norm(b)


def test_unconverged_forward_steps_are_counted_and_warned(small):
    """Precision 64: a forward step that ends above abs_tol (here: an iteration cap far too low)
    is counted in the stats and warned about, since the adjoint assumes converged steps; a
    converged rollout counts none."""
    m, _, _, v0, _, _, _ = small
    sim = DiffSim(m, steps=3, max_iterations=10)
    with pytest.warns(RuntimeWarning, match="forward steps ended"):
        sim.rollout(m.tet_mu, m.tet_lambda, v0)
    assert sim.stats.unconverged_steps == 3
    ok = DiffSim(m, steps=3)
    ok.rollout(m.tet_mu, m.tet_lambda, v0)
    assert ok.stats.unconverged_steps == 0 and max(ok.stats.forward_residuals) <= 1e-6


def test_converged_step_matches_newton(small):
    m, _, _, v0, P, sim, _ = small
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    sim.dev.set_positions(m.rest)
    sim.dev.set_velocities(v0)
    it, res = sim._step()
    x = sim.dev.positions()
    x64, v64, _ = D.solve_step(P, m.rest.astype(np.float64), v0.astype(np.float64), m.tet_mu.astype(np.float64),
                               m.tet_lambda.astype(np.float64))
    disp = np.abs(x64 - m.rest).max()
    err = np.abs(x - x64).max() / disp
    print(f"converged step: {it} iterations, |grad G| {res:.2e}, error vs Newton {err:.2e} of max displacement",
          flush=True)
    assert err <= 1e-3
    assert np.abs(sim.dev.velocities() - v64).max() <= 1e-3 * np.abs(v64).max() + 1e-6


def test_newton_converged_step_matches_vbd_and_the_fp64_newton_oracle(small):
    """vbd_step_converged with the Newton method (precision 64) reaches the same minimiser of G as
    the VBD sweeps (both to |grad G| < 1e-6) and as the FP64 oracle's Newton step; the stats
    count its Newton steps and CG iterations; precision 32 refuses the method."""
    from vbd_tetmesh_amd import _capi as A
    m, _, _, v0, P, _, _ = small
    out = {}
    for method in (A.VBD_CONVERGED_VBD, A.VBD_CONVERGED_NEWTON):
        sim = DiffSim(m, steps=1, precision=64)
        sim.dev.set_converged_method(method)
        sim.dev.set_materials(m.tet_mu, m.tet_lambda)
        sim.dev.set_positions(m.rest)
        sim.dev.set_velocities(v0)
        it, res = sim._step()
        x, _ = sim.dev.state_f64()
        out[method] = (x, it, res, sim.dev.last_step_stats())
    x64, _, _ = D.solve_step(P, m.rest.astype(np.float64), v0.astype(np.float64), m.tet_mu.astype(np.float64),
                             m.tet_lambda.astype(np.float64))
    disp = np.abs(x64 - m.rest).max()
    xv, itv, resv, sv = out[A.VBD_CONVERGED_VBD]
    xn, itn, resn, sn = out[A.VBD_CONVERGED_NEWTON]
    print(f"converged step: vbd {itv} sweeps |grad G| {resv:.2e}; newton {sn[0]} steps, {sn[1]} CG iterations, "
          f"{sn[2]} sweeps, |grad G| {resn:.2e}; newton vs vbd {np.abs(xn - xv).max() / disp:.2e}, vs the FP64 oracle "
          f"{np.abs(xn - x64).max() / disp:.2e} of max displacement", flush=True)
    assert resv < 1e-6 and resn < 1e-6
    assert sv == (0, 0, itv) and sn[0] >= 1 and sn[1] > 0 and sn[2] == 0 and itn == sn[0]
    assert np.abs(xn - xv).max() <= 1e-6 * disp and np.abs(xn - x64).max() <= 1e-5 * disp
    s32 = DiffSim(m, steps=1, precision=32)
    s32.dev.set_converged_method(A.VBD_CONVERGED_NEWTON)
    with pytest.raises(A.VBDError) as ei:
        s32._step()
    assert ei.value.code == A.VBD_ERR_STATE
    with pytest.raises(A.VBDError) as ei:
        s32.dev.set_converged_method(2)
    assert ei.value.code == A.VBD_ERR_ARG
    with pytest.raises(ValueError):
        DiffSim(m, steps=1, precision=32, forward_method="newton")


@pytest.mark.parametrize("forward", ["vbd", "newton"])
def test_checkpoint_recomputation_is_bitwise(small, forward):
    """A segment recomputed from its checkpoint reproduces the forward states exactly (both
    converged-step methods: every reduction has a fixed order)."""
    m, _, _, v0, _, _, _ = small
    sim = DiffSim(m, steps=6, checkpoint_every=3, forward_method=forward)
    x, ck = sim.rollout(m.tet_mu, m.tet_lambda, v0)
    every = DiffSim(m, steps=6, checkpoint_every=1, forward_method=forward)
    _, ck1 = every.rollout(m.tet_mu, m.tet_lambda, v0)
    xs = sim._segment_states(3, 6, ck)
    for j, n in enumerate(range(3, 6)):
        assert np.array_equal(xs[j], ck1[n][0])
    assert np.array_equal(xs[-1], x)


@pytest.mark.parametrize("precision", [32, 64])
def test_small_rollout_gradients_match_fp64_adjoint(small, precision):
    """4^3 cube, 10 converged steps: loss and all gradients vs the FP64 adjoint (relative 1e-3
    with the FP32 state, 1e-5 with the FP64 state)."""
    m, mu_t, lam_t, v0, P, _, _ = small
    x_t = D.forward(P, m.rest.astype(np.float64), v0.astype(np.float64), mu_t.astype(np.float64),
                    lam_t.astype(np.float64), 10)[0][-1]
    sim = DiffSim(m, steps=10, precision=precision)
    L, gmu, glam, gv0 = sim.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
    L64, gmu64, glam64, gv064, _ = D.loss_and_grads(P, m.rest.astype(np.float64), v0.astype(np.float64),
                                                    m.tet_mu.astype(np.float64), m.tet_lambda.astype(np.float64), 10,
                                                    x_t)
    errs = {k: float(np.linalg.norm(a - b) / np.linalg.norm(b)) for k, a, b in
            (("mu", gmu, gmu64), ("lambda", glam, glam64), ("v0", gv0, gv064))}
    tol = 1e-3 if precision == 32 else 1e-5
    print(f"small rollout, precision {precision}: loss {L:.6e} vs {L64:.6e}; gradient rel errors {errs}; "
          f"|grad G| per step max {max(sim.stats.forward_residuals):.2e}", flush=True)
    assert abs(L - L64) <= tol * L64
    assert all(e <= tol for e in errs.values())
    if precision == 64:
        assert max(sim.stats.forward_residuals) < 1e-6


def test_torch_autograd_function(small):
    import torch
    from vbd_tetmesh_amd.adjoint import make_autograd_function
    m, mu_t, lam_t, v0, P, _, _ = small
    sim = DiffSim(m, steps=4, checkpoint_every=2)
    F = make_autograd_function()
    dev = "cuda"
    mu = torch.tensor(m.tet_mu, device=dev, requires_grad=True)
    lam = torch.tensor(m.tet_lambda, device=dev, requires_grad=True)
    v = torch.tensor(v0, device=dev, requires_grad=True)
    target = torch.tensor(m.rest, device=dev)
    x = F.apply(mu, lam, v, sim)
    loss = ((x - target) ** 2).sum()
    loss.backward()
    L, gmu, glam, gv0 = DiffSim(m, steps=4, checkpoint_every=2).loss_and_grads(m.tet_mu, m.tet_lambda, v0, m.rest)
    assert abs(loss.item() - L) <= 1e-6 * L
    assert np.allclose(mu.grad.cpu().numpy(), gmu, rtol=1e-5, atol=1e-6 * np.abs(gmu).max())
    assert np.allclose(lam.grad.cpu().numpy(), glam, rtol=1e-5, atol=1e-6 * np.abs(glam).max())
    assert np.allclose(v.grad.cpu().numpy(), gv0, rtol=1e-5, atol=1e-6 * np.abs(gv0).max())


@pytest.mark.skipif(not os.path.exists(GOLDEN), reason="FP64 golden fixture not generated")
@pytest.mark.timeout(600)
@pytest.mark.parametrize("precision,forward", [(64, "vbd"), (32, "vbd"), (64, "newton")])
def test_diffcube16_gradients_match_fp64_and_finite_differences(precision, forward):
    """Differentiable workload: 16^3 cube (4913 verts, 24576 tets), per-tet mu field, lambda = 4 mu,
    30 converged steps, loss |x_30 - x_target|^2.  GPU adjoint vs the committed FP64 CPU adjoint
    and its central finite differences (step 1e-3 of the parameter, of max |v0| for v0) on 32 seeded
    parameters.  The fixture's own FD-vs-adjoint spread, the FD noise floor, is at most 1.4e-6
    relative on 31 of the 32 and 3.9e-5 on the one whose gradient is 4e-5 of its family's largest.
    Tolerance, FP64 state: every step converged to |grad G| < 1e-6, and per parameter
    |g_gpu - g_fd| / |g_fd| <= 1e-3 and |g_gpu - g_adjoint| / |g_adjoint| <= 1e-5.
    FP32 state (steps end on the stagnation test near 5e-3 |M g|): per parameter
    |g_gpu - g_ref| <= 1e-3 * max|g_ref| of that parameter family.
    Both: loss within 1e-3 and each whole gradient vector within 1e-2, relative.  The FP64 state
    runs twice: converged by VBD sweeps and by Newton steps with conjugate-gradient solves (the
    same tolerances)."""
    z = np.load(GOLDEN)
    m, mu_t, lam_t, v0 = meshgen.diff_workload(int(z["n"]), seed=0)
    sim = DiffSim(m, steps=int(z["steps"]), precision=precision, forward_method=forward)
    L, gmu, glam, gv0 = sim.loss_and_grads(m.tet_mu, m.tet_lambda, v0, z["x_target"])
    fwd, bwd = sim.stats.forward_seconds, sim.stats.backward_seconds
    ref = {"mu": z["grad_mu"].astype(np.float64), "lambda": z["grad_lambda"].astype(np.float64),
           "v0": z["grad_v0"].astype(np.float64)}
    got = {"mu": gmu, "lambda": glam, "v0": gv0}
    vec = {k: float(np.linalg.norm(got[k] - ref[k]) / np.linalg.norm(ref[k])) for k in ref}
    fd = z["fd"]
    picks = [("mu", int(i)) for i in z["idx_mu"]] + [("lambda", int(i)) for i in z["idx_lambda"]] + \
        [("v0", (int(i), int(c))) for i, c in z["idx_v0"]]
    worst_fd, worst_adj, worst_rel_fd, worst_rel_adj = 0.0, 0.0, 0.0, 0.0
    for (kind, i), f in zip(picks, fd):
        g = got[kind][i]
        r = ref[kind][i]
        scale = np.abs(ref[kind]).max()
        worst_fd = max(worst_fd, abs(g - f) / scale)
        worst_adj = max(worst_adj, abs(g - r) / scale)
        worst_rel_fd = max(worst_rel_fd, abs(g - f) / abs(f))
        worst_rel_adj = max(worst_rel_adj, abs(g - r) / abs(r))
        print(f"  {kind} {i}: gpu {g:+.6e} fd {f:+.6e} fp64 adjoint {r:+.6e}; |gpu - fd| / |fd| {abs(g - f) / abs(f):.2e}, "
              f"fixture |fd - adjoint| / |fd| {abs(f - r) / abs(f):.2e}, |fd| / family max {abs(f) / scale:.2e}",
              flush=True)
    res = np.asarray(sim.stats.forward_residuals)
    _, load = sim.dev.residual_norm()
    print(f"  converged steps: |grad G| min {res.min():.3e} median {np.median(res):.3e} max {res.max():.3e} "
          f"(|M g| = {load:.3e}; {sim.stats.forward_iterations[:3]}... iterations)", flush=True)
    ar = np.asarray(sim.stats.adjoint_residuals)
    print(f"  adjoint solves: relative residual max {ar.max():.2e} median {np.median(ar):.2e}, sweeps "
          f"{sim.stats.adjoint_sweeps[:5]}..., fallbacks {sim.stats.adjoint_fallbacks}", flush=True)
    if forward == "newton":
        inner = np.asarray(sim.stats.forward_inner_iterations)
        print(f"  newton: {np.mean(sim.stats.forward_iterations):.2f} steps and {inner.mean():.1f} CG iterations per "
              f"converged step", flush=True)
    print(f"diffcube16, precision {precision}, {forward}: loss {L:.6e} vs {float(z['loss']):.6e}; vector rel errors {vec}; "
          f"32 params: max |gpu - fd| {worst_fd:.2e}, |gpu - fp64 adjoint| {worst_adj:.2e} (of family max); "
          f"per parameter max |gpu - fd| / |fd| {worst_rel_fd:.2e}, |gpu - adjoint| / |adjoint| {worst_rel_adj:.2e}; "
          f"forward {fwd:.2f}s backward {bwd:.2f}s ratio {bwd / fwd:.2f}", flush=True)
    assert abs(L - float(z["loss"])) <= 1e-3 * float(z["loss"])
    assert all(e <= 1e-2 for e in vec.values())
    if precision == 64:
        assert res.max() < 1e-6
        assert worst_rel_fd <= 1e-3 and worst_rel_adj <= 1e-5
    else:
        assert worst_fd <= 1e-3 and worst_adj <= 1e-3


def test_fp64_state_round_trips_and_rounds_into_the_fp32_state(small):
    """vbd_upload_state_f64 / vbd_download_state_f64 are exact; the FP32 arrays hold the rounding
    of the FP64 state after an upload and after a converged step; an FP32 upload replaces the
    FP64 state by its promotion; precision 32 refuses the FP64 state calls."""
    from vbd_tetmesh_amd import _capi as A
    m, _, _, v0, _, _, _ = small
    sim = DiffSim(m, steps=1, precision=64)
    rng = np.random.default_rng(5)
    x = m.rest.astype(np.float64) + rng.normal(scale=1e-3, size=m.rest.shape)
    v = rng.normal(size=m.rest.shape)
    sim.dev.set_state_f64(x, v)
    gx, gv = sim.dev.state_f64()
    assert np.array_equal(gx, x) and np.array_equal(gv, v)
    assert np.array_equal(sim.dev.positions(), x.astype(np.float32))
    sim.dev.set_state_f64(None, np.zeros_like(v))  # velocities only
    assert np.array_equal(sim.dev.state_f64()[0], x)
    sim._step()
    gx, gv = sim.dev.state_f64()
    assert np.array_equal(sim.dev.positions(), gx.astype(np.float32))
    assert np.array_equal(sim.dev.velocities(), gv.astype(np.float32))
    x32 = (m.rest + np.float32(1e-3)).astype(np.float32)
    sim.dev.set_positions(x32)
    assert np.array_equal(sim.dev.state_f64()[0], x32.astype(np.float64))
    sim.dev.set_differentiation_precision(32)
    with pytest.raises(A.VBDError) as ei:
        sim.dev.state_f64()
    assert ei.value.code == A.VBD_ERR_STATE


def test_fp64_kernels_match_the_fp64_oracle(small):
    """With precision 64 the HVP and the parameter gradients equal the FP64 oracle's to 1e-12
    relative, and the adjoint solve a float64 direct solve to 1e-9 (double vectors), the oracle
    given the device's inputs: the FP32 DmInv, rest volumes and masses of the host preparation
    and the FP32 step h."""
    from vbd_tetmesh_amd.solver import prepare_host
    m, _, _, _, _, _, x = small
    t, _, _, dminv32, vol32 = prepare_host(m)
    P = D.make_problem(m, h=float(np.float32(1.0 / 60.0)))
    P.dminv = dminv32.astype(np.float64).reshape(-1, 3, 3)
    P.vol = vol32.astype(np.float64)
    P.mass = t.vertexMass.astype(np.float64)
    D._prepare(P)
    sim = DiffSim(m, steps=1, precision=64)
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    xd = x.astype(np.float64)
    sim.dev.set_state_f64(xd, None)
    mu, lam = m.tet_mu.astype(np.float64), m.tet_lambda.astype(np.float64)
    rng = np.random.default_rng(11)
    w = rng.normal(size=x.shape)
    w[m.fixed == 1] = 0
    fr = m.fixed == 0
    hv = sim.dev.hessian_vector_f64(w)
    e_hv = relmax(hv[fr], D.hessian_vector(P, xd, w, mu, lam)[fr])
    dm, dl = sim.dev.parameter_gradients_f64(w)
    rm, rl = D.param_grads(P, xd, w, mu, lam)
    rhs = rng.normal(size=x.shape)
    rhs[m.fixed == 1] = 0
    ws, sweeps, rr = sim.dev.adjoint_solve_f64(rhs, rel_tol=1e-12, rho=0.9)
    A_ = D.system_matrix(P, xd, mu, lam)
    wd = D.solve_linear(P, A_, rhs.ravel(), mu, lam).reshape(-1, 3)
    e_w = relmax(ws[fr], wd[fr])
    print(f"fp64 kernels: hvp {e_hv:.2e}, param grads {relmax(dm, rm):.2e} {relmax(dl, rl):.2e}, adjoint {sweeps} "
          f"sweeps residual {rr:.1e} vs direct {e_w:.2e}", flush=True)
    assert e_hv <= 1e-12 and relmax(dm, rm) <= 1e-12 and relmax(dl, rl) <= 1e-12
    assert rr <= 1e-12 and e_w <= 1e-9


def test_conjugate_gradient_adjoint_matches_direct_solve_and_gauss_seidel(small):
    """vbd_adjoint_solve_cg_f64 (block-Jacobi preconditioned CG on the FP64 state) against a
    float64 direct solve of the oracle's system on the device's inputs (FP32 DmInv, volumes,
    masses, h): solution within 1e-9 and recomputed residual within 1e-12; fewer iterations than
    the Chebyshev Gauss-Seidel solve has sweeps, which reaches the same w; a zero right-hand side
    returns the start unchanged; precision 32 is a state error."""
    from vbd_tetmesh_amd import _capi as A
    from vbd_tetmesh_amd.solver import prepare_host
    m, _, _, _, _, _, x = small
    t, _, _, dminv32, vol32 = prepare_host(m)
    P = D.make_problem(m, h=float(np.float32(1.0 / 60.0)))
    P.dminv = dminv32.astype(np.float64).reshape(-1, 3, 3)
    P.vol = vol32.astype(np.float64)
    P.mass = t.vertexMass.astype(np.float64)
    D._prepare(P)
    sim = DiffSim(m, steps=1, precision=64)
    sim.dev.set_materials(m.tet_mu, m.tet_lambda)
    xd = x.astype(np.float64)
    sim.dev.set_state_f64(xd, None)
    mu, lam = m.tet_mu.astype(np.float64), m.tet_lambda.astype(np.float64)
    rng = np.random.default_rng(12)
    rhs = rng.normal(size=x.shape)
    rhs[m.fixed == 1] = 0
    fr = m.fixed == 0
    wc, its, rr = sim.dev.adjoint_solve_cg_f64(rhs, rel_tol=1e-12, check_every=5)
    A_ = D.system_matrix(P, xd, mu, lam)
    wd = D.solve_linear(P, A_, rhs.ravel(), mu, lam).reshape(-1, 3)
    e_w = relmax(wc[fr], wd[fr])
    wg, sweeps, rg = sim.dev.adjoint_solve_f64(rhs, rel_tol=1e-12, rho=0.95)
    print(f"cg: {its} iterations, residual {rr:.1e}, vs direct {e_w:.2e}; chebyshev gs: {sweeps} sweeps, "
          f"cg vs gs {relmax(wc[fr], wg[fr]):.2e}", flush=True)
    assert rr <= 1e-12 and e_w <= 1e-9 and relmax(wc[fr], wg[fr]) <= 1e-9
    assert its < sweeps and its % 5 == 0
    assert np.all(wc[m.fixed == 1] == 0)
    # a start that already meets the tolerance is returned without an iteration
    w2, its2, rr2 = sim.dev.adjoint_solve_cg_f64(rhs, w0=wc, rel_tol=1e-10, check_every=5)
    assert its2 == 0 and rr2 <= 1e-10 and np.array_equal(w2, wc)
    start = rng.normal(size=x.shape)
    w0, its0, rr0 = sim.dev.adjoint_solve_cg_f64(np.zeros_like(rhs), w0=start)
    assert its0 == 0 and rr0 == 0.0
    s32 = DiffSim(m, steps=1, precision=32)
    with pytest.raises(A.VBDError) as ei:
        s32.dev.adjoint_solve_cg_f64(rhs)
    assert ei.value.code == A.VBD_ERR_STATE
    # with the penalty plane cutting the state (its k e e^T enters A's diagonal blocks and A p)
    # both solves still reach the same w
    import dataclasses
    plane = float(np.quantile(xd[:, 1], 0.3))
    sim.dev.set_params(dataclasses.replace(sim.params, contact=(1, plane, 1e5)))
    sim.dev.set_state_f64(xd, None)
    assert (xd[fr][:, 1] < plane).sum() > 0
    wc2, its2c, rr2c = sim.dev.adjoint_solve_cg_f64(rhs, rel_tol=1e-12, check_every=5)
    wg2, _, _ = sim.dev.adjoint_solve_f64(rhs, rel_tol=1e-12, rho=0.95)
    print(f"cg with contact: {its2c} iterations, residual {rr2c:.1e}, vs gs {relmax(wc2[fr], wg2[fr]):.2e}, "
          f"vs the contact-free w {relmax(wc2[fr], wc[fr]):.2e}", flush=True)
    assert rr2c <= 1e-12 and relmax(wc2[fr], wg2[fr]) <= 1e-9 and relmax(wc2[fr], wc[fr]) > 1e-6


def test_rollout_gradients_conjugate_gradient_equal_gauss_seidel(small):
    """The 4^3 rollout's gradients with the CG adjoint (the precision 64 default) equal those of
    the Chebyshev Gauss-Seidel adjoint to 1e-8 of their size (both solve to 1e-10)."""
    m, _, _, v0, _, _, _ = small
    x_t = m.rest.copy()
    cg = DiffSim(m, steps=6, precision=64)
    gs = DiffSim(m, steps=6, precision=64, adj_method="gs")
    assert cg.adj_method == "cg" and gs.adj_method == "gs"
    a = cg.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
    b = gs.loss_and_grads(m.tet_mu, m.tet_lambda, v0, x_t)
    print(f"cg {np.mean(cg.stats.adjoint_sweeps):.1f} iterations per solve, gs {np.mean(gs.stats.adjoint_sweeps):.1f} "
          f"sweeps; fallbacks {cg.stats.adjoint_fallbacks}", flush=True)
    assert a[0] == b[0] and cg.stats.adjoint_fallbacks == 0
    for ga, gb in zip(a[1:], b[1:]):
        assert np.linalg.norm(ga - gb) <= 1e-8 * np.linalg.norm(gb)
    with pytest.raises(ValueError):
        DiffSim(m, steps=1, precision=32, adj_method="cg")
