"""GPU: state changes on a live handle.

* vbd_set_materials refreshes every copy of the per-tet materials a kernel reads: after a change,
  the tiled and per-vertex kernels step bitwise like a fresh upload with the new materials, and
  a differentiable rollout with changed materials matches the FP64 oracle's forward.
* The vbd_solve / converged-step graph cache is bounded (VBD_MAX_CACHED_GRAPHS) and its replays
  equal direct launches bitwise, also when the scalars change on every call.
* The 52-byte full pair records exist only while a kernel of the configuration reads them, and
  building or releasing them leaves the results bitwise unchanged.

Tolerances: bitwise where a state change must not alter the arithmetic; the rollout within
1e-3 of the FP64 forward's largest displacement (the FP32 floor of the converged step).
"""
import numpy as np
import pytest

from oracle import diff_oracle as D
from vbd_tetmesh_amd import VBDParams, meshgen
from vbd_tetmesh_amd import _capi as A
from vbd_tetmesh_amd.solver import DeviceMesh, prepare_host

pytestmark = pytest.mark.gpu

H = 1.0 / 60.0
MAX_GRAPHS = 32  # VBD_MAX_CACHED_GRAPHS


def beam_kw(mu_scale=1.0, nx=12):
    m = meshgen.beam(nx, 3, 3)
    t, off, verts, dminv, vol = prepare_host(m)
    mu = (np.linspace(0.5, 1.5, m.num_tets) * m.mu * mu_scale).astype(np.float32)
    lam = (np.linspace(1.5, 0.5, m.num_tets) * m.lam * mu_scale).astype(np.float32)
    return m, dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                   rest_volume=vol, color_off=off, color_verts=verts, tet_mu=mu, tet_lambda=lam)


def params(iterations=15):
    return VBDParams(dt=H, numIterations=iterations, gravity=(0.0, -9.8, 0.0), useAcceleration=True,
                     accelerationRho=0.95)


@pytest.mark.parametrize("cfg", [(0, 256, 1), (0, 128, 2), (8, 256, 1)])
def test_set_materials_equals_a_fresh_upload_bitwise(cfg):
    m, kw = beam_kw()
    _, kw2 = beam_kw(mu_scale=2.0)
    print(f"[set_materials] cfg={cfg} nV={m.num_verts}", flush=True)
    d = DeviceMesh().init(**kw)
    d.set_params(params())
    d.set_launch_config(*cfg)
    d.step(1)  # captures the step with the first materials
    d.set_materials(kw2["tet_mu"], kw2["tet_lambda"])
    d.set_positions(kw["positions"])
    d.set_velocities(np.zeros_like(kw["positions"]))
    d.step(3)
    got = d.positions()
    f = DeviceMesh().init(**kw2)
    f.set_params(params())
    f.set_launch_config(*cfg)
    f.step(3)
    ref = f.positions()
    old = DeviceMesh().init(**kw)
    old.set_params(params())
    old.set_launch_config(*cfg)
    old.step(3)
    assert not np.array_equal(old.positions(), ref)  # the change matters
    assert np.array_equal(got, ref)


def test_rollout_with_changed_materials_matches_fp64_forward():
    """DiffSim.rollout calls set_materials every time: a rollout with mu, lambda scaled by 2 after
    one with the upload's materials follows the FP64 forward with the scaled materials."""
    from vbd_tetmesh_amd.adjoint import DiffSim
    m, mu_t, lam_t, v0 = meshgen.diff_workload(4, seed=0)
    P = D.make_problem(m)
    sim = DiffSim(m, steps=6)
    sim.rollout(m.tet_mu, m.tet_lambda, v0)
    mu2, lam2 = (2.0 * m.tet_mu).astype(np.float32), (2.0 * m.tet_lambda).astype(np.float32)
    x, _ = sim.rollout(mu2, lam2, v0)
    xs64 = D.forward(P, m.rest.astype(np.float64), v0.astype(np.float64), mu2.astype(np.float64),
                     lam2.astype(np.float64), 6)[0]
    x64 = xs64[-1]
    disp = np.abs(x64 - m.rest).max()
    err = np.abs(x - x64).max() / disp
    x1 = D.forward(P, m.rest.astype(np.float64), v0.astype(np.float64), m.tet_mu.astype(np.float64),
                   m.tet_lambda.astype(np.float64), 6)[0][-1]
    gap = np.abs(x1 - x64).max() / disp
    print(f"rollout with 2x materials: vs FP64 {err:.2e} of max displacement (materials change it by {gap:.2e})",
          flush=True)
    assert gap > 1e-2
    assert err <= 1e-3


def test_solve_graph_cache_is_bounded_and_replays_bitwise():
    """vbd_solve (VBD_GPU_Solver::solveVBD) with mu changed on every call: no more than
    VBD_MAX_CACHED_GRAPHS graphs are kept, a repeated call is captured on its second use, and
    the positions equal the same calls made with graphs disabled."""
    m, kw = beam_kw()
    kw = dict(kw)
    kw.pop("tet_mu")
    kw.pop("tet_lambda")
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    runs = []
    for graph in (True, False):
        d = DeviceMesh().init(**kw)
        d.set_graph(graph)
        counts = []
        for k in range(3 * MAX_GRAPHS):
            d.solve(10, sdt, float(m.mu * (1.0 + 0.01 * (k % (2 * MAX_GRAPHS)))), m.lam, True, 0.95, False)
            counts.append(d.info().cached_graphs)
        for _ in range(3):  # one key three times: captured on the second call, replayed on the third
            d.solve(10, sdt, 3.3e5, m.lam, True, 0.95, False)
            counts.append(d.info().cached_graphs)
        d.synchronize()
        runs.append(d.positions())
        print(f"[graph cache] graphs={graph}: max cached {max(counts)}, last {counts[-4:]}", flush=True)
        if graph:
            assert max(counts) <= MAX_GRAPHS
            assert counts[-2] >= 1
        else:
            assert max(counts) == 0
    assert np.array_equal(runs[0], runs[1])


def test_full_records_are_built_on_demand_and_released():
    m, kw = beam_kw()
    d = DeviceMesh().init(**kw)
    d.set_params(params())
    i0 = d.info()
    assert i0.lanes_per_vertex == 0 and i0.full_records == 0
    d.step(2)
    x_tiled = d.positions()
    full_bytes = 36 * i0.num_pairs  # pa 16 + pb 16 + pc 4 per record
    d.set_launch_config(8, 256, 1)
    i1 = d.info()
    assert i1.full_records == 1 and i1.device_bytes == i0.device_bytes + full_bytes
    d.set_launch_config(0, 256, 1)
    assert d.info().full_records == 0 and d.info().device_bytes == i0.device_bytes
    d.set_material_model(A.VBD_MATERIAL_COROTATED_GPU)
    assert d.info().full_records == 1
    d.set_material_model(A.VBD_MATERIAL_STABLE_NEO_HOOKEAN)
    assert d.info().full_records == 0 and d.info().device_bytes == i0.device_bytes
    print(f"[full records] device bytes {i0.device_bytes} -> {i1.device_bytes} with full records", flush=True)
    # the same two steps from the upload state give the same bits after the round trips
    d.set_positions(kw["positions"])
    d.set_velocities(np.zeros_like(kw["positions"]))
    d.step(2)
    assert np.array_equal(d.positions(), x_tiled)
    # the differentiation entry points build them and keep them
    d.residual_norm()
    assert d.info().full_records == 1
