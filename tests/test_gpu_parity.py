"""GPU parity: the gfx950 path through the C ABI against the CPU oracle.

Tolerance (stated per test): positions per frame within 1e-5 relative max-norm,
max|x_gpu - x_oracle| / max|x_oracle|, against the FP32 oracle (VBD.cpp semantics).
Integer/index work (adjacency, color lists) must match bit-exactly.
"""
import numpy as np
import pytest

from oracle import oracle as O
from vbd_tetmesh_amd import VBDParams, meshgen
from vbd_tetmesh_amd.solver import DeviceMesh, TetMesh, VBD_GPU_Solver, prepare_host, upload

pytestmark = pytest.mark.gpu

H = 1.0 / 60.0
TOL = 1e-5


def relmax(a, b):
    return float(np.abs(np.asarray(a, np.float64) - b).max() / max(np.abs(b).max(), 1e-30))


def params_for(m, iterations, accel=True, rho=0.95, contact=None, start_from_inertia=False):
    return VBDParams(dt=H, numIterations=iterations, substeps=1, gravity=m.gravity, useAcceleration=accel,
                     accelerationRho=rho, mu=m.mu, lambda_=m.lam, contact=contact,
                     startFromInertia=start_from_inertia)


def oracle_params(m, iterations, accel=True, rho=0.95, contact=None, mode=O.MODE_CLOSED, init_from_inertia=False):
    return O.make_params(dt=H, iterations=iterations, gravity=m.gravity, accel=accel, rho=rho, mu=m.mu, lam=m.lam,
                         contact=contact, mode=mode, init_from_inertia=init_from_inertia)


def run_gpu_frames(m, params, frames, lanes=None, block=256, graph=True, unroll=1):
    d = upload(m, params=params)
    if lanes is not None:
        d.set_launch_config(lanes, block, unroll)
    d.set_graph(graph)
    out = []
    for _ in range(frames):
        d.step(1)
        out.append(d.positions())
    return np.stack(out), d


@pytest.mark.parametrize("gen", ["beam", "cube16", "sphere"])
def test_device_built_adjacency_and_colors_match_host(gen):
    m = {"beam": lambda: meshgen.beam(), "cube16": lambda: meshgen.noisy_cube(16, seed=1),
         "sphere": lambda: meshgen.voxel_sphere(20, seed=3)}[gen]()
    print(f"[upload {gen}] nV={m.num_verts} nT={m.num_tets}", flush=True)
    t, off, verts, dminv, vol = prepare_host(m)
    d = upload(m)
    aoff, aent = d.adjacency()
    assert np.array_equal(aoff, t.adjOffsets)
    assert np.array_equal(aent >> 2, t.adjTets)
    corner = aent & 3
    assert np.array_equal(m.tets[aent >> 2, corner], np.repeat(np.arange(m.num_verts), np.diff(aoff)))
    coff, cverts = d.color_lists()
    assert np.array_equal(coff, off) and np.array_equal(cverts, verts)
    info = d.info()
    assert info.num_free == int((m.fixed == 0).sum())
    assert info.num_pairs == int(np.diff(t.adjOffsets)[m.fixed == 0].sum())
    # round trip of the vertex state through the slot permutation
    assert np.array_equal(d.positions(), m.x0)


def test_device_matches_the_committed_oracle_fixtures():
    """tests/golden/oracle_small.npz: the beam after one frame (1e-5 rel) and the device's color
    lists of the beam and the 16^3 cube (exact)."""
    import os
    z = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz"))
    m = meshgen.beam()
    got, d = run_gpu_frames(m, params_for(m, 100), 1)
    rel = relmax(got[0], z["beam_x1"].astype(np.float64))
    print(f"beam frame 1 vs fixture {rel:.3e}", flush=True)
    assert rel <= TOL
    for mesh, key in ((m, "beam_color_of"), (meshgen.noisy_cube(16, seed=1), "cube16_color_of")):
        coff, cverts = upload(mesh).color_lists()
        color_of = np.empty(mesh.num_verts, np.int32)
        for c in range(len(coff) - 1):
            color_of[cverts[coff[c]:coff[c + 1]]] = c
        assert np.array_equal(color_of, z[key]), key


def test_one_sweep_matches_oracle_vertex_solves():
    """One VBD_GPU_Solver::solveIteration vs the oracle's per-vertex solves, color by color."""
    m = meshgen.beam()
    prep = O.prepare(m)
    t, off, verts, dminv, vol = prepare_host(m)
    y = m.x0.copy()
    y[m.fixed == 0] += np.float32(H * H) * np.asarray(m.gravity, np.float32)
    g = VBD_GPU_Solver()
    t.mInertia[:] = y
    g.initFromCPU(t, (off, verts), dminv, vol)
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    g.solveIteration(sdt, m.mu, m.lam)
    got = g.device_mesh.positions()
    x = m.x0.copy()
    p = oracle_params(m, 1, accel=False)
    for c in range(prep.n_colors):
        vs = prep.color_verts[prep.color_off[c]:prep.color_off[c + 1]]
        vs = vs[m.fixed[vs] == 0]
        xo, _ = O.solve_subset(prep, p, x, y, vs)
        x[vs] = xo
    disp = np.abs(x - m.x0).max()
    assert disp > 0
    err = np.abs(got.astype(np.float64) - x).max()
    print(f"one sweep: max |dx| {disp:.3e}, gpu-oracle {err:.3e}", flush=True)
    # x + dx is rounded to float: allow one ulp of the largest coordinate on top of 1e-4 |dx|
    ulp = np.spacing(np.float32(np.abs(x).max()))
    assert err <= 1e-4 * disp + ulp


def test_beam_60_frames_matches_fp32_oracle():
    """Cantilever beam 20x4x4 (525 verts), 100 iterations, Chebyshev 0.95, 60 frames."""
    m = meshgen.beam()
    prep = O.prepare(m)
    _, _, of32, _ = O.simulate(prep, oracle_params(m, 100), m.x0, None, 60)
    _, _, of64, _ = O.simulate(prep, oracle_params(m, 100), m.x0, None, 60, f64=True, rest=m.rest,
                               density=m.density)
    got, _ = run_gpu_frames(m, params_for(m, 100), 60)
    rel = [relmax(got[f], of32[f]) for f in range(60)]
    rel64 = [relmax(got[f], of64[f]) for f in range(60)]
    o3264 = [relmax(of32[f], of64[f]) for f in range(60)]
    print(f"beam: max rel vs fp32 oracle {max(rel):.3e}, gpu vs fp64 {max(rel64):.3e}, "
          f"fp32 oracle vs fp64 {max(o3264):.3e}", flush=True)
    assert max(rel) <= TOL


def test_beam_matches_faithful_oracle():
    """Against the Energy.cpp-faithful force path (9x9 directional derivatives), 5 frames."""
    m = meshgen.beam()
    prep = O.prepare(m)
    _, _, of, _ = O.simulate(prep, oracle_params(m, 100, mode=O.MODE_FAITHFUL), m.x0, None, 5)
    got, _ = run_gpu_frames(m, params_for(m, 100), 5)
    rel = max(relmax(got[f], of[f]) for f in range(5))
    print(f"beam faithful: {rel:.3e}", flush=True)
    assert rel <= TOL


def test_main_cpp_hybrid_loop_with_gpu_solver():
    """The reference main.cpp:215-232 pattern: host forwardStep, updateInertiaFromCPU,
    solveVBD (x <- y start), copyPositionsToCPU, host updateVelocity."""
    m = meshgen.beam()
    prep = O.prepare(m)
    frames = 20
    _, _, of, _ = O.simulate(prep, oracle_params(m, 30, rho=0.95, init_from_inertia=True), m.x0, None, frames)
    t, off, verts, dminv, vol = prepare_host(m)
    g = VBD_GPU_Solver()
    t.computeInertia(H)
    g.initFromCPU(t, (off, verts), dminv, vol)
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    rel = 0.0
    for f in range(frames):
        t.savePreviousPosition()
        t.applyGravity(m.gravity, H)
        t.computeInertia(H)
        g.updateInertiaFromCPU(t)
        g.solveVBD(30, sdt, m.mu, m.lam, True, 0.95)
        g.copyPositionsToCPU(t)
        t.updateVelocity(H)
        rel = max(rel, relmax(t.mVertPos, of[f]))
    print(f"hybrid loop: {rel:.3e}", flush=True)
    assert rel <= TOL


def test_caller_stream_gives_identical_results():
    """vbd_set_stream: the same frames on a torch stream (graphs replayed there) and back on the
    handle's own stream."""
    import torch
    m = meshgen.beam(12, 3, 3)
    a, _ = run_gpu_frames(m, params_for(m, 20), 4)
    d = upload(m, params=params_for(m, 20))
    s = torch.cuda.Stream()
    d.set_stream(s)
    for _ in range(3):
        d.step(1)
    d.set_stream(None)
    d.step(1)
    assert np.array_equal(d.positions(), a[-1])


def test_graph_replay_is_bitwise_identical_to_eager():
    m = meshgen.beam(12, 3, 3)
    a, _ = run_gpu_frames(m, params_for(m, 20), 6, graph=True)
    b, _ = run_gpu_frames(m, params_for(m, 20), 6, graph=False)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("lanes,block,unroll", [(4, 64, 1), (4, 128, 3), (8, 128, 1), (8, 256, 2), (8, 192, 3),
                                                (16, 256, 1), (16, 64, 2), (32, 256, 1), (64, 256, 1),
                                                (0, 64, 1), (0, 128, 2), (0, 256, 2), (0, 512, 1)])
def test_launch_configurations_agree(lanes, block, unroll):
    m = meshgen.noisy_cube(12, seed=3)
    prep = O.prepare(m)
    _, _, of, _ = O.simulate(prep, oracle_params(m, 20), m.x0, None, 3)
    got, _ = run_gpu_frames(m, params_for(m, 20), 3, lanes=lanes, block=block, unroll=unroll)
    rel = max(relmax(got[f], of[f]) for f in range(3))
    print(f"cube12 lanes={lanes} block={block} unroll={unroll}: {rel:.3e}", flush=True)
    assert rel <= TOL


def test_unsupported_launch_configuration_is_rejected():
    from vbd_tetmesh_amd import _capi as A
    d = upload(meshgen.beam(3, 1, 1))
    for bad in [(3, 64, 1), (32, 256, 2), (8, 96, 1), (16, 512, 1), (8, 64, 4), (0, 96, 1), (0, 256, 3),
                (0, 1024, 1), (-1, 256, 1)]:
        with pytest.raises(A.VBDError):
            d.set_launch_config(*bad)


def fan_mesh(n):
    """n tets around the segment (0,0,-0.3)-(0,0,1): both segment ends have valence n."""
    ang = np.linspace(0, 2 * np.pi, n, endpoint=False)
    ring = np.stack([np.cos(ang), np.sin(ang), np.zeros(n)], 1)
    rest = np.concatenate([[[0, 0, 1.0], [0, 0, -0.3]], ring]).astype(np.float32)
    i = np.arange(n)
    tets = np.stack([np.zeros(n), 2 + i, 2 + (i + 1) % n, np.ones(n)], 1).astype(np.int32)
    fixed = np.zeros(n + 2, np.uint8)
    fixed[2] = 1
    return meshgen.Mesh(name="fan", rest=rest, x0=rest.copy(), tets=tets, fixed=fixed)


def test_tiled_kernel_rejects_tiles_smaller_than_max_valence():
    from vbd_tetmesh_amd import _capi as A
    m = fan_mesh(100)
    d = upload(m, params=params_for(m, 10))
    assert d.info().max_valence == 100
    with pytest.raises(A.VBDError) as ei:
        d.set_launch_config(0, 64, 1)
    assert ei.value.code == A.VBD_ERR_ARG
    d.set_launch_config(0, 128, 1)  # 128 records per tile hold the valence-100 vertices
    d.step(2)
    prep = O.prepare(m)
    _, _, of, _ = O.simulate(prep, oracle_params(m, 10), m.x0, None, 2)
    assert relmax(d.positions(), of[-1]) <= TOL


def test_tiled_configurations_are_bitwise_identical():
    """Tiled kernel: the per-vertex summation order depends only on the vertex's valence, so every
    (records per thread, block) choice gives the same bits; and it agrees with the oracle."""
    m = meshgen.voxel_sphere(20, seed=5)
    contact = (2, -1.05, 1e4)
    prep = O.prepare(m)
    _, _, of, _ = O.simulate(prep, oracle_params(m, 20, contact=contact), m.x0, None, 6)
    outs = []
    for lanes, block, k in [(0, 64, 1), (0, 256, 1), (0, 128, 2), (0, 512, 2), (0, 64, 2)]:
        got, _ = run_gpu_frames(m, params_for(m, 20, contact=contact), 6, lanes=lanes, block=block, unroll=k)
        outs.append(got)
    for o in outs[1:]:
        assert np.array_equal(o, outs[0])
    rel = max(relmax(outs[0][f], of[f]) for f in range(6))
    print(f"sphere20 tiled: {rel:.3e}", flush=True)
    assert rel <= TOL


@pytest.mark.parametrize("gen", ["beam13", "cube20", "cube64"])
def test_autotune_restores_the_state_and_caches_its_choice(gen):
    """vbd_autotune times candidate sweeps on the live mesh: afterwards the positions and
    velocities are the ones before the call, the chosen mapping continues the trajectory (bitwise
    for a tiled choice, whose bits do not depend on the configuration; within the parity
    tolerance otherwise), and a second mesh with the same valence key reuses the choice without
    timing."""
    # the small meshes pick a per-vertex mapping; the 64^3 cube picked another tiled
    # configuration (a new tile cut) when this was written: both continue the trajectory
    m = meshgen.beam(13, 3, 3) if gen == "beam13" else meshgen.noisy_cube(int(gen[4:]), seed=7)
    p = params_for(m, 30)
    a = upload(m, params=p)
    a.step(2)
    x, v = a.positions(), a.velocities()
    before = (a.info().lanes_per_vertex, a.info().pairs_per_lane, a.info().block_size)
    lanes, unroll, block, ms = a.autotune(2)
    print(f"autotune {gen}: {before} -> lanes={lanes} pairs/lane={unroll} block={block} {ms:.4f} ms/sweep "
          f"({'current cut kept' if (lanes, unroll, block) == before else 'new configuration'})", flush=True)
    assert ms > 0
    assert np.array_equal(a.positions(), x) and np.array_equal(a.velocities(), v)
    ref = upload(m, params=p)
    ref.step(2)
    assert np.array_equal(ref.positions(), x)
    a.step(3)
    ref.step(3)
    if lanes == 0:
        assert np.array_equal(a.positions(), ref.positions())
    else:
        assert relmax(a.positions(), ref.positions()) <= TOL
    c = upload(m, params=p)
    assert c.autotune(2) == (lanes, unroll, block, -1.0)


def test_noisy_cube_32_matches_oracle():
    """The 128^3 workload's setup (noise, pins, Chebyshev) at 32^3, 20 iterations, 3 frames."""
    m = meshgen.noisy_cube(32, seed=0)
    prep = O.prepare(m)
    p = O.make_params(dt=H, iterations=20, rho=0.95, threads=8)
    _, _, of, _ = O.simulate(prep, p, m.x0, None, 3)
    _, _, of64, _ = O.simulate(prep, p, m.x0, None, 3, f64=True, rest=m.rest, density=m.density)
    got, _ = run_gpu_frames(m, params_for(m, 20), 3)
    rel = max(relmax(got[f], of[f]) for f in range(3))
    rel64 = max(relmax(got[f], of64[f]) for f in range(3))
    o64 = max(relmax(of[f], of64[f]) for f in range(3))
    print(f"cube32: {rel:.3e}, gpu vs fp64 {rel64:.3e}, fp32 oracle vs fp64 {o64:.3e}", flush=True)
    assert rel <= TOL and rel64 <= TOL


@pytest.mark.parametrize("material", [0, 1])
def test_cube128_color_launches_match_oracle_on_sampled_vertices(material):
    """Full-size 128^3 cube (2.15M verts, 12.6M tets): one device sweep, every color launch
    checked on a seeded sample of 2000 of its vertices against the oracle's local solves from
    the state that launch saw (earlier colors already updated by the device, later colors
    and this color still at the start positions).  Material 0: stable Neo-Hookean (compact
    records); 1: the reference GPU kernel's corotated model (full records)."""
    m = meshgen.noisy_cube(128, seed=0)
    print(f"[cube128] material {material} nV={m.num_verts} nT={m.num_tets}", flush=True)
    prep = O.prepare(m)
    t, off, verts, dminv, vol = prepare_host(m)
    y = m.x0.copy()
    y[m.fixed == 0] += np.float32(H * H) * np.asarray(m.gravity, np.float32)
    d = DeviceMesh().init(t.mVertPos, t.vertexMass, t.isFixed, t.tets, dminv, vol, off, verts, inertia=y)
    d.set_material_model(material)
    info = d.info()
    assert info.num_verts == 129 ** 3 and info.num_tets == 6 * 128 ** 3 and info.lanes_per_vertex == 0
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    d.solve_iteration(sdt, m.mu, m.lam)
    got = d.positions()
    rng = np.random.default_rng(0)
    p = oracle_params(m, 1, accel=False)
    p.material = material
    worst = 0.0
    for c in range(prep.n_colors):
        vs = prep.color_verts[prep.color_off[c]:prep.color_off[c + 1]]
        vs = vs[m.fixed[vs] == 0]
        sample = np.sort(rng.choice(vs, size=min(2000, vs.size), replace=False)).astype(np.int32)
        xin = got.copy()
        later = prep.color_verts[prep.color_off[c]:]
        xin[later] = m.x0[later]
        xo, _ = O.solve_subset(prep, p, xin, y, sample)
        disp = np.abs(xo - m.x0[sample]).max()
        err = np.abs(got[sample].astype(np.float64) - xo).max()
        worst = max(worst, err / max(disp, 1e-30))
        print(f"cube128 material {material} color {c}: |dx| {disp:.3e} err {err:.3e}", flush=True)
    assert worst <= 1e-4


def test_sphere96_contact_color_launches_match_oracle_on_sampled_vertices():
    """Full-size 96^3 voxel sphere (485k verts, 2.78M tets, 11 colors, shuffled ids) with the
    penalty plane cutting its bottom cap, so contact is active in the first sweep: one device
    sweep, each color launch checked on its penetrating vertices plus a seeded sample (2000 in
    all) against the oracle's local solves from the state that launch saw."""
    m = meshgen.voxel_sphere(96, seed=0)
    contact = (2, -0.95, 1e4)
    print(f"[sphere96] nV={m.num_verts} nT={m.num_tets}", flush=True)
    prep = O.prepare(m)
    t, off, verts, dminv, vol = prepare_host(m)
    y = m.x0.copy()
    y[m.fixed == 0] += np.float32(H * H) * np.asarray(m.gravity, np.float32)
    d = DeviceMesh().init(t.mVertPos, t.vertexMass, t.isFixed, t.tets, dminv, vol, off, verts, inertia=y)
    d.set_params(params_for(m, 1, contact=contact))
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    d.solve_iteration(sdt, m.mu, m.lam)
    got = d.positions()
    rng = np.random.default_rng(1)
    p = oracle_params(m, 1, accel=False, contact=contact)
    worst, touching = 0.0, 0
    for c in range(prep.n_colors):
        vs = prep.color_verts[prep.color_off[c]:prep.color_off[c + 1]]
        vs = vs[m.fixed[vs] == 0]
        low = vs[m.x0[vs, 2] < contact[1]]
        rest = np.setdiff1d(vs, low)
        k = max(0, min(2000 - low.size, rest.size))
        sample = np.sort(np.concatenate([low, rng.choice(rest, size=k, replace=False)])).astype(np.int32)
        touching += low.size
        xin = got.copy()
        later = prep.color_verts[prep.color_off[c]:]
        xin[later] = m.x0[later]
        xo, _ = O.solve_subset(prep, p, xin, y, sample)
        disp = np.abs(xo - m.x0[sample]).max()
        err = np.abs(got[sample].astype(np.float64) - xo).max()
        worst = max(worst, err / max(disp, 1e-30))
        print(f"sphere96 color {c}: {sample.size} checked ({low.size} in contact), |dx| {disp:.3e} err {err:.3e}",
              flush=True)
    assert touching > 0 and worst <= 1e-4


def test_sphere_with_contact_matches_oracle():
    """Voxel sphere (shuffled ids, irregular valence) dropped on a penalty plane, 24^3 grid."""
    m = meshgen.voxel_sphere(24, seed=0)
    contact = (2, -1.2, 1e4)
    prep = O.prepare(m)
    frames = 20  # contact starts near frame 12
    _, _, of, _ = O.simulate(prep, oracle_params(m, 30, contact=contact), m.x0, None, frames)
    _, _, of64, _ = O.simulate(prep, oracle_params(m, 30, contact=contact), m.x0, None, frames, f64=True, rest=m.rest,
                               density=m.density)
    got, _ = run_gpu_frames(m, params_for(m, 30, contact=contact), frames)
    rel = max(relmax(got[f], of[f]) for f in range(frames))
    rel64 = max(relmax(got[f], of64[f]) for f in range(frames))
    o64 = max(relmax(of[f], of64[f]) for f in range(frames))
    assert of[-1][:, 2].min() < -1.19  # contact is active by the last frame
    print(f"sphere24 contact: {rel:.3e}, gpu vs fp64 {rel64:.3e}, fp32 oracle vs fp64 {o64:.3e}", flush=True)
    assert rel <= TOL and rel64 <= TOL


@pytest.mark.parametrize("gen", ["sphere", "cube"])
def test_slot_order_does_not_change_results(gen):
    """Auto, ascending-id and Morton vertex order inside a color give bitwise identical frames:
    same-color vertices share no tet, so the update of one never reads another."""
    m = meshgen.voxel_sphere(20, seed=4) if gen == "sphere" else meshgen.noisy_cube(12, seed=2)
    contact = (2, -1.1, 1e4) if gen == "sphere" else None
    out = []
    for order in (0, 1, 2):
        d = upload(m, params=params_for(m, 25, contact=contact), slot_order=order)
        d.step(15)
        out.append(d.positions())
    assert np.array_equal(out[0], out[1]) and np.array_equal(out[0], out[2])


def test_invalid_slot_order_is_rejected():
    from vbd_tetmesh_amd import _capi as A
    with pytest.raises(A.VBDError) as ei:
        upload(meshgen.beam(3, 1, 1), slot_order=3)
    assert ei.value.code == A.VBD_ERR_ARG


def test_uniform_per_tet_material_equals_scalar_bitwise():
    m = meshgen.beam(10, 3, 3)
    a, _ = run_gpu_frames(m, params_for(m, 15), 3)
    m2 = meshgen.beam(10, 3, 3)
    m2.tet_mu = np.full(m2.num_tets, m2.mu, np.float32)
    m2.tet_lambda = np.full(m2.num_tets, m2.lam, np.float32)
    b, _ = run_gpu_frames(m2, params_for(m2, 15), 3)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("cfg", [(8, 256, 1), (0, 256, 2), (0, 128, 1)])
def test_batch_equals_per_mesh_runs_bitwise(cfg):
    ms = meshgen.batch_beams(12, seed=5)
    params = VBDParams(dt=H, numIterations=20, useAcceleration=True, accelerationRho=0.95, gravity=(0, -9.8, 0))
    descs = []
    singles = []
    for m in ms:
        t, off, verts, dminv, vol = prepare_host(m)
        kw = dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                  rest_volume=vol, color_off=off, color_verts=verts, tet_mu=t.tetMu, tet_lambda=t.tetLambda)
        descs.append(kw)
        d = DeviceMesh().init(**kw)
        d.set_params(params)
        d.set_launch_config(*cfg)
        d.step(4)
        singles.append(d.positions())
    b = DeviceMesh().init_batch(descs)
    b.set_params(params)
    b.set_launch_config(*cfg)
    b.step(4)
    got = b.positions()
    ref = np.concatenate(singles)
    assert np.array_equal(got, ref)
    # and the batch agrees with the oracle on one of its meshes
    m = ms[3]
    prep = O.prepare(m)
    p = O.make_params(dt=H, iterations=20, rho=0.95)
    _, _, of, _ = O.simulate(prep, p, m.x0, None, 4)
    _, _, of64, _ = O.simulate(prep, p, m.x0, None, 4, f64=True, rest=m.rest, density=m.density)
    o = sum(x.num_verts for x in ms[:3])
    rel, rel64 = relmax(got[o:o + m.num_verts], of[-1]), relmax(got[o:o + m.num_verts], of64[-1])
    print(f"batch mesh 3: {rel:.3e}, gpu vs fp64 {rel64:.3e}, fp32 oracle vs fp64 {relmax(of[-1], of64[-1]):.3e}",
          flush=True)
    assert rel <= TOL and rel64 <= TOL


def test_benchmark_batch_of_256_beams_equals_per_mesh_runs_bitwise():
    """The batched benchmark workload at full size (256 uneven beams, 394,633 verts, per-tet
    materials, default launch configuration): 2 batched frames == the 256 meshes stepped alone."""
    ms = meshgen.batch_beams(256, seed=0)
    params = VBDParams(dt=H, numIterations=20, useAcceleration=True, accelerationRho=0.95, gravity=(0, -9.8, 0))
    descs, singles = [], []
    for m in ms:
        t, off, verts, dminv, vol = prepare_host(m)
        kw = dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                  rest_volume=vol, color_off=off, color_verts=verts, tet_mu=t.tetMu, tet_lambda=t.tetLambda)
        descs.append(kw)
    b = DeviceMesh().init_batch(descs)
    b.set_params(params)
    info = b.info()
    print(f"[batch256] nV={info.num_verts} lanes={info.lanes_per_vertex} block={info.block_size} "
          f"k={info.pairs_per_lane}", flush=True)
    b.step(2)
    got = b.positions()
    for kw in descs:
        d = DeviceMesh().init(**kw)
        d.set_params(params)
        d.set_launch_config(info.lanes_per_vertex, info.block_size, info.pairs_per_lane)
        d.step(2)
        singles.append(d.positions())
        d.close()
    assert info.num_verts == 394633
    assert np.array_equal(got, np.concatenate(singles))


def test_device_energy_matches_oracle():
    m = meshgen.noisy_cube(10, seed=1)
    prep = O.prepare(m)
    d = upload(m)
    e = d.total_energy(m.mu, m.lam)
    eo = O.total_energy(prep, O.make_params(mu=m.mu, lam=m.lam), m.x0)
    assert np.isclose(e, eo, rtol=1e-4), (e, eo)


def test_energy_timing_leaves_the_energies_and_reports_the_yardstick():
    m = meshgen.noisy_cube(10, seed=1)
    d = upload(m, params=params_for(m, 5))
    e0, g0 = d.total_energy(m.mu, m.lam), d.incremental_potential()
    info = d.info()
    for mode, extra in ((0, 0), (1, 28 * info.num_free)):
        ms, nb = d.time_energy(mode, 5)
        assert ms > 0 and nb == 56 * info.num_tets + 12 * info.num_verts + extra
    assert d.total_energy(m.mu, m.lam) == e0 and d.incremental_potential() == g0
    from vbd_tetmesh_amd import _capi as A
    for bad in ((2, 5), (0, 0)):
        with pytest.raises(A.VBDError) as ei:
            d.time_energy(*bad)
        assert ei.value.code == A.VBD_ERR_ARG


def _potential_fp64(m, x, y, mass, contact=None):
    from oracle import diff_oracle as D
    P = D.make_problem(m)
    mu = np.full(m.num_tets, m.mu) if m.tet_mu is None else m.tet_mu.astype(np.float64)
    lam = np.full(m.num_tets, m.lam) if m.tet_lambda is None else m.tet_lambda.astype(np.float64)
    el = D.energy(P, x.astype(np.float64), mu, lam)
    fr = m.fixed == 0
    d = (x.astype(np.float64) - y)[fr]
    ine = 0.5 * float(np.sum(mass.astype(np.float64)[fr] * np.sum(d * d, axis=1))) / (H * H)
    if contact is not None:
        ax, plane, k = contact
        pen = np.maximum(plane - x.astype(np.float64)[fr, ax], 0.0)
        ine += 0.5 * k * float(np.sum(pen * pen))
    return el + ine, el


@pytest.mark.parametrize("gen", ["cube", "sphere_contact", "cube90"])
def test_incremental_potential_matches_fp64(gen):
    """G = inertia + stable NH elastic (+ penalty) on the device (double, fixed order) vs float64
    numpy with the FP64 oracle's DmInv and volumes (1e-5 relative).  cube90 (4.37M tets) has
    more than 16384 per-block partials, so its final sums take the two-stage path."""
    if gen == "cube":
        m, contact = meshgen.noisy_cube(10, seed=2), None
    elif gen == "cube90":
        m, contact = meshgen.noisy_cube(90, seed=2), None
    else:
        m, contact = meshgen.voxel_sphere(14, seed=1), (2, -0.9, 1e4)
    rng = np.random.default_rng(7)
    t, _, _, _, _ = prepare_host(m)
    x = (m.rest + rng.normal(scale=0.02, size=m.rest.shape)).astype(np.float32)
    y = (m.rest + rng.normal(scale=0.02, size=m.rest.shape)).astype(np.float32)
    d = upload(m, params=params_for(m, 10, contact=contact))
    d.set_positions(x)
    d.set_inertia(y)
    G, el = d.incremental_potential()
    G64, el64 = _potential_fp64(m, x, y, t.vertexMass, contact)
    print(f"{gen}: G {G:.9e} vs {G64:.9e}, elastic {el:.9e} vs {el64:.9e}", flush=True)
    assert abs(G - G64) <= 1e-5 * abs(G64) and abs(el - el64) <= 1e-5 * abs(el64)
    assert d.incremental_potential() == (G, el)  # fixed summation order: bit-reproducible
    e1 = d.total_energy(m.mu, m.lam)
    assert d.total_energy(m.mu, m.lam) == e1


def test_two_stage_residual_sum_matches_numpy():
    """|M g| over 35k free vertices (the two-stage fixed-order sum) vs float64 numpy."""
    m = meshgen.noisy_cube(32, seed=3)
    t, _, _, _, _ = prepare_host(m)
    d = upload(m, params=params_for(m, 10))
    assert d.info().num_free > 16384
    _, load = d.residual_norm()
    fr = m.fixed == 0
    g2 = sum(float(np.float32(c)) ** 2 for c in m.gravity)
    ref = np.sqrt(np.sum(t.vertexMass[fr].astype(np.float64) ** 2) * g2)
    print(f"|Mg| {load:.12e} vs {ref:.12e}", flush=True)
    assert abs(load - ref) <= 1e-12 * ref
    assert d.residual_norm()[1] == load


def test_sweeps_lower_the_incremental_potential():
    """Each plain Gauss-Seidel sweep (no Chebyshev) lowers G from a perturbed state."""
    m = meshgen.noisy_cube(16, seed=4)
    d = upload(m, params=params_for(m, 10, accel=False))
    y = m.x0.copy()
    y[m.fixed == 0] += np.float32(H * H) * np.asarray(m.gravity, np.float32)
    d.set_inertia(y)
    sdt = float(np.float32(1.0 / np.float64(np.float32(H) * np.float32(H))))
    gs = [d.incremental_potential()[0]]
    for _ in range(6):
        d.solve_iteration(sdt, m.mu, m.lam)
        gs.append(d.incremental_potential()[0])
    print("G per sweep:", " ".join(f"{g:.6e}" for g in gs), flush=True)
    assert all(b < a for a, b in zip(gs, gs[1:]))


def test_solve_before_init_is_a_state_error():
    from vbd_tetmesh_amd import _capi as A
    d = DeviceMesh()
    with pytest.raises(A.VBDError) as ei:
        d.solve_iteration(1.0, 1.0, 1.0)
    assert ei.value.code == A.VBD_ERR_STATE


def test_invalid_coloring_is_rejected():
    from vbd_tetmesh_amd import _capi as A
    m = meshgen.beam(3, 1, 1)
    t, off, verts, dminv, vol = prepare_host(m)
    bad_off = np.array([0, m.num_verts], np.int32)  # everything in one color
    with pytest.raises(A.VBDError) as ei:
        DeviceMesh().init(t.mVertPos, t.vertexMass, t.isFixed, t.tets, dminv, vol, bad_off,
                          np.arange(m.num_verts, dtype=np.int32))
    assert ei.value.code == A.VBD_ERR_ARG
