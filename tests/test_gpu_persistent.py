"""GPU: the persistent sweep (vbd_set_persistent) against the per-color launches.

One persistent launch per Gauss-Seidel sweep runs every color through the tiled kernel's tile
code, with a device-wide barrier between colors, then the Chebyshev blend.  The order of the
arithmetic inside a tile, and the tile each vertex belongs to, are those of the per-color
launches, so every frame must be bitwise equal to theirs: per-tet and uniform materials, the
contact plane, acceleration on and off, both pairs-per-lane, several block sizes, the batched
mesh, graph capture on and off, the energy log, vbd_solve, and grids capped so that each block
loops over several tiles of a color.

Tolerance: bitwise (np.array_equal) for positions and velocities.
"""
import numpy as np
import pytest

from vbd_tetmesh_amd import VBDParams, meshgen
from vbd_tetmesh_amd import _capi as A
from vbd_tetmesh_amd.solver import DeviceMesh, prepare_host, upload

pytestmark = pytest.mark.gpu

H = 1.0 / 60.0


def beam_kw(nx=14, ny=3, nz=3):
    m = meshgen.beam(nx, ny, nz)
    t, off, verts, dminv, vol = prepare_host(m)
    mu = (np.linspace(0.5, 1.5, m.num_tets) * m.mu).astype(np.float32)
    lam = (np.linspace(1.5, 0.5, m.num_tets) * m.lam).astype(np.float32)
    return m, dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                   rest_volume=vol, color_off=off, color_verts=verts, tet_mu=mu, tet_lambda=lam)


def params(iterations=12, accel=True, contact=None, gravity=(0.0, -9.8, 0.0), mu=1e5, lam=4e5):
    kw = dict(dt=H, numIterations=iterations, gravity=gravity, useAcceleration=accel, accelerationRho=0.95,
              mu=mu, lambda_=lam)
    if contact is not None:
        kw["contact"] = contact
    return VBDParams(**kw)


def frames_of(make, cfg, persistent, frames=4, graph=True):
    d = make()
    d.set_launch_config(*cfg)
    d.set_graph(graph)
    d.set_persistent(persistent)
    xs, vs = [], []
    for _ in range(frames):
        d.step(1)
        xs.append(d.positions())
        vs.append(d.velocities())
    d.synchronize()
    return xs, vs


def assert_frames_equal(a, b):
    for f, (xa, xb) in enumerate(zip(a[0], b[0])):
        assert np.all(np.isfinite(xa)), f"frame {f}"
        assert np.array_equal(xa, xb), f"positions differ at frame {f}: max {np.abs(xa - xb).max()}"
    for f, (va, vb) in enumerate(zip(a[1], b[1])):
        assert np.array_equal(va, vb), f"velocities differ at frame {f}"


@pytest.mark.parametrize("cfg", [(0, 256, 1), (0, 128, 2), (0, 64, 1), (0, 512, 2), (0, 512, 1)])
@pytest.mark.parametrize("accel", [True, False])
def test_beam_per_tet_materials_bitwise(cfg, accel):
    m, kw = beam_kw()
    print(f"[persistent beam] cfg={cfg} accel={accel} nV={m.num_verts}", flush=True)

    def make():
        d = DeviceMesh().init(**kw)
        d.set_params(params(accel=accel))
        return d

    ref = frames_of(make, cfg, 0)
    got = frames_of(make, cfg, 1)
    assert_frames_equal(got, ref)
    moved = np.abs(ref[0][-1] - kw["positions"]).max()
    assert moved > 1e-4  # the frames do something


@pytest.mark.parametrize("graph", [True, False])
def test_sphere_contact_uniform_materials_bitwise(graph):
    """Voxel sphere 24^3 resting on the plane z = -0.98: the contact branch active from frame 1."""
    m = meshgen.voxel_sphere(24, seed=0)
    p = params(iterations=10, contact=(2, -0.98, 1e4), gravity=m.gravity, mu=m.mu, lam=m.lam)
    print(f"[persistent sphere] graph={graph} nV={m.num_verts}", flush=True)

    def make():
        return upload(m, params=p)

    ref = frames_of(make, (0, 256, 1), 0, graph=graph)
    got = frames_of(make, (0, 256, 1), 1, graph=graph)
    assert_frames_equal(got, ref)
    assert ref[0][-1][:, 2].min() < -0.98  # in contact


def test_batched_beams_bitwise():
    ms = meshgen.batch_beams(24, seed=1)
    descs = []
    for mm in ms:
        t, off, verts, dminv, vol = prepare_host(mm)
        descs.append(dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                          rest_volume=vol, color_off=off, color_verts=verts, tet_mu=t.tetMu, tet_lambda=t.tetLambda))
    print(f"[persistent batch] meshes={len(ms)} nV={sum(mm.num_verts for mm in ms)}", flush=True)

    def make():
        b = DeviceMesh().init_batch(descs)
        b.set_params(params(iterations=20))
        return b

    for cfg in ((0, 256, 1), (0, 128, 2)):
        assert_frames_equal(frames_of(make, cfg, 1, frames=3), frames_of(make, cfg, 0, frames=3))


@pytest.mark.parametrize("cap", [8, 16, 64])
def test_capped_grid_loops_over_tiles_bitwise(cap):
    """Noisy cube 40^3 with at most `cap` blocks: every block solves many tiles of each color."""
    m = meshgen.noisy_cube(40, seed=0)
    p = params(iterations=6, gravity=m.gravity, mu=m.mu, lam=m.lam)
    print(f"[persistent cap] cap={cap} nV={m.num_verts}", flush=True)

    def make():
        return upload(m, params=p)

    ref = frames_of(make, (0, 256, 2), 0, frames=2)
    got = frames_of(make, (0, 256, 2), cap, frames=2)
    assert_frames_equal(got, ref)


def test_energy_log_and_solve_bitwise():
    """The energy log between sweeps, and vbd_solve (solveVBD), as the per-color launches."""
    m, kw = beam_kw(nx=10)
    out = {}
    for mode in (0, 1):
        d = DeviceMesh().init(**kw)
        d.set_params(params(iterations=10))
        d.set_launch_config(0, 128, 1)
        d.set_persistent(mode)
        d.set_energy_log(3)
        d.step(2)
        log = d.energy_log()
        x1 = d.positions()
        d.forward_step()
        d.solve(8, 1.0 / (H * H), 1e5, 4e5, accel=True, rho=0.9)
        out[mode] = (log, x1, d.positions())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


def test_set_persistent_arguments_and_other_configs():
    """Mode 0, 1 or a multiple of 8; the per-vertex kernels and the corotated material keep their
    per-color launches (bitwise as with the mode off)."""
    m, kw = beam_kw(nx=8)
    d = DeviceMesh().init(**kw)
    for bad in (-1, 2, 12):
        with pytest.raises(A.VBDError):
            d.set_persistent(bad)
    for cfg, mat in (((8, 256, 1), A.VBD_MATERIAL_STABLE_NEO_HOOKEAN), ((0, 256, 1), A.VBD_MATERIAL_COROTATED_GPU)):
        res = []
        for mode in (0, 1):
            e = DeviceMesh().init(**kw)
            e.set_params(params(iterations=8))
            e.set_material_model(mat)
            e.set_launch_config(*cfg)
            e.set_persistent(mode)
            e.step(2)
            res.append(e.positions())
        assert np.array_equal(res[0], res[1])


def test_auto_grid_on_a_chip_filling_workload_bitwise():
    """Noisy cube 56^3 (185,193 vertices) at the library's default tiled configuration: more
    tiles per color than the automatic grid (2 blocks per CU), so mode 1 runs on that grid."""
    m = meshgen.noisy_cube(56, seed=0)
    p = params(iterations=10, gravity=m.gravity, mu=m.mu, lam=m.lam)
    d = upload(m, params=p)
    i = d.info()
    cfg = (i.lanes_per_vertex, i.block_size, i.pairs_per_lane)
    print(f"[persistent auto] cfg={cfg} nV={m.num_verts}", flush=True)
    d.set_persistent(1)
    grid = d.info().persistent_grid
    assert grid >= 8 and grid % 8 == 0
    ref = frames_of(lambda: upload(m, params=p), cfg, 0, frames=2)
    got = frames_of(lambda: upload(m, params=p), cfg, 1, frames=2)
    assert_frames_equal(got, ref)
