"""CPU tests of the product library: host helpers behind the C ABI are bit-identical to the
oracle, the library loads and exports every declared symbol, and misuse fails loudly."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from oracle import oracle as O
from vbd_tetmesh_amd import _capi as A
from vbd_tetmesh_amd import meshgen
from vbd_tetmesh_amd.solver import prepare_host

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    src = open(os.path.join(ROOT, "include", "vbd_capi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(vbd_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = header_functions()
    assert len(names) >= 30
    lib = C.CDLL(A.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    bound = {s[0] for s in A.SIGNATURES}
    assert set(names) == bound, set(names) ^ bound


def test_struct_layouts_match_header(tmp_path):
    """ctypes mirrors of the ABI structs have the sizes and offsets a C compiler gives."""
    import subprocess
    prog = tmp_path / "sz.c"
    prog.write_text('#include <stdio.h>\n#include <stddef.h>\n#include "vbd_capi.h"\n'
                    'int main(void){printf("%zu %zu %zu %zu %zu %zu\\n", sizeof(vbd_mesh_desc), '
                    'sizeof(vbd_step_params), sizeof(vbd_info), offsetof(vbd_mesh_desc, color_offsets), '
                    'offsetof(vbd_step_params, contact_plane), offsetof(vbd_info, device_bytes));return 0;}\n')
    exe = tmp_path / "sz"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(prog), "-o", str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    assert got == [C.sizeof(A.MeshDesc), C.sizeof(A.StepParams), C.sizeof(A.Info), A.MeshDesc.color_offsets.offset,
                   A.StepParams.contact_plane.offset, A.Info.device_bytes.offset]


@pytest.mark.parametrize("gen", ["beam", "cube", "sphere", "diff"])
def test_host_preparation_is_bit_identical_to_oracle(gen):
    m = {"beam": lambda: meshgen.beam(), "cube": lambda: meshgen.noisy_cube(10, seed=4),
         "sphere": lambda: meshgen.voxel_sphere(12, seed=2), "diff": lambda: meshgen.diff_cube(6)}[gen]()
    t, off, verts, dminv, vol = prepare_host(m)
    p = O.prepare(m)
    assert np.array_equal(t.vertexMass, p.mass)
    assert np.array_equal(t.tetRestVol, p.rest_vol)
    assert np.array_equal(t.adjOffsets, p.adj_off) and np.array_equal(t.adjTets, p.adj)
    assert np.array_equal(off, p.color_off) and np.array_equal(verts, p.color_verts)
    assert np.array_equal(dminv, p.dminv) and np.array_equal(vol, p.vol)


def test_omega_matches_oracle_bitwise():
    prev_a = prev_b = 1.0
    for k in range(1, 40):
        a = A.accelerator_omega(k, 0.95, prev_a)
        b = O.omega(k, 0.95, prev_b)
        assert a == b
        prev_a, prev_b = a, b


def test_host_helpers_reject_bad_input():
    tets = np.array([[0, 1, 2, 7]], np.int32)
    rest = np.zeros((4, 3), np.float32)
    with pytest.raises(A.VBDError):
        A.mesh_initialize(rest, tets, 1000.0)


def test_no_device_is_a_clean_error_or_works():
    """Without a GPU, vbd_create reports an error instead of crashing."""
    h = C.c_void_p()
    rc = A.lib.vbd_create(C.byref(h), 0)
    if rc == A.VBD_OK:
        A.lib.vbd_destroy(h)
    else:
        assert rc == A.VBD_ERR_HIP and A.lib.vbd_global_last_error()


def test_null_handle_calls_fail_cleanly():
    assert A.lib.vbd_step(None, 1) == A.VBD_ERR_ARG
    assert A.lib.vbd_solve_iteration(None, 1.0, 1.0, 1.0) == A.VBD_ERR_ARG
    assert A.lib.vbd_download_positions(None, None) == A.VBD_ERR_ARG


def test_batch_generator_and_union():
    ms = meshgen.batch_beams(6, seed=1)
    u, vo, to = meshgen.concat(ms)
    assert u.num_verts == vo[-1] and u.num_tets == to[-1]
    for i, m in enumerate(ms):
        assert np.array_equal(u.tets[to[i]:to[i + 1]] - vo[i], m.tets)
        assert 8 <= m.meta["nx"] <= 64 and 2 <= m.meta["ny"] <= 8
        nu = m.meta["nu"]
        assert np.isclose(m.lam, 2 * m.mu * nu / (1 - 2 * nu))


def test_sphere_generator_properties():
    s = meshgen.voxel_sphere(16, seed=0)
    cells = s.meta["cells"]
    assert s.num_tets == 6 * cells
    r = np.linalg.norm(s.rest.astype(np.float64), axis=1)
    assert r.max() <= 1.0 + np.sqrt(3) * s.meta["cell"]
    assert s.fixed.sum() == 0 and s.gravity == (0.0, 0.0, -9.8)
    # ids shuffled: tets reference every vertex
    assert np.unique(s.tets).size == s.num_verts


def test_surface_extraction_is_closed_outward_and_complete(tmp_path):
    """TetMesh.h:45-186 restated on the host: the boundary of the 20x4x4 beam has
    4 (20*4 + 4*4 + 20*4) = 704 triangles, every directed edge appears once (closed and
    consistently oriented), and the divergence theorem over the triangles gives the tet volume
    (outward normals)."""
    from vbd_tetmesh_amd import meshgen
    from vbd_tetmesh_amd.solver import mesh_from_generator
    m = meshgen.beam()
    t = mesh_from_generator(m)
    f = t.extractSurfaceFaces()
    assert f.shape == (704, 3)
    edges = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert len({tuple(e) for e in edges}) == len(edges)
    assert {tuple(e) for e in edges} == {(b, a) for a, b in edges}
    x = t.mVertPos.astype(np.float64)
    vol_surface = np.sum(np.einsum("ij,ij->i", x[f[:, 0]], np.cross(x[f[:, 1]], x[f[:, 2]]))) / 6.0
    assert np.isclose(vol_surface, t.tetRestVol.astype(np.float64).sum(), rtol=1e-5)
    path = tmp_path / "beam.obj"
    assert t.exportToObj(str(path))
    lines = path.read_text().splitlines()
    assert sum(1 for ln in lines if ln.startswith("v ")) == m.num_verts
    assert sum(1 for ln in lines if ln.startswith("f ")) == 704


def test_python_wrappers_reject_arrays_of_the_wrong_size():
    """The C ABI takes plain pointers and reads num_verts / num_tets entries from each array, so
    the Python wrappers check the sizes first: a short array is a ValueError, never read past its
    end (no GPU needed: the checks run before any ABI call)."""
    from vbd_tetmesh_amd.solver import DeviceMesh, _mesh_arrays
    m = meshgen.beam(3, 1, 1)
    t, off, verts, dminv, vol = prepare_host(m)
    good = dict(positions=t.mVertPos, mass=t.vertexMass, is_fixed=t.isFixed, tets=t.tets, dminv=dminv,
                rest_volume=vol, color_off=off, color_verts=verts)
    a = _mesh_arrays(**good)
    assert a["positions"].shape == (m.num_verts, 3) and a["tets"].shape == (m.num_tets, 4)
    nt = m.num_tets
    for key, bad in (("mass", t.vertexMass[:-1]), ("is_fixed", t.isFixed[:-2]), ("color_verts", verts[:-1]),
                     ("dminv", dminv[:-9]), ("rest_volume", vol[:-1]), ("color_off", np.zeros(0, np.int32))):
        with pytest.raises(ValueError):
            _mesh_arrays(**{**good, key: bad})
    with pytest.raises(ValueError):
        _mesh_arrays(**good, inertia=t.mVertPos[:-1])
    with pytest.raises(ValueError):
        _mesh_arrays(**good, tet_mu=np.ones(nt, np.float32))  # lambda missing
    with pytest.raises(ValueError):
        _mesh_arrays(**good, tet_mu=np.ones(nt, np.float32), tet_lambda=np.ones(nt - 1, np.float32))
    # the state and differentiation wrappers of a handle (its counts set, no device call made)
    d = object.__new__(DeviceMesh)
    d.h, d.num_verts, d.num_tets = None, m.num_verts, nt
    for call in (lambda: d.set_positions(np.zeros((m.num_verts - 1, 3))),
                 lambda: d.set_state_f64(x=np.zeros((m.num_verts + 1, 3))),
                 lambda: d.set_materials(np.ones(nt), np.ones(nt + 1)),
                 lambda: d.hessian_vector(np.zeros((2, 3))),
                 lambda: d.adjoint_solve(np.zeros((m.num_verts, 3)), w0=np.zeros((1, 3))),
                 lambda: d.adjoint_solve_f64(np.zeros((3, 3))),
                 lambda: d.parameter_gradients_f64(np.zeros((m.num_verts - 1, 3)))):
        with pytest.raises(ValueError):
            call()
