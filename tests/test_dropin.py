"""The source-compatible C++ drop-in (include/vbd/*.h) and the TetGen loader path."""
import os
import subprocess

import numpy as np
import pytest

from oracle import oracle as O
from vbd_tetmesh_amd import meshgen

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DEMO = os.path.join(ROOT, "vbd_tetmesh_amd", "vbd_dropin_demo")


def test_dropin_headers_compile_cleanly(tmp_path):
    src = tmp_path / "tu.cpp"
    src.write_text('#include "vbd/VBD.h"\n#include "vbd/VBD_GPU_Wrapper.h"\n#include "vbd/GraphColoring.h"\n'
                   '#include "vbd/TetGenIO.h"\nint main() { VBD::VBDParams p; (void)p; return 0; }\n')
    r = subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.parametrize("one_based", [True, False])
@pytest.mark.parametrize("chunk", [7, 65536])
def test_tetgen_round_trip(tmp_path, one_based, chunk):
    """write_tetgen in chunks of `chunk` rows (several, with a remainder, at 7) reads back
    bitwise: noisy float32 coordinates, tets, one header and one row per node / tet."""
    m = meshgen.noisy_cube(4, seed=3)
    prefix = str(tmp_path / "cube")
    meshgen.write_tetgen(prefix, m.x0, m.tets, one_based=one_based, chunk=chunk)
    pos, tets = meshgen.read_tetgen(prefix)
    assert np.array_equal(pos, m.x0) and np.array_equal(tets, m.tets)
    with open(prefix + ".node") as f:
        assert len(f.read().splitlines()) == 1 + m.num_verts
    with open(prefix + ".ele") as f:
        assert len(f.read().splitlines()) == 1 + m.num_tets


def _build_tetgen_check(tmp_path):
    exe = str(tmp_path / "tetgen_io_check")
    lib = os.path.join(ROOT, "vbd_tetmesh_amd")
    r = subprocess.run(["g++", "-O1", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "tetgen_io_check.cpp"), "-L", lib, "-lvbd_hip",
                        f"-Wl,-rpath,{lib}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _read_check(path):
    b = open(path, "rb").read()
    nv, nt = np.frombuffer(b, np.int32, 2)
    o = 8
    rest = np.frombuffer(b, np.float32, 3 * nv, o).reshape(-1, 3); o += 12 * nv
    x = np.frombuffer(b, np.float32, 3 * nv, o).reshape(-1, 3); o += 12 * nv
    tets = np.frombuffer(b, np.int32, 4 * nt, o).reshape(-1, 4); o += 16 * nt
    fixed = np.frombuffer(b, np.uint8, nv, o); o += nv
    mass = np.frombuffer(b, np.float32, nv, o); o += 4 * nv
    nf = int(np.frombuffer(b, np.int32, 1, o)[0]); o += 4
    faces = np.frombuffer(b, np.int32, 3 * nf, o).reshape(-1, 3)
    return rest, x, tets, fixed, mass, faces


def test_cpp_tetgen_loader_parses_the_format_and_pins_like_main_cpp(tmp_path):
    """include/vbd/TetGenIO.h on hand-written files: comments, blank lines, attribute and
    boundary-marker columns, 1-based ids, a separate current-shape node file.  Positions and
    tets exactly as written (ids shifted to 0-based), vertices within 5% of the y extent of the
    current shape pinned (main.cpp:146-164), lumped masses equal to the oracle's TetMesh.h:213-275
    restatement, and the 12 outward boundary triangles of the one-cell Kuhn cube."""
    m = meshgen.beam(1, 1, 1, cell=1.0)
    rest = m.rest
    cur = rest.copy()
    cur[:, 1] += 0.5 * rest[:, 0]  # shear: the pin rule must use these y values
    node = ["# rest shape", f"{rest.shape[0]}  3  1  1", ""]
    node += [f"{i + 1}  {p[0]!r}  {p[1]!r}  {p[2]!r}  7.5  {i % 2}" for i, p in enumerate(rest.tolist())]
    (tmp_path / "m.node").write_text("\n".join(node) + "\n")
    curl = [f"{rest.shape[0]} 3 0 0"] + [f"{i + 1} {p[0]!r} {p[1]!r} {p[2]!r}" for i, p in enumerate(cur.tolist())]
    (tmp_path / "cur.node").write_text("\n".join(curl) + "\n")
    ele = ["# tets", f"{m.num_tets} 4 1"] + [f"{i + 1} {t[0] + 1} {t[1] + 1} {t[2] + 1} {t[3] + 1} 3"
                                             for i, t in enumerate(m.tets.tolist())]
    (tmp_path / "m.ele").write_text("\n".join(ele) + "\n")
    exe = _build_tetgen_check(tmp_path)
    out = str(tmp_path / "out.bin")
    r = subprocess.run([exe, str(tmp_path / "m"), str(tmp_path / "cur.node"), "1", "0.05", out],
                       capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    r_rest, r_x, r_tets, r_fixed, r_mass, faces = _read_check(out)
    assert np.array_equal(r_rest, rest) and np.array_equal(r_x, cur.astype(np.float32))
    assert np.array_equal(r_tets, m.tets)
    y = cur[:, 1].astype(np.float32)
    thr = np.float32(y.min() + (y.max() - y.min()) * np.float32(0.05))
    assert np.array_equal(r_fixed, (y <= thr).astype(np.uint8)) and 0 < r_fixed.sum() < len(y)
    mass, _, _, _ = O.mesh_init(rest, m.tets, 1000.0)
    assert np.array_equal(r_mass, mass)
    # the unit cube's boundary: 6 faces x 2 triangles, each outward
    assert faces.shape == (12, 3)
    c = rest.mean(axis=0)
    for f in faces:
        a, b, d = rest[f[0]], rest[f[1]], rest[f[2]]
        assert np.dot(np.cross(b - a, d - a), (a + b + d) / 3 - c) > 0


def reference_surface_order(tets):
    """TetMesh::extractSurfaceFaces as the reference states it (TetMesh.h:80-141), brute force:
    every tet's four faces in the order (0,2,1), (0,1,3), (0,3,2), (1,2,3), keyed by the sorted
    vertex triple in a dict (the std::map), the first occurrence's winding kept; the faces that
    occur exactly once, in ascending key order (the map's iteration order)."""
    count, orient = {}, {}
    for t in np.asarray(tets).tolist():
        v0, v1, v2, v3 = t
        for tri in ((v0, v2, v1), (v0, v1, v3), (v0, v3, v2), (v1, v2, v3)):
            key = tuple(sorted(tri))
            if key not in count:
                count[key] = 1
                orient[key] = tri
            else:
                count[key] += 1
    return np.array([orient[k] for k in sorted(count) if count[k] == 1], np.int32).reshape(-1, 3)


def reference_obj_text(x, tets, faces):
    """The bytes TetMesh::exportToObj writes (TetMesh.h:147-181): float coordinates through an
    ostream's default (printf %g) formatting, 1-based face ids."""
    out = [f"# TetMesh OBJ Export\n# Vertices: {len(x)}\n# Tetrahedra: {len(tets)}\n\n"]
    out += ["v %g %g %g\n" % (float(p[0]), float(p[1]), float(p[2])) for p in x]
    out.append(f"\n# Surface Triangles: {len(faces)}\n")
    out += [f"f {a + 1} {b + 1} {c + 1}\n" for a, b, c in faces.tolist()]
    return "".join(out)


def surface_cases():
    rng = np.random.default_rng(7)
    beam = meshgen.beam(4, 2, 2)
    sph = meshgen.voxel_sphere(6, seed=1)  # shuffled vertex ids
    # tets in a shuffled order with their corners rotated by an even permutation (same
    # orientation, another first corner): another first-occurrence winding for interior faces
    t = sph.tets[rng.permutation(sph.num_tets)][:, [1, 2, 0, 3]]
    # a face shared by three tets and an isolated tet: counts 3 and 1
    x = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0, 0, -1], [0.3, 0.3, 0.8], [5, 5, 5], [6, 5, 5],
                  [5, 6, 5], [5, 5, 6]], np.float32)
    odd = np.array([[0, 1, 2, 3], [0, 2, 1, 4], [0, 1, 2, 5], [6, 7, 8, 9]], np.int32)
    return [("beam", beam.rest, beam.tets), ("sphere", sph.rest, sph.tets), ("sphere_permuted", sph.rest, t),
            ("three_tets_on_a_face", x, odd)]


@pytest.mark.parametrize("case", range(4))
def test_surface_faces_and_obj_bytes_follow_the_reference_map_order(tmp_path, case):
    """Face list (ids, winding and order) of the C++ drop-in and the Python mirror against the
    brute-force restatement of the reference's std::map; the C++ exportToObj bytes against the
    reference format; the Python exportToObj bytes equal to the C++ ones."""
    from vbd_tetmesh_amd.solver import TetMesh
    name, x, tets = surface_cases()[case]
    prefix = str(tmp_path / name)
    meshgen.write_tetgen(prefix, x, tets)
    exe = _build_tetgen_check(tmp_path)
    obj = str(tmp_path / "cpp.obj")
    r = subprocess.run([exe, prefix, "-", "-1", "0.05", str(tmp_path / "o.bin"), obj], capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    _, r_x, r_tets, _, _, faces = _read_check(str(tmp_path / "o.bin"))
    ref = reference_surface_order(tets)
    assert np.array_equal(faces, ref), name
    t = TetMesh().initialize(x, x, 1000.0, tets, 1e5, 4e5)
    assert np.array_equal(t.extractSurfaceFaces(), ref), name
    expect = reference_obj_text(r_x, tets, ref)
    assert open(obj).read() == expect
    py = str(tmp_path / "py.obj")
    assert t.exportToObj(py)
    assert open(py, "rb").read() == open(obj, "rb").read()
    print(f"[surface] {name}: {len(ref)} boundary faces in the reference order", flush=True)


def test_cpp_tetgen_loader_rejects_bad_files(tmp_path):
    exe = _build_tetgen_check(tmp_path)
    (tmp_path / "bad.node").write_text("4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n")  # one point short
    (tmp_path / "bad.ele").write_text("1 4 0\n1 1 2 3 4\n")
    r = subprocess.run([exe, str(tmp_path / "bad"), "-", "1", "0.05", str(tmp_path / "o.bin")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1
    (tmp_path / "oob.node").write_text("4 3 0 0\n1 0 0 0\n2 1 0 0\n3 0 1 0\n4 0 0 1\n")
    (tmp_path / "oob.ele").write_text("1 4 0\n1 1 2 3 5\n")  # vertex 5 does not exist
    r = subprocess.run([exe, str(tmp_path / "oob"), "-", "1", "0.05", str(tmp_path / "o.bin")], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 1


def test_cpp_host_solver_helpers_match_the_fp64_closed_form(tmp_path):
    """include/vbd/VBD.h host helpers on a deformed 8x2x2 beam: initializeColoring == the
    oracle's greedy coloring, forwardStep's inertia target == float32 y = x + h (h g), and
    solveVertex's H = m/h^2 I + sum H_kk, f = m/h^2 (y - x) - sum g_k against the oracle's FP64
    closed form on the same FP32 DmInv (1e-5 of the largest block entry, 1e-4 of the largest
    force)."""
    m = meshgen.beam(8, 2, 2)
    prefix = str(tmp_path / "b")
    meshgen.write_tetgen(prefix, m.rest, m.tets)
    rng = np.random.default_rng(3)
    cur = (m.rest + rng.normal(scale=0.004, size=m.rest.shape) * (m.fixed == 0)[:, None]).astype(np.float32)
    curl = [f"{m.num_verts} 3 0 0"] + [f"{i + 1} {p[0]!r} {p[1]!r} {p[2]!r}" for i, p in enumerate(cur.tolist())]
    (tmp_path / "cur.node").write_text("\n".join(curl) + "\n")
    exe = str(tmp_path / "host_solver_check")
    lib = os.path.join(ROOT, "vbd_tetmesh_amd")
    r = subprocess.run(["g++", "-O2", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"),
                        os.path.join(ROOT, "tests", "cpp", "host_solver_check.cpp"), "-L", lib, "-lvbd_hip",
                        f"-Wl,-rpath,{lib}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = str(tmp_path / "o.bin")
    r = subprocess.run([exe, prefix, str(tmp_path / "cur.node"), str(m.mu), str(m.lam), out], capture_output=True,
                       text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    b = open(out, "rb").read()
    nv, nc = np.frombuffer(b, np.int32, 2)
    o = 8
    color_of = np.frombuffer(b, np.int32, nv, o); o += 4 * nv
    y = np.frombuffer(b, np.float32, 3 * nv, o).reshape(-1, 3); o += 12 * nv
    Hc = np.frombuffer(b, np.float32, 9 * nv, o).reshape(-1, 3, 3); o += 36 * nv
    fc = np.frombuffer(b, np.float32, 3 * nv, o).reshape(-1, 3)
    prep = O.prepare(m)
    assert np.array_equal(color_of, prep.color_of)
    h = np.float32(1.0 / 60.0)
    g = np.asarray(m.gravity, np.float32)
    free = m.fixed == 0
    y_ref = cur.copy()
    y_ref[free] = cur[free] + h * (h * g)
    assert np.array_equal(y, y_ref)
    dminv, vol = O.precompute(m.rest, m.tets)
    mass, _, _, _ = O.mesh_init(m.rest, m.tets, m.density)
    sdt = float(np.float32(1.0 / np.float64(h * h)))
    H_ref = np.zeros((nv, 3, 3))
    f_ref = np.zeros((nv, 3))
    for v in np.nonzero(free)[0]:
        H_ref[v] = np.eye(3) * float(mass[v]) * sdt
        f_ref[v] = float(mass[v]) * (y[v].astype(np.float64) - cur[v]) * sdt
    for t, tet in enumerate(m.tets):
        x12 = cur[tet].astype(np.float64).reshape(12)
        for k, v in enumerate(tet):
            if not free[v]:
                continue
            Hk, gk = O.tet_block(x12, dminv[t].astype(np.float64), m.mu, m.lam, float(vol[t]), k, faithful=False)
            H_ref[v] += Hk
            f_ref[v] -= gk
    eH = np.abs(Hc - H_ref).max() / np.abs(H_ref).max()
    ef = np.abs(fc - f_ref).max() / np.abs(f_ref).max()
    print(f"host solveVertex vs FP64 closed form: H {eH:.2e}, f {ef:.2e}")
    assert eH <= 1e-5 and ef <= 1e-4


@pytest.mark.parametrize("openmp", [False, True])
def test_cpp_host_time_step_helpers_equal_the_float32_forms(tmp_path, openmp):
    """include/vbd/TetMesh.h applyGravity, computeInertia, updateVelocity and
    savePreviousPosition (TetMesh.h:305-348) on 100,003 vertices (above the size at which the
    loops go to OpenMP threads), built without and with OpenMP: bitwise the float32 forms
    v += h g, y = x + h v (pinned: y = x), v = ((x - x_prev) / h) * 0.999, x_prev = x on the free
    vertices, pinned velocities untouched."""
    n = 100003
    rng = np.random.default_rng(7)
    x = rng.normal(size=(n, 3)).astype(np.float32)
    xp = (x + rng.normal(scale=1e-3, size=(n, 3))).astype(np.float32)
    v = rng.normal(size=(n, 3)).astype(np.float32)
    fixed = (rng.random(n) < 0.1).astype(np.uint8)
    h = np.float32(1.0 / 60.0)
    g = np.asarray([0.0, -9.8, 0.0], np.float32)
    inp = tmp_path / "in.bin"
    inp.write_bytes(np.int32(n).tobytes() + x.tobytes() + xp.tobytes() + v.tobytes() + fixed.tobytes() +
                    h.tobytes() + g.tobytes())
    exe = str(tmp_path / "host_steps_check")
    lib = os.path.join(ROOT, "vbd_tetmesh_amd")
    flags = ["-fopenmp"] if openmp else []
    r = subprocess.run(["g++", "-O2", "-std=c++17", *flags, "-Wall", "-Wextra", "-Werror", "-I",
                        os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "cpp", "host_steps_check.cpp"),
                        "-L", lib, "-lvbd_hip", f"-Wl,-rpath,{lib}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = tmp_path / "out.bin"
    env = dict(os.environ, OMP_NUM_THREADS="4")
    r = subprocess.run([exe, str(inp), str(out)], capture_output=True, text=True, timeout=60, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.frombuffer(out.read_bytes(), np.float32).reshape(4, n, 3)
    free = fixed == 0
    v1 = v.copy()
    v1[free] = v[free] + h * g
    y = x.copy()
    y[free] = x[free] + h * v1[free]
    v2 = v1.copy()
    v2[free] = ((x[free] - xp[free]) / h) * np.float32(0.999)
    assert np.array_equal(got[0], v1)
    assert np.array_equal(got[1], y)
    assert np.array_equal(got[2], v2)
    assert np.array_equal(got[3], x)


def test_demo_binary_is_built():
    assert os.path.exists(DEMO), "vbd_dropin_demo not built (run __graft_entry__.build())"


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["hybrid", "step"])
def test_cpp_dropin_matches_oracle(tmp_path, mode):
    """main.cpp-style C++ program on the drop-in headers, TetGen input, beam 20x4x4."""
    m = meshgen.beam()
    prefix = str(tmp_path / "beam")
    meshgen.write_tetgen(prefix, m.rest, m.tets)
    out = str(tmp_path / "frames.bin")
    frames, iters = 12, 40
    print(f"[dropin {mode}]", flush=True)
    r = subprocess.run([DEMO, prefix, mode, str(frames), str(iters), "0.95", "0", "0.01", str(m.mu), str(m.lam), out,
                        str(tmp_path / "last.obj")], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, dtype=np.float32).reshape(frames, m.num_verts, 3)
    prep = O.prepare(m)
    p = O.make_params(dt=1.0 / 60.0, iterations=iters, rho=0.95, mu=m.mu, lam=m.lam,
                      init_from_inertia=(mode == "hybrid"))
    _, _, of, _ = O.simulate(prep, p, m.x0, None, frames)
    rel = max(float(np.abs(got[f].astype(np.float64) - of[f]).max() / np.abs(of[f]).max()) for f in range(frames))
    print(f"dropin {mode}: {rel:.3e}", flush=True)
    assert rel <= 1e-5
    obj = open(tmp_path / "last.obj").read()
    assert obj.count("\nv ") == m.num_verts
    # boundary of a 20x4x4 box: 2*(20*4 + 20*4 + 4*4) quads = 352 quads = 704 triangles
    assert obj.count("\nf ") == 704


def steel_block_mm():
    """A block in mm (8 x 10 x 8 one-mm cells, 891 verts, 3840 tets) and a warm-start current
    shape: compressed 8% along y and sheared 2% in x over y, like the compressed current-shape
    node file reference main.cpp:183 loads."""
    rest = meshgen.grid_positions(8, 10, 8, 1.0)
    tets = meshgen.grid_tets(8, 10, 8)
    cur = rest.copy()
    cur[:, 1] *= np.float32(0.92)
    cur[:, 0] += np.float32(0.02) * rest[:, 1]
    return rest, cur.astype(np.float32), tets


@pytest.mark.gpu
@pytest.mark.parametrize("material", [0, 1])
def test_cpp_reference_main_configuration_matches_oracle(tmp_path, material):
    """Reference main.cpp:171-232 as written, through the C++ drop-in: steel in mm (E = 200e3,
    nu = 0.3, density 7800e-12), warm start from the current-shape node file, the y <= min + 5%
    pin on the current positions (main.cpp:146-164), dt = 1e-3, 2 substeps, 50 iterations,
    Chebyshev rho 0.75, no gravity, host forwardStep / updateInertiaFromCPU / solveVBD (x <- y) /
    copyPositionsToCPU / updateVelocity, 20 frames; material 1 is the reference GPU kernel's
    corotated model, 0 the CPU solver's stable Neo-Hookean.  Here m/h^2 is about 1e-7 of the
    elastic diagonal.  Against the FP32 and FP64 oracles: per frame max|x_gpu - x_oracle| /
    max|x_oracle| <= 1e-5; the FP32 oracle's distance from FP64 is printed beside it."""
    rest, cur, tets = steel_block_mm()
    meshgen.write_tetgen(str(tmp_path / "block"), rest, tets)
    meshgen.write_tetgen(str(tmp_path / "block_compressed"), cur, tets)
    out = str(tmp_path / "frames.bin")
    frames = 20
    print(f"[reference main configuration] material {material}", flush=True)
    r = subprocess.run([DEMO, "main", str(tmp_path / "block.node"), str(tmp_path / "block_compressed.node"),
                        str(tmp_path / "block.ele"), str(frames), str(material), out],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, dtype=np.float32).reshape(frames, -1, 3)
    E, nu = np.float32(200e3), np.float32(0.3)
    mu = float(np.float32(E / (np.float32(2) * (np.float32(1) + nu))))
    lam = float(np.float32(E * nu / ((np.float32(1) + nu) * (np.float32(1) - np.float32(2) * nu))))
    lo, hi = cur[:, 1].min(), cur[:, 1].max()
    fixed = (cur[:, 1] <= lo + (hi - lo) * np.float32(0.05)).astype(np.uint8)
    m = meshgen.Mesh(name="steel_block_mm", rest=rest, x0=cur, tets=tets, fixed=fixed, mu=mu, lam=lam,
                     density=float(np.float32(7800e-12)), gravity=(0.0, 0.0, 0.0))
    prep = O.prepare(m)
    p = O.make_params(dt=1e-3, iterations=50, substeps=2, gravity=(0.0, 0.0, 0.0), rho=0.75, mu=mu, lam=lam,
                      init_from_inertia=True, material=material)
    _, _, of32, _ = O.simulate(prep, p, m.x0, None, frames)
    _, _, of64, _ = O.simulate(prep, p, m.x0, None, frames, f64=True, rest=m.rest, density=m.density)
    r32 = max(float(np.abs(got[f].astype(np.float64) - of32[f]).max() / np.abs(of32[f]).max()) for f in range(frames))
    r64 = max(float(np.abs(got[f].astype(np.float64) - of64[f]).max() / np.abs(of64[f]).max()) for f in range(frames))
    o = max(float(np.abs(of32[f] - of64[f]).max() / np.abs(of64[f]).max()) for f in range(frames))
    disp = float(np.abs(of64[-1] - cur).max())
    print(f"reference main, material {material}: gpu vs fp32 oracle {r32:.2e}, vs fp64 {r64:.2e}, fp32 oracle vs "
          f"fp64 {o:.2e}; {int(fixed.sum())} pinned, displacement after {frames} frames {disp:.3f} mm", flush=True)
    assert disp > 0.1  # the warm start relaxes
    assert r32 <= 1e-5 and r64 <= 1e-5
