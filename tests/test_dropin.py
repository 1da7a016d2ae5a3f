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
def test_tetgen_round_trip(tmp_path, one_based):
    m = meshgen.beam(4, 2, 2)
    prefix = str(tmp_path / "beam")
    meshgen.write_tetgen(prefix, m.rest, m.tets, one_based=one_based)
    pos, tets = meshgen.read_tetgen(prefix)
    assert np.array_equal(pos, m.rest) and np.array_equal(tets, m.tets)


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
