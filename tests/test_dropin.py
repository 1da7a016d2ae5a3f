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
