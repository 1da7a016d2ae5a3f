"""Writes tests/golden/oracle_small.npz: small fixtures of the CPU oracle (oracle/), committed so
that later changes to the oracle or to the device path are checked against fixed data.

Contents (all from the oracle; none from the reference, which ships no fixtures):
  tet1_*   one tet: rest and deformed corner positions (12), DmInv (9, FP32 precompute), V0;
           per corner k the 3x3 block and gradient of the closed form and of the Energy.cpp
           faithful form, in FP64 (mu = 1e5, lambda = 4e5)
  tet2_*   two tets sharing a face (5 vertices), the same per tet and corner
  beam_x1  the 20x4x4 beam after one frame (100 iterations, Chebyshev 0.95, h = 1/60), FP32
  beam_color_of, cube16_color_of   greedy colorings of the beam and of the 16^3 noisy cube

usage: python tests/golden/make_oracle_golden.py [out.npz]
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from vbd_tetmesh_amd import meshgen  # noqa: E402

MU, LAM = 1e5, 4e5
H = 1.0 / 60.0


def tet_fixture(rest, tets, x, prefix, out):
    dminv, vol = O.precompute(rest, tets)
    blocks, grads, fblocks, fgrads = [], [], [], []
    for t in range(tets.shape[0]):
        x12 = x[tets[t]].reshape(12)
        for k in range(4):
            Hc, gc = O.tet_block(x12, dminv[t], MU, LAM, float(vol[t]), k, faithful=False)
            Hf, gf = O.tet_block(x12, dminv[t], MU, LAM, float(vol[t]), k, faithful=True)
            blocks.append(Hc)
            grads.append(gc)
            fblocks.append(Hf)
            fgrads.append(gf)
    out[prefix + "rest"] = rest
    out[prefix + "x"] = x
    out[prefix + "tets"] = tets
    out[prefix + "dminv"] = np.asarray(dminv, np.float32).reshape(-1, 9)
    out[prefix + "vol"] = np.asarray(vol, np.float32)
    out[prefix + "block"] = np.array(blocks).reshape(tets.shape[0], 4, 3, 3)
    out[prefix + "grad"] = np.array(grads).reshape(tets.shape[0], 4, 3)
    out[prefix + "block_faithful"] = np.array(fblocks).reshape(tets.shape[0], 4, 3, 3)
    out[prefix + "grad_faithful"] = np.array(fgrads).reshape(tets.shape[0], 4, 3)


def build():
    rng = np.random.default_rng(11)
    out = {}
    rest1 = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1]], np.float32)
    x1 = (rest1 + rng.normal(scale=0.01, size=rest1.shape)).astype(np.float32)
    tet_fixture(rest1, np.array([[0, 1, 2, 3]], np.int32), x1, "tet1_", out)
    rest2 = np.array([[0, 0, 0], [0.1, 0, 0], [0, 0.1, 0], [0, 0, 0.1], [0.1, 0.1, 0.1]], np.float32)
    x2 = (rest2 + rng.normal(scale=0.01, size=rest2.shape)).astype(np.float32)
    tet_fixture(rest2, np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32), x2, "tet2_", out)
    m = meshgen.beam()
    prep = O.prepare(m)
    _, _, of, _ = O.simulate(prep, O.make_params(dt=H, iterations=100, rho=0.95, mu=m.mu, lam=m.lam), m.x0, None, 1)
    out["beam_x1"] = np.asarray(of[0], np.float32)
    out["beam_color_of"] = np.asarray(prep.color_of, np.int32)
    c = meshgen.noisy_cube(16, seed=1)
    out["cube16_color_of"] = np.asarray(O.prepare(c).color_of, np.int32)
    return out


def main():
    path = sys.argv[1] if len(sys.argv) > 1 else os.path.join(os.path.dirname(os.path.abspath(__file__)),
                                                               "oracle_small.npz")
    out = build()
    np.savez_compressed(path, **out)
    print(path, {k: v.shape for k, v in out.items()})


if __name__ == "__main__":
    main()
