"""Writes tests/golden/diffcube16_fp64.npz: the FP64 CPU oracle (oracle/diff_oracle.py) on the
differentiable workload -- 16^3-cell cube (4913 vertices, 24576 tets), per-tet mu from a seeded
smooth field in [5e4, 2e5], lambda = 4 mu, bottom face pinned, gravity -y, initial velocity a
seeded smooth field, 30 steps at h = 1/60 each solved to |grad G| <= 1e-10 |M g|; loss
L = |x_30 - x_target|^2 with x_target from the seed-drawn ground-truth mu field.

Stores: x_target, the loss, the full adjoint gradients (float32) w.r.t. per-tet mu, lambda and
v0, 32 seeded parameter indices and their central finite differences (float64 forward
warm-started from the base trajectory; step 1e-3 of the parameter for mu and lambda, 1e-3 of
max |v0| for v0 components, some of which are near zero).

usage: python tests/golden/make_diffcube_golden.py [n] [steps] [out.npz] [--only mu,lam,v0]
  --only recomputes the finite differences of the listed parameter families and keeps the
  others from an existing out.npz (same n, steps and seeds).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import diff_oracle as D  # noqa: E402
from vbd_tetmesh_amd import meshgen  # noqa: E402


def main():
    argv = list(sys.argv[1:])
    only = None
    if "--only" in argv:
        i = argv.index("--only")
        only = set(argv[i + 1].split(","))
        del argv[i:i + 2]
    n = int(argv[0]) if len(argv) > 0 else 16
    steps = int(argv[1]) if len(argv) > 1 else 30
    out = argv[2] if len(argv) > 2 else os.path.join(ROOT, "tests", "golden", f"diffcube{n}_fp64.npz")
    old_fd = None
    if only is not None:
        old = np.load(out)
        assert int(old["n"]) == n and int(old["steps"]) == steps, "existing fixture has other sizes"
        old_fd = list(old["fd"])
    m, mu_true, lam_true, v0 = meshgen.diff_workload(n, seed=0)
    P = D.make_problem(m)
    mu = m.tet_mu.astype(np.float64)
    lam = m.tet_lambda.astype(np.float64)
    x0 = m.rest.astype(np.float64)
    v0d = v0.astype(np.float64)
    t0 = time.time()
    xs_t, _ = D.forward(P, x0, v0d, mu_true.astype(np.float64), lam_true.astype(np.float64), steps)
    x_target = xs_t[-1]
    print(f"target done {time.time() - t0:.1f}s", flush=True)
    L, gmu, glam, gv0, xs = D.loss_and_grads(P, x0, v0d, mu, lam, steps, x_target)
    print(f"loss {L:.6e} grads done {time.time() - t0:.1f}s", flush=True)
    rng = np.random.default_rng(1234)
    T = m.num_tets
    free = np.nonzero(m.fixed == 0)[0]
    idx_mu = rng.choice(T, 12, replace=False)
    idx_lam = rng.choice(T, 12, replace=False)
    idx_v = np.stack([rng.choice(free, 8, replace=False), rng.integers(0, 3, 8)], axis=1)
    fd = []
    eps_rel = 1e-3

    def loss_of(mu_, lam_, v_):
        xs_, _ = D.forward(P, x0, v_, mu_, lam_, steps, guesses=xs)
        d = xs_[-1] - x_target
        return float(np.sum(d * d))

    j = -1
    for kind, ids in (("mu", idx_mu), ("lam", idx_lam), ("v0", idx_v)):
        for i in ids:
            j += 1
            if only is not None and kind not in only:
                fd.append(old_fd[j])
                continue
            mp, lp, vp = mu.copy(), lam.copy(), v0d.copy()
            mm, lm, vm = mu.copy(), lam.copy(), v0d.copy()
            if kind == "mu":
                e = eps_rel * mu[i]; mp[i] += e; mm[i] -= e
            elif kind == "lam":
                e = eps_rel * lam[i]; lp[i] += e; lm[i] -= e
            else:
                e = eps_rel * np.abs(v0d).max(); vp[i[0], i[1]] += e; vm[i[0], i[1]] -= e
            d = (loss_of(mp, lp, vp) - loss_of(mm, lm, vm)) / (2 * e)
            fd.append(d)
            print(f"{kind} {i} fd {d:.6e} adjoint "
                  f"{(gmu[i] if kind == 'mu' else glam[i] if kind == 'lam' else gv0[i[0], i[1]]):.6e} "
                  f"{time.time() - t0:.0f}s", flush=True)
    np.savez_compressed(out, n=n, steps=steps, x_target=x_target, loss=L, grad_mu=gmu.astype(np.float32),
                        grad_lambda=glam.astype(np.float32), grad_v0=gv0.astype(np.float32), idx_mu=idx_mu,
                        idx_lambda=idx_lam, idx_v0=idx_v, fd=np.asarray(fd), eps_rel=eps_rel)
    print("wrote", out, flush=True)


if __name__ == "__main__":
    main()
