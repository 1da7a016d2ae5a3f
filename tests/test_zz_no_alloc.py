"""GPU: vbd_total_energy_noalloc, the energy entry point that allocates nothing, against
vbd_total_energy through the C ABI: bitwise the same value (same kernels, same fixed summation
order), on two calls in a row, the second on a smaller mesh (the shared scratch keeps the first
call's partial sums, which the second must not read)."""
import ctypes as C

import numpy as np
import pytest

from vbd_tetmesh_amd import VBDParams, meshgen
from vbd_tetmesh_amd import _capi as A
from vbd_tetmesh_amd.solver import upload

pytestmark = pytest.mark.gpu


def noalloc_energy(d, mu, lam, out):
    """Calls the no-allocation entry point with a device-resident output double."""
    rc = A.lib.vbd_total_energy_noalloc(d.h, mu, lam, C.c_void_p(out.data_ptr()))
    A.check(rc, d.h)
    d.synchronize()
    return float(out.cpu().item())


def test_noalloc_energy_equals_total_energy_on_two_calls():
    import torch
    out = torch.full((1,), -1.0, dtype=torch.float64, device="cuda")
    big = meshgen.noisy_cube(90, seed=2)     # 4,374,000 tets, 17,086 partial sums: two-stage final sum
    small = meshgen.noisy_cube(6, seed=3)    # 1,296 tets: one block sums the partials
    for m, name in ((big, "cube90"), (small, "cube6")):
        print(f"[noalloc] {name}: nV={m.num_verts} nT={m.num_tets}", flush=True)
        d = upload(m, params=VBDParams(dt=1.0 / 60.0, numIterations=5, gravity=m.gravity, mu=m.mu, lambda_=m.lam))
        d.step(1)  # a deformed state
        ref = d.total_energy(m.mu, m.lam)
        got = noalloc_energy(d, m.mu, m.lam, out)
        print(f"[noalloc] {name}: energy {got!r} vs vbd_total_energy {ref!r}", flush=True)
        assert ref != 0.0
        assert got == ref


def test_noalloc_energy_rejects_bad_arguments():
    m = meshgen.beam(4, 2, 2)
    d = upload(m)
    assert A.lib.vbd_total_energy_noalloc(d.h, m.mu, m.lam, None) == A.VBD_ERR_ARG
    from vbd_tetmesh_amd.solver import DeviceMesh
    e = DeviceMesh(0)  # not initialised
    assert A.lib.vbd_total_energy_noalloc(e.h, m.mu, m.lam, C.c_void_p(1)) == A.VBD_ERR_STATE
