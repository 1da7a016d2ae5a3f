"""ctypes front-end of the VBD CPU oracle (vbd_oracle.cpp) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this module,
and only as the checker / the timed CPU baseline.  The product package never imports it.

Each wrapper cites the reference function it restates (paths relative to the reference repo).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class Params(C.Structure):
    """Mirror of VBDParams (VBD.h:14-25) plus the oracle's extension fields."""

    _fields_ = [
        ("dt", C.c_float),
        ("numIterations", C.c_int32),
        ("substeps", C.c_int32),
        ("gravity", C.c_float * 3),
        ("useAcceleration", C.c_int32),
        ("accelerationRho", C.c_float),
        ("mu", C.c_float),
        ("lambda_", C.c_float),
        ("initFromInertia", C.c_int32),
        ("contactOn", C.c_int32),
        ("contactAxis", C.c_int32),
        ("contactK", C.c_float),
        ("contactPlane", C.c_float),
        ("mode", C.c_int32),
        ("threads", C.c_int32),
        ("energyEvery", C.c_int32),
        ("material", C.c_int32),
    ]


MODE_CLOSED, MODE_FAITHFUL, MODE_FAITHFUL_COST = 0, 1, 2


def make_params(dt=1.0 / 60.0, iterations=100, substeps=1, gravity=(0.0, -9.8, 0.0), accel=True, rho=0.95,
                mu=1e5, lam=4e5, init_from_inertia=False, contact=None, mode=MODE_CLOSED, threads=1,
                energy_every=0, material=0) -> Params:
    p = Params()
    p.dt = dt
    p.numIterations = iterations
    p.substeps = substeps
    p.gravity[:] = list(gravity)
    p.useAcceleration = int(bool(accel))
    p.accelerationRho = rho
    p.mu = mu
    p.lambda_ = lam
    p.initFromInertia = int(bool(init_from_inertia))
    if contact is not None:
        axis, plane, k = contact
        p.contactOn, p.contactAxis, p.contactPlane, p.contactK = 1, int(axis), float(plane), float(k)
    p.mode = mode
    p.threads = threads
    p.energyEvery = energy_every
    p.material = material
    return p


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "_oracle.so")
        if not os.path.exists(path):
            subprocess.run(["make", "-s", "-C", _HERE], check=True)
        L = C.CDLL(path)
        f32p, f64p, i32p, u8p = C.POINTER(C.c_float), C.POINTER(C.c_double), C.POINTER(C.c_int32), C.POINTER(C.c_uint8)
        L.orc_params_size.restype = C.c_int
        L.orc_mesh_init_f32.argtypes = [C.c_int, f32p, C.c_int, i32p, C.c_float, f32p, f32p, i32p, i32p]
        L.orc_mesh_init_f64.argtypes = [C.c_int, f32p, C.c_int, i32p, C.c_double, f64p, f64p]
        L.orc_color.argtypes = [C.c_int, i32p, i32p, i32p, i32p, i32p, i32p]
        L.orc_color.restype = C.c_int
        L.orc_precompute_f32.argtypes = [f32p, C.c_int, i32p, f32p, f32p]
        L.orc_precompute_f64.argtypes = [f32p, C.c_int, i32p, f64p, f64p]
        L.orc_tet_full_f64.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double, f64p, f64p]
        L.orc_tet_full_f32.argtypes = [f32p, f32p, C.c_float, C.c_float, C.c_float, f32p, f32p]
        L.orc_tet_block_f64.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, f64p, f64p]
        L.orc_tet_block_f32.argtypes = [f32p, f32p, C.c_float, C.c_float, C.c_float, C.c_int, C.c_int, f32p, f32p]
        L.orc_tet_block_corot_f64.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double, C.c_int, f64p, f64p]
        L.orc_polar_rotation_f64.argtypes = [f64p, f64p]
        L.orc_tet_energy_f64.argtypes = [f64p, f64p, C.c_double, C.c_double, C.c_double]
        L.orc_tet_energy_f64.restype = C.c_double
        L.orc_omega_f32.argtypes = [C.c_int, C.c_float, C.c_float]
        L.orc_omega_f32.restype = C.c_double
        sim_common = [C.POINTER(Params), C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, i32p, i32p, u8p]
        L.orc_simulate_f32.argtypes = sim_common + [f32p] * 5 + [f32p, f32p, C.c_int, f32p]
        L.orc_simulate_f32.restype = C.c_double
        L.orc_simulate_f64.argtypes = sim_common + [f64p] * 5 + [f64p, f64p, C.c_int, f64p]
        L.orc_simulate_f64.restype = C.c_double
        L.orc_solve_subset_f32.argtypes = [C.POINTER(Params), C.c_int, C.c_int, i32p, i32p, i32p, u8p] + \
            [f32p] * 5 + [f32p, f32p, C.c_int, i32p, f32p]
        L.orc_solve_subset_f32.restype = C.c_double
        L.orc_sweep_over_colors_f32.argtypes = [C.POINTER(Params), C.c_int, C.c_int, i32p, i32p, i32p, C.c_int, i32p,
                                                i32p, u8p] + [f32p] * 5 + [f32p, f32p]
        L.orc_sweep_over_colors_f32.restype = C.c_double
        L.orc_total_energy_f32.argtypes = [C.POINTER(Params), C.c_int, i32p, f32p, f32p, f32p, f32p, f32p]
        L.orc_total_energy_f32.restype = C.c_double
        assert L.orc_params_size() == C.sizeof(Params), "oracle Params layout mismatch"
        _LIB = L
    return _LIB


@dataclass
class Prepared:
    """Host-side precomputation of one mesh, as the reference CPU solver holds it."""

    nV: int
    nT: int
    tets: np.ndarray      # (T,4) int32
    fixed: np.ndarray     # (N,) uint8
    mass: np.ndarray      # (N,) float32   TetMesh::vertexMass
    rest_vol: np.ndarray  # (T,) float32   TetMesh::tetRestVol
    adj_off: np.ndarray   # (N+1,) int32   TetMesh::vertAdjacentTets as CSR
    adj: np.ndarray       # (4T,) int32
    color_of: np.ndarray  # (N,) int32
    color_off: np.ndarray  # (C+1,) int32   VBDsolver::vertexColors as CSR
    color_verts: np.ndarray
    dminv: np.ndarray     # (T,9) float32 VBDsolver::tetDmInv (row-major)
    vol: np.ndarray       # (T,) float32   VBDsolver::tetRestVolume
    tet_mu: Optional[np.ndarray] = None
    tet_lambda: Optional[np.ndarray] = None

    @property
    def n_colors(self) -> int:
        return int(self.color_off.size - 1)


def mesh_init(rest: np.ndarray, tets: np.ndarray, density: float):
    """TetMesh::initialize (TetMesh.h:213-275): (mass, tetRestVol, adj_off, adj)."""
    L = lib()
    rest = np.ascontiguousarray(rest, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    nV, nT = rest.shape[0], tets.shape[0]
    mass = np.zeros(nV, np.float32)
    vol = np.zeros(nT, np.float32)
    off = np.zeros(nV + 1, np.int32)
    adj = np.zeros(4 * nT, np.int32)
    L.orc_mesh_init_f32(nV, _p(rest, C.c_float), nT, _p(tets, C.c_int32), density, _p(mass, C.c_float),
                        _p(vol, C.c_float), _p(off, C.c_int32), _p(adj, C.c_int32))
    return mass, vol, off, adj


def mesh_init_f64(rest, tets, density):
    L = lib()
    rest = np.ascontiguousarray(rest, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    mass = np.zeros(rest.shape[0], np.float64)
    vol = np.zeros(tets.shape[0], np.float64)
    L.orc_mesh_init_f64(rest.shape[0], _p(rest, C.c_float), tets.shape[0], _p(tets, C.c_int32), density,
                        _p(mass, C.c_double), _p(vol, C.c_double))
    return mass, vol


def color(nV, tets, adj_off, adj):
    """GraphColoring::colorVertices (GraphColoring.cpp:28-66): (color_of, color_off, color_verts)."""
    L = lib()
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    color_of = np.zeros(nV, np.int32)
    off = np.zeros(nV + 1, np.int32)
    verts = np.zeros(nV, np.int32)
    nc = L.orc_color(nV, _p(tets, C.c_int32), _p(adj_off, C.c_int32), _p(adj, C.c_int32), _p(color_of, C.c_int32),
                     _p(off, C.c_int32), _p(verts, C.c_int32))
    return color_of, off[: nc + 1].copy(), verts


def precompute(rest, tets, f64=False):
    """VBDsolver::precomputeTetData (VBD.cpp:36-58): (DmInv row-major (T,9), V0 (T,))."""
    L = lib()
    rest = np.ascontiguousarray(rest, dtype=np.float32)
    tets = np.ascontiguousarray(tets, dtype=np.int32)
    nT = tets.shape[0]
    if f64:
        d = np.zeros((nT, 9), np.float64)
        v = np.zeros(nT, np.float64)
        L.orc_precompute_f64(_p(rest, C.c_float), nT, _p(tets, C.c_int32), _p(d, C.c_double), _p(v, C.c_double))
    else:
        d = np.zeros((nT, 9), np.float32)
        v = np.zeros(nT, np.float32)
        L.orc_precompute_f32(_p(rest, C.c_float), nT, _p(tets, C.c_int32), _p(d, C.c_float), _p(v, C.c_float))
    return d, v


def prepare(mesh) -> Prepared:
    """Everything VBDsolver::initializeColoring + precomputeTetData derive from a mesh."""
    mass, rvol, off, adj = mesh_init(mesh.rest, mesh.tets, mesh.density)
    color_of, coff, cverts = color(mesh.num_verts, mesh.tets, off, adj)
    dminv, vol = precompute(mesh.rest, mesh.tets)
    return Prepared(mesh.num_verts, mesh.num_tets, np.ascontiguousarray(mesh.tets, np.int32),
                    np.ascontiguousarray(mesh.fixed, np.uint8), mass, rvol, off, adj, color_of, coff, cverts, dminv,
                    vol, mesh.tet_mu, mesh.tet_lambda)


def simulate(prep: Prepared, params: Params, x0: np.ndarray, v0: Optional[np.ndarray], frames: int,
             f64: bool = False, record: bool = True, rest: Optional[np.ndarray] = None, density: float = 1000.0):
    """Run `frames` calls of VBDsolver::step (VBD.cpp:60-101).

    Returns (x_final (N,3), v_final (N,3), per-frame positions (frames,N,3) or None, seconds).
    With f64=True everything (mass, DmInv, V0, state) is recomputed and carried in double;
    `rest` and `density` are then required.
    """
    L = lib()
    nV, nT = prep.nV, prep.nT
    dt = np.float64 if f64 else np.float32
    ct = C.c_double if f64 else C.c_float
    x = np.ascontiguousarray(x0, dtype=dt).copy()
    v = np.zeros((nV, 3), dt) if v0 is None else np.ascontiguousarray(v0, dtype=dt).copy()
    if f64:
        mass, _ = mesh_init_f64(rest, prep.tets, density)
        dminv, vol = precompute(rest, prep.tets, f64=True)
    else:
        mass, dminv, vol = prep.mass, prep.dminv, prep.vol
    tmu = None if prep.tet_mu is None else np.ascontiguousarray(prep.tet_mu, dtype=dt)
    tlam = None if prep.tet_lambda is None else np.ascontiguousarray(prep.tet_lambda, dtype=dt)
    out = np.zeros((frames, nV, 3), dt) if record else None
    fn = L.orc_simulate_f64 if f64 else L.orc_simulate_f32
    secs = fn(C.byref(params), nV, nT, _p(prep.tets, C.c_int32), _p(prep.adj_off, C.c_int32), _p(prep.adj, C.c_int32),
              prep.n_colors, _p(prep.color_off, C.c_int32), _p(prep.color_verts, C.c_int32),
              _p(prep.fixed, C.c_uint8), _p(np.ascontiguousarray(mass, dt), ct),
              _p(np.ascontiguousarray(dminv, dt), ct), _p(np.ascontiguousarray(vol, dt), ct), _p(tmu, ct),
              _p(tlam, ct), _p(x, ct), _p(v, ct), frames, _p(out, ct))
    return x, v, out, secs


def solve_subset(prep: Prepared, params: Params, x: np.ndarray, y: np.ndarray, verts: np.ndarray):
    """Per-vertex local solves (VBD.cpp:124-203) for a same-color vertex subset against
    positions x; returns (new positions (n,3), seconds)."""
    L = lib()
    x = np.ascontiguousarray(x, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    verts = np.ascontiguousarray(verts, np.int32)
    out = np.zeros((verts.size, 3), np.float32)
    tmu = None if prep.tet_mu is None else np.ascontiguousarray(prep.tet_mu, np.float32)
    tlam = None if prep.tet_lambda is None else np.ascontiguousarray(prep.tet_lambda, np.float32)
    secs = L.orc_solve_subset_f32(C.byref(params), prep.nV, prep.nT, _p(prep.tets, C.c_int32),
                                  _p(prep.adj_off, C.c_int32), _p(prep.adj, C.c_int32), _p(prep.fixed, C.c_uint8),
                                  _p(prep.mass, C.c_float), _p(prep.dminv, C.c_float), _p(prep.vol, C.c_float),
                                  _p(tmu, C.c_float), _p(tlam, C.c_float), _p(x, C.c_float), _p(y, C.c_float),
                                  verts.size, _p(verts, C.c_int32), _p(out, C.c_float))
    return out, secs


def sweep_over_colors(prep: Prepared, params: Params, x: np.ndarray, y: np.ndarray):
    """One sweep in the reference's scheme (VBD.cpp:119-121: OpenMP over colors, params.threads
    threads; racy, timing only).  Returns (positions after the sweep, seconds)."""
    L = lib()
    x = np.ascontiguousarray(x, np.float32).copy()
    y = np.ascontiguousarray(y, np.float32)
    tmu = None if prep.tet_mu is None else np.ascontiguousarray(prep.tet_mu, np.float32)
    tlam = None if prep.tet_lambda is None else np.ascontiguousarray(prep.tet_lambda, np.float32)
    secs = L.orc_sweep_over_colors_f32(C.byref(params), prep.nV, prep.nT, _p(prep.tets, C.c_int32),
                                       _p(prep.adj_off, C.c_int32), _p(prep.adj, C.c_int32), prep.n_colors,
                                       _p(prep.color_off, C.c_int32), _p(prep.color_verts, C.c_int32),
                                       _p(prep.fixed, C.c_uint8), _p(prep.mass, C.c_float), _p(prep.dminv, C.c_float),
                                       _p(prep.vol, C.c_float), _p(tmu, C.c_float), _p(tlam, C.c_float),
                                       _p(x, C.c_float), _p(y, C.c_float))
    return x, secs


def tet_full(x12, dminv9, mu, lam, V, f64=True):
    """Energy.cpp:90-158 on one tet: (H (12,12), g (12,))."""
    L = lib()
    dt, ct = (np.float64, C.c_double) if f64 else (np.float32, C.c_float)
    x12 = np.ascontiguousarray(x12, dt).reshape(12)
    d9 = np.ascontiguousarray(dminv9, dt).reshape(9)
    H = np.zeros((12, 12), dt)
    g = np.zeros(12, dt)
    fn = L.orc_tet_full_f64 if f64 else L.orc_tet_full_f32
    fn(_p(x12, ct), _p(d9, ct), mu, lam, V, _p(H, ct), _p(g, ct))
    return H, g


def tet_block(x12, dminv9, mu, lam, V, k, faithful, f64=True):
    """(k,k) block and gradient of corner k: faithful (Energy.cpp path) or closed form."""
    L = lib()
    dt, ct = (np.float64, C.c_double) if f64 else (np.float32, C.c_float)
    x12 = np.ascontiguousarray(x12, dt).reshape(12)
    d9 = np.ascontiguousarray(dminv9, dt).reshape(9)
    H = np.zeros((3, 3), dt)
    g = np.zeros(3, dt)
    fn = L.orc_tet_block_f64 if f64 else L.orc_tet_block_f32
    fn(_p(x12, ct), _p(d9, ct), mu, lam, V, int(k), int(bool(faithful)), _p(H, ct), _p(g, ct))
    return H, g


def tet_block_corotated(x12, dminv9, mu, lam, V, k):
    """(k,k) block and gradient of corner k for the GPU wrapper's corotated model
    (VBD_DeviceMath.cuh:127-198), in double."""
    L = lib()
    x12 = np.ascontiguousarray(x12, np.float64).reshape(12)
    d9 = np.ascontiguousarray(dminv9, np.float64).reshape(9)
    H = np.zeros((3, 3))
    g = np.zeros(3)
    L.orc_tet_block_corot_f64(_p(x12, C.c_double), _p(d9, C.c_double), mu, lam, V, int(k), _p(H, C.c_double),
                              _p(g, C.c_double))
    return H, g


def polar_rotation(F):
    """The corotated model's rotation: 5 steps of R <- (R + cof(R)/det R) / 2 from R = F."""
    L = lib()
    F9 = np.ascontiguousarray(F, np.float64).reshape(9)
    R = np.zeros(9)
    L.orc_polar_rotation_f64(_p(F9, C.c_double), _p(R, C.c_double))
    return R.reshape(3, 3)


def tet_energy(x12, dminv9, mu, lam, V) -> float:
    """V * [mu/2(|F|^2-3) - mu(J-1) + lambda/2 (J-1)^2], the potential of Energy.cpp:97."""
    L = lib()
    x12 = np.ascontiguousarray(x12, np.float64).reshape(12)
    d9 = np.ascontiguousarray(dminv9, np.float64).reshape(9)
    return float(L.orc_tet_energy_f64(_p(x12, C.c_double), _p(d9, C.c_double), mu, lam, V))


def omega(order: int, rho: float, prev: float) -> float:
    """VBDsolver::getAcceleratorOmega (VBD.cpp:302-311) in the reference's float/double mix."""
    return float(lib().orc_omega_f32(order, rho, prev))


def total_energy(prep: Prepared, params: Params, x: np.ndarray) -> float:
    """VBDsolver::computeTotalEnergy (VBD.cpp:258-288), |det F| in place of the SVD product."""
    L = lib()
    x = np.ascontiguousarray(x, np.float32)
    tmu = None if prep.tet_mu is None else np.ascontiguousarray(prep.tet_mu, np.float32)
    tlam = None if prep.tet_lambda is None else np.ascontiguousarray(prep.tet_lambda, np.float32)
    return float(L.orc_total_energy_f32(C.byref(params), prep.nT, _p(prep.tets, C.c_int32), _p(prep.dminv, C.c_float),
                                        _p(prep.vol, C.c_float), _p(tmu, C.c_float), _p(tlam, C.c_float),
                                        _p(x, C.c_float)))
