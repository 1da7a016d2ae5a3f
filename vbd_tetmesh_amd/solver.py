"""Host-side mirror of the reference VBD interfaces over the gfx950 C ABI.

Names, argument meaning and call order follow the reference so code written against it
reads the same (paths are relative to the reference repository):

  * ``VBDParams``      -- VBD.h:14-25
  * ``TetMesh``        -- TetMesh.h:19-349 (state arrays, initialize, time-step helpers)
  * ``VBD_GPU_Solver`` -- VBD_GPU_Wrapper.h:7-45 (initFromCPU, solveIteration, solveVBD,
                          copyPositionsToCPU, updateInertiaFromCPU)
  * ``VBDsolver``      -- VBD.h:33-140 (setParams, initializeColoring, precomputeTetData,
                          step, simulate, computeTotalEnergy, getters), with the whole
                          step running on the device.

Vertex arrays are (N, 3) float32 numpy arrays (rows = vertices), the transpose view of the
reference's 3 x N column-major TVerticesMat, i.e. the same memory layout.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import Callable, List, Optional, Sequence, Tuple

import numpy as np

from . import _capi as A
from ._capi import check, lib, ptr


@dataclass
class VBDParams:
    """VBDParams (VBD.h:14-25) plus the build's extensions."""

    dt: float = 0.01666666
    numIterations: int = 50
    substeps: int = 1
    gravity: tuple = (0.0, -9.8, 0.0)
    useAcceleration: bool = False
    accelerationRho: float = 0.75
    numThreads: int = 4          # accepted for source compatibility; the device path ignores it
    mu: float = 1e6
    lambda_: float = 1e6
    density: float = 1000.0
    # extensions (no reference counterpart)
    velocityDamping: float = 0.999   # TetMesh.h:308 hard-codes 0.999
    startFromInertia: bool = False   # True reproduces VBD_GPU_Wrapper.cu:334 (x <- y)
    contact: Optional[tuple] = None  # (axis, plane, stiffness) penalty half-space x[axis] >= plane
    material: int = A.VBD_MATERIAL_STABLE_NEO_HOOKEAN  # or A.VBD_MATERIAL_COROTATED_GPU
    energyLogEvery: int = 10  # VBDsolver.step records the diagnostic energy every n-th iteration; 0: off

    def to_c(self) -> A.StepParams:
        p = A.StepParams()
        p.dt = self.dt
        p.num_iterations = int(self.numIterations)
        p.substeps = int(self.substeps)
        p.gravity[:] = [float(g) for g in self.gravity]
        p.use_acceleration = int(bool(self.useAcceleration))
        p.acceleration_rho = self.accelerationRho
        p.mu = self.mu
        p.lambda_ = self.lambda_
        p.velocity_damping = self.velocityDamping
        p.start_from_inertia = int(bool(self.startFromInertia))
        if self.contact is not None:
            axis, plane, k = self.contact
            p.contact_on, p.contact_axis, p.contact_plane, p.contact_k = 1, int(axis), float(plane), float(k)
        return p


class TetMesh:
    """TetMesh (TetMesh.h:19-349).  Adjacency is held as CSR (vertAdjacentTets)."""

    def __init__(self):
        self.numVerts = 0
        self.numTets = 0

    def initialize(self, rest_positions, current_positions, density, tetIndices, mu, lambda_, fixedVerts=()):
        """TetMesh::initialize (TetMesh.h:213-275)."""
        rest = np.ascontiguousarray(rest_positions, np.float32).reshape(-1, 3)
        cur = np.ascontiguousarray(current_positions, np.float32).reshape(-1, 3)
        if rest.shape != cur.shape:
            raise ValueError("rest and current positions differ in size")
        self.numVerts = rest.shape[0]
        self.tets = np.ascontiguousarray(tetIndices, np.int32).reshape(-1, 4)
        self.numTets = self.tets.shape[0]
        self.mRestPos = rest.copy()
        self.mVertPos = cur.copy()
        self.mVertPrevPos = cur.copy()
        self.mVelocity = np.zeros_like(cur)
        self.mInertia = np.zeros_like(cur)
        mass, vol, off, adj = A.mesh_initialize(rest, self.tets, density)
        self.vertexMass = mass
        self.tetRestVol = vol
        self.tetMu = np.full(self.numTets, mu, np.float32)
        self.tetLambda = np.full(self.numTets, lambda_, np.float32)
        self.adjOffsets = off
        self.adjTets = adj
        self.isFixed = np.zeros(self.numVerts, np.uint8)
        fv = np.asarray(list(fixedVerts), dtype=np.int64)
        fv = fv[(fv >= 0) & (fv < self.numVerts)]
        self.isFixed[fv] = 1
        return self

    @property
    def vertAdjacentTets(self) -> List[np.ndarray]:
        return [self.adjTets[self.adjOffsets[i]: self.adjOffsets[i + 1]] for i in range(self.numVerts)]

    # time-step helpers (host), TetMesh.h:305-348; float32 arithmetic in the reference's order,
    # evaluated in place over all rows and stored to the free rows only
    def updateVelocity(self, dt):
        free = (self.isFixed == 0)[:, None]
        t = np.subtract(self.mVertPos, self.mVertPrevPos)
        t /= np.float32(dt)
        t *= np.float32(0.999)
        np.copyto(self.mVelocity, t, where=free)

    def applyGravity(self, gravity, dt):
        free = (self.isFixed == 0)[:, None]
        np.add(self.mVelocity, np.float32(dt) * np.asarray(gravity, np.float32), out=self.mVelocity, where=free)

    def computeInertia(self, dt):
        fixed = (self.isFixed != 0)[:, None]
        np.multiply(self.mVelocity, np.float32(dt), out=self.mInertia)
        np.add(self.mVertPos, self.mInertia, out=self.mInertia)
        np.copyto(self.mInertia, self.mVertPos, where=fixed)

    def savePreviousPosition(self):
        np.copyto(self.mVertPrevPos, self.mVertPos)

    # surface output, TetMesh.h:45-186
    def extractSurfaceFaces(self) -> np.ndarray:
        """Boundary triangles (faces used by exactly one tet), oriented as the tet's outward face
        ((0,2,1), (0,1,3), (0,3,2), (1,2,3) of a positively oriented tet), ordered by sorted ids."""
        if self.numTets == 0:
            return np.zeros((0, 3), np.int32)
        t = self.tets
        tri = np.concatenate([t[:, [0, 2, 1]], t[:, [0, 1, 3]], t[:, [0, 3, 2]], t[:, [1, 2, 3]]])
        key = np.sort(tri, axis=1)
        _, first, counts = np.unique(key, axis=0, return_index=True, return_counts=True)
        return np.ascontiguousarray(tri[first[counts == 1]], np.int32)

    def exportToObj(self, filepath: str) -> bool:
        """Wavefront OBJ of the current positions and the surface triangles (1-based)."""
        try:
            faces = self.extractSurfaceFaces()
            with open(filepath, "w") as f:
                f.write(f"# TetMesh OBJ Export\n# Vertices: {self.numVerts}\n# Tetrahedra: {self.numTets}\n\n")
                for v in self.mVertPos:
                    f.write(f"v {v[0]:g} {v[1]:g} {v[2]:g}\n")
                f.write(f"\n# Surface Triangles: {len(faces)}\n")
                for a, b, c in faces + 1:
                    f.write(f"f {a} {b} {c}\n")
        except OSError:
            return False
        return True


def mesh_from_generator(m, mu=None, lam=None) -> TetMesh:
    """TetMesh from a vbd_tetmesh_amd.meshgen.Mesh (per-tet materials carried over)."""
    t = TetMesh().initialize(m.rest, m.x0, m.density, m.tets, m.mu if mu is None else mu, m.lam if lam is None else lam,
                             np.nonzero(m.fixed)[0])
    if m.tet_mu is not None:
        t.tetMu = np.ascontiguousarray(m.tet_mu, np.float32)
        t.tetLambda = np.ascontiguousarray(m.tet_lambda, np.float32)
    if m.v0 is not None:
        t.mVelocity[:] = m.v0
    return t


def _colors_to_csr(colors) -> tuple:
    if isinstance(colors, tuple) and len(colors) == 2:
        off, verts = colors
        return np.ascontiguousarray(off, np.int32), np.ascontiguousarray(verts, np.int32)
    off = np.zeros(len(colors) + 1, np.int32)
    for i, c in enumerate(colors):
        off[i + 1] = off[i] + len(c)
    verts = np.concatenate([np.asarray(c, np.int32) for c in colors]) if len(colors) else np.zeros(0, np.int32)
    return off, np.ascontiguousarray(verts, np.int32)


def _mesh_arrays(positions, mass, is_fixed, tets, dminv, rest_volume, color_off, color_verts, inertia=None,
                 velocity=None, tet_mu=None, tet_lambda=None) -> dict:
    """Contiguous arrays of one mesh description, checked against each other's sizes: the C ABI
    takes plain pointers and reads num_verts / num_tets entries from each, so a short array
    would be read past its end."""
    a = dict(
        positions=np.ascontiguousarray(positions, np.float32).reshape(-1, 3),
        mass=np.ascontiguousarray(mass, np.float32).reshape(-1),
        is_fixed=np.ascontiguousarray(is_fixed, np.uint8).reshape(-1),
        tets=np.ascontiguousarray(tets, np.int32).reshape(-1, 4),
        dminv=np.ascontiguousarray(dminv, np.float32).reshape(-1, 9),
        rest_volume=np.ascontiguousarray(rest_volume, np.float32).reshape(-1),
        color_off=np.ascontiguousarray(color_off, np.int32).reshape(-1),
        color_verts=np.ascontiguousarray(color_verts, np.int32).reshape(-1),
    )
    nv, nt = a["positions"].shape[0], a["tets"].shape[0]
    for k, n in (("mass", nv), ("is_fixed", nv), ("color_verts", nv), ("dminv", nt), ("rest_volume", nt)):
        if a[k].shape[0] != n:
            raise ValueError(f"{k} has {a[k].shape[0]} entries, expected {n}")
    if a["color_off"].size < 1:
        raise ValueError("color_off needs num_colors + 1 >= 1 entries")
    for k, v in (("inertia", inertia), ("velocity", velocity)):
        if v is not None:
            a[k] = np.ascontiguousarray(v, np.float32).reshape(-1, 3)
            if a[k].shape[0] != nv:
                raise ValueError(f"{k} has {a[k].shape[0]} rows, expected {nv}")
    if (tet_mu is None) != (tet_lambda is None):
        raise ValueError("tet_mu and tet_lambda must both be given or both be None")
    if tet_mu is not None:
        a["tet_mu"] = np.ascontiguousarray(tet_mu, np.float32).reshape(-1)
        a["tet_lambda"] = np.ascontiguousarray(tet_lambda, np.float32).reshape(-1)
        for k in ("tet_mu", "tet_lambda"):
            if a[k].shape[0] != nt:
                raise ValueError(f"{k} has {a[k].shape[0]} entries, expected {nt}")
    return a


def _fill_desc(d, a: dict, slot_order: int):
    d.num_verts = a["positions"].shape[0]
    d.num_tets = a["tets"].shape[0]
    d.positions = ptr(a["positions"], C.c_float)
    d.inertia = ptr(a.get("inertia"), C.c_float)
    d.velocity = ptr(a.get("velocity"), C.c_float)
    d.mass = ptr(a["mass"], C.c_float)
    d.is_fixed = ptr(a["is_fixed"], C.c_uint8)
    d.tets = ptr(a["tets"], C.c_int32)
    d.dminv = ptr(a["dminv"], C.c_float)
    d.rest_volume = ptr(a["rest_volume"], C.c_float)
    d.tet_mu = ptr(a.get("tet_mu"), C.c_float)
    d.tet_lambda = ptr(a.get("tet_lambda"), C.c_float)
    d.num_colors = a["color_off"].size - 1
    d.color_offsets = ptr(a["color_off"], C.c_int32)
    d.color_verts = ptr(a["color_verts"], C.c_int32)
    d.slot_order = int(slot_order or 0)


class _Handle:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        check(lib.vbd_create(C.byref(h), device))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib.vbd_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DeviceMesh(_Handle):
    """One uploaded mesh (or batch) on the device; thin object wrapper of the C ABI."""

    def __init__(self, device: int = 0):
        super().__init__(device)
        self.num_verts = 0
        self.num_tets = 0
        self._keep = None

    # -- upload
    def init(self, positions, mass, is_fixed, tets, dminv, rest_volume, color_off, color_verts, inertia=None,
             velocity=None, tet_mu=None, tet_lambda=None, slot_order: int = 0):
        arrs = _mesh_arrays(positions, mass, is_fixed, tets, dminv, rest_volume, color_off, color_verts, inertia,
                            velocity, tet_mu, tet_lambda)
        d = A.MeshDesc()
        _fill_desc(d, arrs, slot_order)
        check(lib.vbd_init(self.h, C.byref(d)), self.h)
        self.num_verts = d.num_verts
        self.num_tets = d.num_tets
        return self

    def init_batch(self, descs: Sequence[dict]):
        """descs: per-mesh dicts with the keyword arguments of init() (local numbering)."""
        keep = []
        arr = (A.MeshDesc * len(descs))()
        nv = nt = 0
        for i, kw in enumerate(descs):
            fields = _mesh_arrays(**{k: v for k, v in kw.items() if k != "slot_order"})
            keep.append(fields)
            _fill_desc(arr[i], fields, kw.get("slot_order", 0))
            nv += arr[i].num_verts
            nt += arr[i].num_tets
        check(lib.vbd_init_batch(self.h, len(descs), arr), self.h)
        self.num_verts = nv
        self.num_tets = nt
        return self

    def info(self) -> A.Info:
        i = A.Info()
        check(lib.vbd_get_info(self.h, C.byref(i)), self.h)
        return i

    # -- transfers
    def _down(self, fn, out=None) -> np.ndarray:
        """Into `out` when it is a C-contiguous float32 (N, 3) array, else into a new one."""
        if not (isinstance(out, np.ndarray) and out.dtype == np.float32 and out.flags.c_contiguous
                and out.shape == (self.num_verts, 3) and out.flags.writeable):
            out = np.empty((self.num_verts, 3), np.float32)
        check(fn(self.h, ptr(out, C.c_float)), self.h)
        return out

    def _vec3(self, a, dtype=np.float32) -> np.ndarray:
        """(num_verts, 3) contiguous array of dtype; other row counts are rejected (the C ABI reads
        num_verts rows)."""
        a = np.ascontiguousarray(a, dtype).reshape(-1, 3)
        if a.shape[0] != self.num_verts:
            raise ValueError(f"expected {self.num_verts} rows of 3, got {a.shape[0]}")
        return a

    def _per_tet(self, a, name: str) -> np.ndarray:
        a = np.ascontiguousarray(a, np.float32).reshape(-1)
        if a.shape[0] != self.num_tets:
            raise ValueError(f"{name} has {a.shape[0]} entries, expected {self.num_tets}")
        return a

    def _up(self, fn, a):
        a = self._vec3(a)
        check(fn(self.h, ptr(a, C.c_float)), self.h)

    def positions(self, out=None) -> np.ndarray:
        return self._down(lib.vbd_download_positions, out)

    def velocities(self, out=None) -> np.ndarray:
        return self._down(lib.vbd_download_velocities, out)

    def inertia(self, out=None) -> np.ndarray:
        return self._down(lib.vbd_download_inertia, out)

    def previous_positions(self, out=None) -> np.ndarray:
        return self._down(lib.vbd_download_previous_positions, out)

    def set_positions(self, x):
        self._up(lib.vbd_upload_positions, x)

    def set_inertia(self, y):
        self._up(lib.vbd_upload_inertia, y)

    def set_velocities(self, v):
        self._up(lib.vbd_upload_velocities, v)

    def adjacency(self):
        info = self.info()
        off = np.zeros(info.num_verts + 1, np.int32)
        ent = np.zeros(max(4 * info.num_tets, 1), np.int32)
        check(lib.vbd_get_adjacency(self.h, ptr(off, C.c_int32), ptr(ent, C.c_int32)), self.h)
        return off, ent[: 4 * info.num_tets]

    def color_lists(self):
        info = self.info()
        off = np.zeros(info.num_colors + 1, np.int32)
        verts = np.zeros(max(info.num_verts, 1), np.int32)
        check(lib.vbd_get_color_lists(self.h, ptr(off, C.c_int32), ptr(verts, C.c_int32)), self.h)
        return off, verts[: info.num_verts]

    # -- solve
    def solve_iteration(self, dt_sqr_recip, mu, lam):
        check(lib.vbd_solve_iteration(self.h, dt_sqr_recip, mu, lam), self.h)

    def solve(self, iters, dt_sqr_recip, mu, lam, accel=True, rho=0.99, start_from_inertia=True):
        check(lib.vbd_solve(self.h, iters, dt_sqr_recip, mu, lam, int(bool(accel)), rho, int(bool(start_from_inertia))),
              self.h)

    def set_params(self, p: VBDParams):
        cp = p.to_c()
        check(lib.vbd_set_params(self.h, C.byref(cp)), self.h)
        self.set_material_model(p.material)

    def set_material_model(self, model: int):
        """VBD_MATERIAL_STABLE_NEO_HOOKEAN (the CPU solver's material) or VBD_MATERIAL_COROTATED_GPU
        (the reference GPU kernel's corotated model and regularised inertia term)."""
        check(lib.vbd_set_material_model(self.h, int(model)), self.h)

    def step(self, frames: int = 1):
        check(lib.vbd_step(self.h, frames), self.h)

    def forward_step(self):
        check(lib.vbd_forward_step(self.h), self.h)

    def update_velocity(self):
        check(lib.vbd_update_velocity(self.h), self.h)

    def synchronize(self):
        check(lib.vbd_synchronize(self.h), self.h)

    def set_stream(self, stream=None):
        """Run on a caller's stream: a torch.cuda.Stream, a raw hipStream_t (int), or None for the
        handle's own stream."""
        raw = getattr(stream, "cuda_stream", stream)
        check(lib.vbd_set_stream(self.h, C.c_void_p(raw) if raw else None), self.h)

    def set_persistent(self, mode: int):
        """1: one persistent launch per sweep (colors + blend, grid barriers) on the resident grid;
        8k: the same on at most 8k blocks; 0: one launch per color."""
        check(lib.vbd_set_persistent(self.h, int(mode)), self.h)

    def set_graph(self, enabled: bool):
        check(lib.vbd_set_graph_enabled(self.h, int(bool(enabled))), self.h)

    def set_launch_config(self, lanes: int, block: int = 256, pairs_per_lane: int = 1):
        check(lib.vbd_set_launch_config(self.h, lanes, block, pairs_per_lane), self.h)

    def autotune(self, iterations: int = 3):
        """Returns (lanes_per_vertex, pairs_per_lane, block_size, ms per sweep or -1 if cached)."""
        l, u, b, ms = C.c_int32(), C.c_int32(), C.c_int32(), C.c_float()
        check(lib.vbd_autotune(self.h, iterations, C.byref(l), C.byref(u), C.byref(b), C.byref(ms)), self.h)
        return l.value, u.value, b.value, ms.value

    def time_iterations(self, iterations: int):
        cms, bms = C.c_double(), C.c_double()
        cn, bn = C.c_int32(), C.c_int32()
        check(lib.vbd_time_iterations(self.h, iterations, C.byref(cms), C.byref(cn), C.byref(bms), C.byref(bn)), self.h)
        return cms.value, cn.value, bms.value, bn.value

    def time_color_launches(self, sweeps: int):
        """(ms per color launch, launches) of `sweeps` sweeps of the step's color kernels replayed
        from one hipGraph between two HIP events on the handle's stream.  Moves the positions."""
        ms, n = C.c_double(), C.c_int32()
        check(lib.vbd_time_color_launches(self.h, int(sweeps), C.byref(ms), C.byref(n)), self.h)
        return ms.value, n.value

    def layout_bytes(self):
        """(color-launch bytes per sweep, blend bytes per sweep) the current layout must move."""
        c, b = C.c_double(), C.c_double()
        check(lib.vbd_layout_bytes(self.h, C.byref(c), C.byref(b)), self.h)
        return c.value, b.value

    def algorithmic_bytes(self):
        cb, bb = C.c_double(), C.c_double()
        check(lib.vbd_algorithmic_bytes(self.h, C.byref(cb), C.byref(bb)), self.h)
        return cb.value, bb.value

    def total_energy(self, mu: float, lam: float) -> float:
        """VBDsolver::computeTotalEnergy's log-Neo-Hookean diagnostic (VBD.cpp:258-288)."""
        e = C.c_double()
        check(lib.vbd_total_energy(self.h, mu, lam, C.byref(e)), self.h)
        return e.value

    def incremental_potential(self) -> Tuple[float, float]:
        """(G, elastic part) of the incremental potential at the current positions and inertia."""
        g, e = C.c_double(), C.c_double()
        check(lib.vbd_incremental_potential(self.h, C.byref(g), C.byref(e)), self.h)
        return g.value, e.value

    def set_energy_log(self, every: int):
        """Record the diagnostic energy after every `every`-th iteration of each step (0: off)."""
        check(lib.vbd_set_energy_log(self.h, int(every)), self.h)

    def energy_log(self) -> np.ndarray:
        """The energies recorded by the last step (VBD.cpp:92-94)."""
        n = C.c_int32()
        check(lib.vbd_energy_log(self.h, None, 0, C.byref(n)), self.h)
        out = np.zeros(max(n.value, 1), np.float64)
        check(lib.vbd_energy_log(self.h, ptr(out, C.c_double), n.value, C.byref(n)), self.h)
        return out[:n.value]

    def time_energy(self, mode: int = 0, reps: int = 20) -> Tuple[float, float]:
        """(ms per evaluation, algorithmic bytes per evaluation) of the energy reduction:
        mode 0 the log-Neo-Hookean diagnostic, mode 1 the incremental potential."""
        ms, nb = C.c_double(), C.c_double()
        check(lib.vbd_time_energy(self.h, mode, reps, C.byref(ms), C.byref(nb)), self.h)
        return ms.value, nb.value

    # -- differentiation
    def set_differentiation_precision(self, bits: int):
        """32 (FP32 state) or 64 (an FP64 copy of the state for the differentiation entry points)."""
        check(lib.vbd_set_differentiation_precision(self.h, int(bits)), self.h)

    def state_f64(self) -> Tuple[np.ndarray, np.ndarray]:
        """(positions, velocities) of the FP64 state, (N, 3) float64 each (precision 64)."""
        x = np.zeros((self.num_verts, 3), np.float64)
        v = np.zeros((self.num_verts, 3), np.float64)
        check(lib.vbd_download_state_f64(self.h, ptr(x, C.c_double), ptr(v, C.c_double)), self.h)
        return x, v

    def set_state_f64(self, x=None, v=None):
        """Sets the FP64 state's positions and / or velocities (None keeps them)."""
        xa = None if x is None else self._vec3(x, np.float64)
        va = None if v is None else self._vec3(v, np.float64)
        check(lib.vbd_upload_state_f64(self.h, ptr(xa, C.c_double), ptr(va, C.c_double)), self.h)

    def set_materials(self, tet_mu, tet_lambda):
        mu = self._per_tet(tet_mu, "tet_mu")
        lam = self._per_tet(tet_lambda, "tet_lambda")
        check(lib.vbd_set_materials(self.h, ptr(mu, C.c_float), ptr(lam, C.c_float)), self.h)

    def set_converged_method(self, method: int):
        """VBD_CONVERGED_VBD (VBD sweeps, the default) or VBD_CONVERGED_NEWTON (Newton steps with
        conjugate-gradient solves, precision 64) for step_converged."""
        check(lib.vbd_set_converged_method(self.h, int(method)), self.h)

    def last_step_stats(self) -> Tuple[int, int, int]:
        """(Newton steps, CG iterations inside them, VBD sweeps) of the last step_converged."""
        a, b, c = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib.vbd_last_step_stats(self.h, C.byref(a), C.byref(b), C.byref(c)), self.h)
        return a.value, b.value, c.value

    def residual_norm(self):
        """(|grad G| at the current state, |M g|)."""
        r, l = C.c_double(), C.c_double()
        check(lib.vbd_residual_norm(self.h, C.byref(r), C.byref(l)), self.h)
        return r.value, l.value

    def step_converged(self, max_iterations: int, check_every: int = 10, abs_tol: float = 0.0, rel_tol: float = 1e-5,
                       stall_checks: int = 5):
        """One converged time step; returns (iterations, final |grad G|)."""
        it, nrm = C.c_int32(), C.c_double()
        check(lib.vbd_step_converged(self.h, max_iterations, check_every, abs_tol, rel_tol, stall_checks, C.byref(it),
                                     C.byref(nrm)), self.h)
        return it.value, nrm.value

    def hessian_vector(self, w) -> np.ndarray:
        w = self._vec3(w)
        out = np.zeros_like(w)
        check(lib.vbd_hessian_vector(self.h, ptr(w, C.c_float), ptr(out, C.c_float)), self.h)
        return out

    def adjoint_solve(self, rhs, w0=None, max_iterations: int = 5000, check_every: int = 10, rel_tol: float = 1e-6,
                      rho: float = 0.0, stall_checks: int = 0):
        """Solve (M/h^2 + K) w = rhs; returns (w, sweeps, relative residual)."""
        rhs = self._vec3(rhs)
        w = np.zeros_like(rhs) if w0 is None else self._vec3(w0).copy()
        it, rr = C.c_int32(), C.c_double()
        check(lib.vbd_adjoint_solve(self.h, ptr(rhs, C.c_float), ptr(w, C.c_float), max_iterations, check_every,
                                    rel_tol, rho, stall_checks, C.byref(it), C.byref(rr)), self.h)
        return w, it.value, rr.value

    def hessian_vector_f64(self, w) -> np.ndarray:
        """A w with double vectors (precision 64)."""
        w = self._vec3(w, np.float64)
        out = np.zeros_like(w)
        check(lib.vbd_hessian_vector_f64(self.h, ptr(w, C.c_double), ptr(out, C.c_double)), self.h)
        return out

    def adjoint_solve_f64(self, rhs, w0=None, max_iterations: int = 5000, check_every: int = 10,
                          rel_tol: float = 1e-10, rho: float = 0.0, stall_checks: int = 0):
        """adjoint_solve with double vectors (precision 64); returns (w, sweeps, relative residual)."""
        rhs = self._vec3(rhs, np.float64)
        w = np.zeros_like(rhs) if w0 is None else self._vec3(w0, np.float64).copy()
        it, rr = C.c_int32(), C.c_double()
        check(lib.vbd_adjoint_solve_f64(self.h, ptr(rhs, C.c_double), ptr(w, C.c_double), max_iterations, check_every,
                                        rel_tol, rho, stall_checks, C.byref(it), C.byref(rr)), self.h)
        return w, it.value, rr.value

    def adjoint_solve_cg_f64(self, rhs, w0=None, max_iterations: int = 5000, check_every: int = 10,
                             rel_tol: float = 1e-10):
        """(M/h^2 + K) w = rhs by block-Jacobi preconditioned conjugate gradients on the FP64 state
        (precision 64); returns (w, iterations, relative residual of the returned w)."""
        rhs = self._vec3(rhs, np.float64)
        w = np.zeros_like(rhs) if w0 is None else self._vec3(w0, np.float64).copy()
        it, rr = C.c_int32(), C.c_double()
        check(lib.vbd_adjoint_solve_cg_f64(self.h, ptr(rhs, C.c_double), ptr(w, C.c_double), max_iterations,
                                           check_every, rel_tol, C.byref(it), C.byref(rr)), self.h)
        return w, it.value, rr.value

    def parameter_gradients_f64(self, w):
        """parameter_gradients with double w and outputs (precision 64)."""
        w = self._vec3(w, np.float64)
        nt = self.info().num_tets
        dmu = np.zeros(max(nt, 1), np.float64)
        dlam = np.zeros(max(nt, 1), np.float64)
        check(lib.vbd_parameter_gradients_f64(self.h, ptr(w, C.c_double), ptr(dmu, C.c_double),
                                              ptr(dlam, C.c_double)), self.h)
        return dmu[:nt], dlam[:nt]

    def parameter_gradients(self, w):
        w = self._vec3(w)
        nt = self.info().num_tets
        dmu = np.zeros(max(nt, 1), np.float32)
        dlam = np.zeros(max(nt, 1), np.float32)
        check(lib.vbd_parameter_gradients(self.h, ptr(w, C.c_float), ptr(dmu, C.c_float), ptr(dlam, C.c_float)),
              self.h)
        return dmu[:nt], dlam[:nt]


class VBD_GPU_Solver:
    """VBD_GPU_Solver (VBD_GPU_Wrapper.h:7-45) on the gfx950 C ABI."""

    def __init__(self, device: int = 0, material: int = A.VBD_MATERIAL_STABLE_NEO_HOOKEAN):
        """material=A.VBD_MATERIAL_COROTATED_GPU runs the reference GPU kernel's corotated model
        (VBD_GPU_Wrapper.cu:181); the default is the CPU solver's stable Neo-Hookean."""
        self._dev = DeviceMesh(device)
        self._material = material

    @property
    def device_mesh(self) -> DeviceMesh:
        return self._dev

    def initFromCPU(self, mesh: TetMesh, vertexColors, tetDmInv, tetRestVol):
        """VBD_GPU_Wrapper.cu:246-315.  Per-tet materials of `mesh` are not uploaded (the
        reference solve takes scalar mu, lambda per call)."""
        off, verts = _colors_to_csr(vertexColors)
        self._dev.init(mesh.mVertPos, mesh.vertexMass, mesh.isFixed, mesh.tets, tetDmInv, tetRestVol, off, verts,
                       inertia=mesh.mInertia)
        self._dev.set_material_model(self._material)

    def solveIteration(self, dtSqrReciprocal, mu, lambda_):
        """VBD_GPU_Wrapper.cu:317-329: one sweep over all colors."""
        self._dev.solve_iteration(dtSqrReciprocal, mu, lambda_)

    def solveVBD(self, numIterations, dtSqrReciprocal, mu, lambda_, useAcceleration=True, rho=0.99):
        """VBD_GPU_Wrapper.cu:332-362: x <- y, then iterations with Chebyshev."""
        self._dev.solve(numIterations, dtSqrReciprocal, mu, lambda_, useAcceleration, rho, start_from_inertia=True)

    def copyPositionsToCPU(self, mesh: TetMesh):
        """VBD_GPU_Wrapper.cu:364-366."""
        got = self._dev.positions(out=mesh.mVertPos)
        if got is not mesh.mVertPos:
            mesh.mVertPos[:] = got

    def updateInertiaFromCPU(self, mesh: TetMesh):
        """VBD_GPU_Wrapper.cu:369-371."""
        self._dev.set_inertia(mesh.mInertia)


class VBDsolver:
    """VBDsolver (VBD.h:33-140) whose step runs entirely on the device.

    ``step(mesh)`` uploads the mesh state on the first call (and whenever the host arrays
    were replaced), runs forwardStep / iterations / updateVelocity as one captured hipGraph,
    then downloads positions, velocities and the inertia target back into ``mesh`` so the
    host view matches the reference after every call.  ``step_device(n)`` advances n frames
    without any host transfer.
    """

    def __init__(self, device: int = 0):
        self.params = VBDParams()
        self.vertexColors: List[np.ndarray] = []
        self.colorOffsets = None
        self.colorVerts = None
        self.tetDmInv = None
        self.tetRestVolume = None
        self.energyHistory: List[float] = []
        self._device = device
        self._dev: Optional[DeviceMesh] = None
        self._uploaded_ids = None

    def setParams(self, params: VBDParams):
        self.params = params
        if self._dev is not None:
            self._dev.set_params(self._device_params())
            self._dev.set_energy_log(params.energyLogEvery)

    def initializeColoring(self, mesh: TetMesh):
        """VBD.cpp:23-34 (greedy coloring on the host, GraphColoring.cpp:28-66)."""
        _, off, verts = A.color_vertices(mesh.numVerts, mesh.tets, mesh.adjOffsets, mesh.adjTets)
        self.colorOffsets, self.colorVerts = off, verts
        self.vertexColors = [verts[off[c]: off[c + 1]] for c in range(off.size - 1)]

    def precomputeTetData(self, mesh: TetMesh):
        """VBD.cpp:36-58."""
        self.tetDmInv, self.tetRestVolume = A.precompute_tet_data(mesh.mRestPos, mesh.tets)

    def getVertexColors(self):
        return self.vertexColors

    def getTetDmInv(self):
        return self.tetDmInv

    def getTetRestVol(self):
        return self.tetRestVolume

    def _device_params(self) -> VBDParams:
        return self.params

    def _ensure_device(self, mesh: TetMesh, per_tet: bool):
        ids = (id(mesh.mVertPos), id(mesh.mVelocity), mesh.numVerts)
        if self._dev is None:
            self._dev = DeviceMesh(self._device)
            mu = mesh.tetMu if per_tet else None
            lam = mesh.tetLambda if per_tet else None
            self._dev.init(mesh.mVertPos, mesh.vertexMass, mesh.isFixed, mesh.tets, self.tetDmInv, self.tetRestVolume,
                           self.colorOffsets, self.colorVerts, inertia=mesh.mInertia, velocity=mesh.mVelocity,
                           tet_mu=mu, tet_lambda=lam)
            self._dev.set_params(self._device_params())
            self._dev.set_energy_log(self.params.energyLogEvery)
            self._uploaded_ids = ids

    def step(self, mesh: TetMesh, per_tet_material: bool = False):
        """VBD.cpp:60-101."""
        if not self.vertexColors:
            self.initializeColoring(mesh)
            self.precomputeTetData(mesh)
        self._ensure_device(mesh, per_tet_material)
        self._dev.set_positions(mesh.mVertPos)
        self._dev.set_velocities(mesh.mVelocity)
        self._dev.step(1)
        self.syncToHost(mesh)
        if self.params.energyLogEvery > 0:  # VBD.cpp:92-94
            self.energyHistory.extend(float(e) for e in self._dev.energy_log())

    def syncToHost(self, mesh: TetMesh):
        for dst, fn in ((mesh.mVertPos, self._dev.positions), (mesh.mVelocity, self._dev.velocities),
                        (mesh.mInertia, self._dev.inertia), (mesh.mVertPrevPos, self._dev.previous_positions)):
            got = fn(out=dst)  # straight into the mesh's array when it is float32 (N, 3)
            if got is not dst:
                dst[:] = got

    def step_device(self, frames: int = 1):
        self._dev.step(frames)

    def simulate(self, mesh: TetMesh, numFrames: int, callback: Optional[Callable] = None,
                 per_tet_material: bool = False):
        """VBD.cpp:313-330."""
        self.energyHistory.clear()
        for f in range(numFrames):
            self.step(mesh, per_tet_material)
            if callback is not None:
                callback(f, mesh)

    def computeTotalEnergy(self, mesh: TetMesh) -> float:
        """VBD.cpp:258-288 as a device reduction (|det F| for the singular-value product)."""
        if self._dev is None:
            self._ensure_device(mesh, False)
        self._dev.set_positions(mesh.mVertPos)
        return self._dev.total_energy(self.params.mu, self.params.lambda_)

    @property
    def device_mesh(self) -> Optional[DeviceMesh]:
        return self._dev


def prepare_host(m, mu=None, lam=None):
    """Mesh generator output -> (TetMesh, color offsets, color verts, DmInv, V0) using the
    product's host C++ helpers (the same preparation VBDsolver does before its first step)."""
    t = mesh_from_generator(m, mu, lam)
    _, off, verts = A.color_vertices(t.numVerts, t.tets, t.adjOffsets, t.adjTets)
    dminv, vol = A.precompute_tet_data(t.mRestPos, t.tets)
    return t, off, verts, dminv, vol


def upload(m, device: int = 0, params: Optional[VBDParams] = None, per_tet: Optional[bool] = None,
           slot_order: int = 0) -> DeviceMesh:
    """Generator mesh -> DeviceMesh with params set (host prep by the product C++ helpers)."""
    t, off, verts, dminv, vol = prepare_host(m)
    use_mat = (m.tet_mu is not None) if per_tet is None else per_tet
    d = DeviceMesh(device).init(t.mVertPos, t.vertexMass, t.isFixed, t.tets, dminv, vol, off, verts,
                                velocity=t.mVelocity, tet_mu=t.tetMu if use_mat else None,
                                tet_lambda=t.tetLambda if use_mat else None, slot_order=slot_order)
    if params is not None:
        d.set_params(params)
    return d
