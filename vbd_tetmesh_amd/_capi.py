"""ctypes binding of the C ABI declared in include/vbd_capi.h (libvbd_hip.so).

The library is the only compute path: if it is missing or fails to load, importing this
module raises.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvbd_hip.so")

VBD_OK, VBD_ERR_ARG, VBD_ERR_HIP, VBD_ERR_STATE, VBD_ERR_ALLOC, VBD_ERR_NUMERIC = range(6)

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
i32p = C.POINTER(C.c_int32)
u8p = C.POINTER(C.c_uint8)


class VBDError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"vbd error {code}: {msg}")
        self.code = code


class MeshDesc(C.Structure):
    _fields_ = [
        ("num_verts", C.c_int32),
        ("num_tets", C.c_int32),
        ("positions", f32p),
        ("inertia", f32p),
        ("velocity", f32p),
        ("mass", f32p),
        ("is_fixed", u8p),
        ("tets", i32p),
        ("dminv", f32p),
        ("rest_volume", f32p),
        ("tet_mu", f32p),
        ("tet_lambda", f32p),
        ("num_colors", C.c_int32),
        ("color_offsets", i32p),
        ("color_verts", i32p),
        ("slot_order", C.c_int32),
    ]


VBD_MATERIAL_STABLE_NEO_HOOKEAN = 0
VBD_MATERIAL_COROTATED_GPU = 1
VBD_CONVERGED_VBD = 0
VBD_CONVERGED_NEWTON = 1


class StepParams(C.Structure):
    _fields_ = [
        ("dt", C.c_float),
        ("num_iterations", C.c_int32),
        ("substeps", C.c_int32),
        ("gravity", C.c_float * 3),
        ("use_acceleration", C.c_int32),
        ("acceleration_rho", C.c_float),
        ("mu", C.c_float),
        ("lambda_", C.c_float),
        ("velocity_damping", C.c_float),
        ("start_from_inertia", C.c_int32),
        ("contact_on", C.c_int32),
        ("contact_axis", C.c_int32),
        ("contact_k", C.c_float),
        ("contact_plane", C.c_float),
    ]


class Info(C.Structure):
    _fields_ = [
        ("num_verts", C.c_int32),
        ("num_tets", C.c_int32),
        ("num_free", C.c_int32),
        ("num_colors", C.c_int32),
        ("num_pairs", C.c_int64),
        ("max_valence", C.c_int32),
        ("per_tet_material", C.c_int32),
        ("lanes_per_vertex", C.c_int32),
        ("pairs_per_lane", C.c_int32),
        ("block_size", C.c_int32),
        ("device_bytes", C.c_int64),
        ("cached_graphs", C.c_int32),
        ("full_records", C.c_int32),
        ("persistent_grid", C.c_int32),
    ]


# (name, restype, argtypes) of every exported symbol declared in include/vbd_capi.h
H = C.c_void_p
SIGNATURES = [
    ("vbd_capi_version", C.c_int32, []),
    ("vbd_global_last_error", C.c_char_p, []),
    ("vbd_last_error", C.c_char_p, [H]),
    ("vbd_create", C.c_int, [C.POINTER(C.c_void_p), C.c_int32]),
    ("vbd_destroy", None, [H]),
    ("vbd_init", C.c_int, [H, C.POINTER(MeshDesc)]),
    ("vbd_init_batch", C.c_int, [H, C.c_int32, C.POINTER(MeshDesc)]),
    ("vbd_upload_positions", C.c_int, [H, f32p]),
    ("vbd_upload_inertia", C.c_int, [H, f32p]),
    ("vbd_upload_velocities", C.c_int, [H, f32p]),
    ("vbd_download_positions", C.c_int, [H, f32p]),
    ("vbd_download_velocities", C.c_int, [H, f32p]),
    ("vbd_download_inertia", C.c_int, [H, f32p]),
    ("vbd_download_previous_positions", C.c_int, [H, f32p]),
    ("vbd_get_info", C.c_int, [H, C.POINTER(Info)]),
    ("vbd_get_adjacency", C.c_int, [H, i32p, i32p]),
    ("vbd_get_color_lists", C.c_int, [H, i32p, i32p]),
    ("vbd_solve_iteration", C.c_int, [H, C.c_float, C.c_float, C.c_float]),
    ("vbd_solve", C.c_int, [H, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_int32, C.c_float, C.c_int32]),
    ("vbd_set_params", C.c_int, [H, C.POINTER(StepParams)]),
    ("vbd_forward_step", C.c_int, [H]),
    ("vbd_update_velocity", C.c_int, [H]),
    ("vbd_step", C.c_int, [H, C.c_int32]),
    ("vbd_synchronize", C.c_int, [H]),
    ("vbd_set_graph_enabled", C.c_int, [H, C.c_int32]),
    ("vbd_set_persistent", C.c_int, [H, C.c_int32]),
    ("vbd_set_launch_config", C.c_int, [H, C.c_int32, C.c_int32, C.c_int32]),
    ("vbd_autotune", C.c_int, [H, C.c_int32, i32p, i32p, i32p, f32p]),
    ("vbd_time_iterations", C.c_int, [H, C.c_int32, f64p, i32p, f64p, i32p]),
    ("vbd_algorithmic_bytes", C.c_int, [H, f64p, f64p]),
    ("vbd_time_color_launches", C.c_int, [H, C.c_int32, f64p, i32p]),
    ("vbd_layout_bytes", C.c_int, [H, f64p, f64p]),
    ("vbd_mesh_initialize", C.c_int, [C.c_int32, f32p, C.c_int32, i32p, C.c_float, f32p, f32p, i32p, i32p]),
    ("vbd_color_vertices", C.c_int, [C.c_int32, i32p, i32p, i32p, i32p, i32p, i32p, i32p]),
    ("vbd_precompute_tet_data", C.c_int, [f32p, C.c_int32, i32p, f32p, f32p]),
    ("vbd_accelerator_omega", C.c_float, [C.c_int32, C.c_float, C.c_float]),
    ("vbd_total_energy", C.c_int, [H, C.c_float, C.c_float, f64p]),
    ("vbd_total_energy_noalloc", C.c_int, [H, C.c_float, C.c_float, C.c_void_p]),
    ("vbd_incremental_potential", C.c_int, [H, f64p, f64p]),
    ("vbd_time_energy", C.c_int, [H, C.c_int32, C.c_int32, f64p, f64p]),
    ("vbd_set_energy_log", C.c_int, [H, C.c_int32]),
    ("vbd_energy_log", C.c_int, [H, f64p, C.c_int32, i32p]),
    ("vbd_set_material_model", C.c_int, [H, C.c_int32]),
    ("vbd_set_stream", C.c_int, [H, C.c_void_p]),
    ("vbd_set_materials", C.c_int, [H, f32p, f32p]),
    ("vbd_residual_norm", C.c_int, [H, f64p, f64p]),
    ("vbd_step_converged", C.c_int, [H, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, i32p, f64p]),
    ("vbd_hessian_vector", C.c_int, [H, f32p, f32p]),
    ("vbd_adjoint_solve", C.c_int, [H, f32p, f32p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, i32p,
                                    f64p]),
    ("vbd_parameter_gradients", C.c_int, [H, f32p, f32p, f32p]),
    ("vbd_set_differentiation_precision", C.c_int, [H, C.c_int32]),
    ("vbd_download_state_f64", C.c_int, [H, f64p, f64p]),
    ("vbd_upload_state_f64", C.c_int, [H, f64p, f64p]),
    ("vbd_hessian_vector_f64", C.c_int, [H, f64p, f64p]),
    ("vbd_adjoint_solve_f64", C.c_int, [H, f64p, f64p, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, i32p,
                                        f64p]),
    ("vbd_parameter_gradients_f64", C.c_int, [H, f64p, f64p, f64p]),
    ("vbd_adjoint_solve_cg_f64", C.c_int, [H, f64p, f64p, C.c_int32, C.c_int32, C.c_float, i32p, f64p]),
    ("vbd_set_converged_method", C.c_int, [H, C.c_int32]),
    ("vbd_last_step_stats", C.c_int, [H, i32p, i32p, i32p]),
]


def _share_torch_hip_runtime():
    """PyTorch wheels bundle their own HIP/HSA runtime (same sonames as the system ROCm).  Two
    runtimes in one process cannot both open the GPU, so when torch is installed its runtime is
    loaded first (RTLD_GLOBAL) and libvbd_hip.so binds to it; torch and this library then share
    one runtime whichever is imported first.  VBD_SYSTEM_HIP=1 keeps the system runtime."""
    if os.environ.get("VBD_SYSTEM_HIP") == "1":
        return None
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return None
    path = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if not os.path.exists(path):
        return None
    return C.CDLL(path, mode=C.RTLD_GLOBAL)


_runtime = _share_torch_hip_runtime()


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run __graft_entry__.build() (hipcc --offload-arch=gfx950)")
    lib = C.CDLL(LIB_PATH)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def ptr(a, ct):
    if a is None:
        return None
    return a.ctypes.data_as(C.POINTER(ct))


def check(rc: int, handle=None):
    if rc != VBD_OK:
        if handle is not None:
            msg = lib.vbd_last_error(handle)
        else:
            msg = lib.vbd_global_last_error()
        raise VBDError(rc, (msg or b"").decode(errors="replace"))


# ---- host helpers (product C++ code behind the ABI) -------------------------------------

def mesh_initialize(rest: np.ndarray, tets: np.ndarray, density: float):
    """TetMesh::initialize (TetMesh.h:213-275): (mass, tetRestVol, adj_offsets, adj_tets)."""
    rest = np.ascontiguousarray(rest, np.float32)
    tets = np.ascontiguousarray(tets, np.int32)
    nV, nT = rest.shape[0], tets.shape[0]
    mass = np.zeros(nV, np.float32)
    vol = np.zeros(nT, np.float32)
    off = np.zeros(nV + 1, np.int32)
    adj = np.zeros(max(4 * nT, 1), np.int32)
    check(lib.vbd_mesh_initialize(nV, ptr(rest, C.c_float), nT, ptr(tets, C.c_int32), density, ptr(mass, C.c_float),
                                  ptr(vol, C.c_float), ptr(off, C.c_int32), ptr(adj, C.c_int32)))
    return mass, vol, off, adj[: 4 * nT]


def color_vertices(nV: int, tets: np.ndarray, adj_off: np.ndarray, adj: np.ndarray):
    """GraphColoring::colorVertices (GraphColoring.cpp:28-66): (color_of, color_offsets, color_verts)."""
    tets = np.ascontiguousarray(tets, np.int32)
    adj_off = np.ascontiguousarray(adj_off, np.int32)
    adj = np.ascontiguousarray(adj, np.int32)
    color_of = np.zeros(max(nV, 1), np.int32)
    off = np.zeros(nV + 1, np.int32)
    verts = np.zeros(max(nV, 1), np.int32)
    nc = C.c_int32(0)
    check(lib.vbd_color_vertices(nV, ptr(tets, C.c_int32), ptr(adj_off, C.c_int32), ptr(adj, C.c_int32),
                                 ptr(color_of, C.c_int32), ptr(off, C.c_int32), ptr(verts, C.c_int32), C.byref(nc)))
    return color_of[:nV], off[: nc.value + 1].copy(), verts[:nV]


def precompute_tet_data(rest: np.ndarray, tets: np.ndarray):
    """VBDsolver::precomputeTetData (VBD.cpp:36-58): (DmInv row-major (T,9), V0 (T,))."""
    rest = np.ascontiguousarray(rest, np.float32)
    tets = np.ascontiguousarray(tets, np.int32)
    nT = tets.shape[0]
    d = np.zeros((max(nT, 1), 9), np.float32)
    v = np.zeros(max(nT, 1), np.float32)
    check(lib.vbd_precompute_tet_data(ptr(rest, C.c_float), nT, ptr(tets, C.c_int32), ptr(d, C.c_float),
                                      ptr(v, C.c_float)))
    return d[:nT], v[:nT]


def accelerator_omega(order: int, rho: float, prev: float) -> float:
    """VBDsolver::getAcceleratorOmega (VBD.cpp:302-311)."""
    return float(lib.vbd_accelerator_omega(order, rho, prev))
