"""ctypes binding of liboe.so (the C ABI declared in include/octopus_engine.h).

The library is loaded from the package directory.  There is no fallback: if
liboe.so is missing or fails to load, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(os.environ.get("OE_LIB", Path(__file__).resolve().parent / "liboe.so"))

if not LIB_PATH.exists():
    raise ImportError(f"{LIB_PATH} not found: build it with `python octopusengine_amd/build.py`")

lib = C.CDLL(str(LIB_PATH))

# status codes
OE_OK, OE_ERR_INVALID, OE_ERR_HIP, OE_ERR_BUDGET, OE_ERR_STATE, OE_ERR_INTERNAL = range(6)
# element / material / solver ids (reference enum values)
TETRA, HEXA, TETRA10, TETRA20, HEXA27 = 3, 6, 7, 8, 9
NEOHOOKE, STABLE_NEOHOOKE = 2, 3
MASS_UNIFORM, MASS_FEMSHAPE = 0, 1
SOLVER_VBD, SOLVER_XPBD, SOLVER_EXPLICIT = 0, 1, 2
VBD_BASE, VBD_THREADED_QUADRATURE, VBD_REDUCTION_SYMMETRY, VBD_BETTER_COLORATION = 0, 1, 2, 3
XPBD_XPBD, XPBD_QUADRATURE, XPBD_OPTISHARED, XPBD_OPTIGROUP = 0, 1, 2, 3

NODES = {TETRA: 4, HEXA: 8, TETRA10: 10, TETRA20: 20, HEXA27: 27}


class OEError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"[oe status {status}] {msg}")
        self.status = status


class SceneDesc(C.Structure):
    _fields_ = [
        ("element", C.c_int32),
        ("n_vertices", C.c_int32),
        ("xyz", C.POINTER(C.c_float)),
        ("n_elements", C.c_int32),
        ("topology", C.POINTER(C.c_int32)),
        ("material", C.c_int32),
        ("young", C.c_float),
        ("poisson", C.c_float),
        ("density", C.c_float),
        ("mass_distribution", C.c_int32),
        ("solver", C.c_int32),
        ("iterations", C.c_int32),
        ("rho", C.c_float),
        ("damping", C.c_float),
        ("global_damping", C.c_float),
        ("vbd_version", C.c_int32),
        ("target_color", C.c_int32),
        ("xpbd_version", C.c_int32),
        ("gravity", C.c_float * 3),
        ("shuffle_seed", C.c_uint64),
    ]


class Tile(C.Structure):
    """oe_tile: the full colour-pass layout."""
    _fields_ = [
        ("lanes", C.c_int32),
        ("qpl", C.c_int32),
        ("mode", C.c_int32),
        ("inc_per_block", C.c_int32),
        ("latency_blocks", C.c_int64),
    ]

    def as_dict(self):
        return {k: int(getattr(self, k)) for k, _ in self._fields_}


P = C.c_void_p
f32p = C.POINTER(C.c_float)
i32p = C.POINTER(C.c_int32)
f64p = C.POINTER(C.c_double)
i64p = C.POINTER(C.c_int64)
u64p = C.POINTER(C.c_uint64)

_SIGS = {
    "oe_scene_desc_default": (None, [C.POINTER(SceneDesc)]),
    "oe_last_error": (C.c_char_p, []),
    "oe_version": (C.c_char_p, []),
    "oe_engine_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(P)]),
    "oe_engine_destroy": (C.c_int32, [P]),
    "oe_engine_synchronize": (C.c_int32, [P]),
    "oe_engine_memory": (C.c_int32, [P, u64p, u64p]),
    "oe_scene_create": (C.c_int32, [P, C.POINTER(SceneDesc), C.POINTER(P)]),
    "oe_scene_destroy": (C.c_int32, [P]),
    "oe_scene_fix_vertices": (C.c_int32, [P, i32p, C.c_int32]),
    "oe_scene_fix_plane": (C.c_int32, [P, f32p, f32p, i32p]),
    "oe_scene_box_limit": (C.c_int32, [P, f32p, f32p]),
    "oe_scene_step": (C.c_int32, [P, C.c_float, C.c_int32]),
    "oe_scene_set_graph": (C.c_int32, [P, C.c_int32]),
    "oe_scene_set_tile": (C.c_int32, [P, C.c_int32, C.c_int32, C.c_int32]),
    "oe_scene_get_tile": (C.c_int32, [P, i32p, i32p, i32p]),
    "oe_scene_get_layout": (C.c_int32, [P, i32p]),
    "oe_scene_info": (C.c_int32, [P, i32p, i32p, i32p, i32p, i32p, u64p]),
    "oe_scene_get_positions": (C.c_int32, [P, f32p]),
    "oe_scene_get_last_positions": (C.c_int32, [P, f32p]),
    "oe_scene_get_init_positions": (C.c_int32, [P, f32p]),
    "oe_scene_get_velocities": (C.c_int32, [P, f32p]),
    "oe_scene_get_forces": (C.c_int32, [P, f32p]),
    "oe_scene_get_masses": (C.c_int32, [P, f32p]),
    "oe_scene_get_masks": (C.c_int32, [P, i32p]),
    "oe_scene_get_colors": (C.c_int32, [P, i32p]),
    "oe_scene_set_state": (C.c_int32, [P, f32p, f32p, f32p]),
    "oe_scene_set_timing": (C.c_int32, [P, C.c_int32]),
    "oe_scene_get_timing": (C.c_int32, [P, f64p, i64p]),
    "oe_batch_create": (C.c_int32, [P, C.POINTER(P), C.c_int32, C.POINTER(P)]),
    "oe_batch_destroy": (C.c_int32, [P]),
    "oe_batch_step": (C.c_int32, [P, C.c_float, C.c_int32]),
    "oe_batch_set_graph": (C.c_int32, [P, C.c_int32]),
    "oe_batch_set_streams": (C.c_int32, [P, C.c_int32]),
    "oe_batch_set_tile": (C.c_int32, [P, C.c_int32, C.c_int32, C.c_int32, C.c_int32]),
    "oe_batch_set_timing": (C.c_int32, [P, C.c_int32]),
    "oe_batch_get_timing": (C.c_int32, [P, f64p, i64p]),
    "oe_batch_info": (C.c_int32, [P, i32p, i64p, f64p]),
    "oe_batch_get_timing_element": (C.c_int32, [P, C.c_int32, f64p, i64p]),
    "oe_batch_set_tile_ex": (C.c_int32, [P, C.c_int32, C.POINTER(Tile)]),
    "oe_batch_get_tile_ex": (C.c_int32, [P, C.c_int32, C.POINTER(Tile)]),
    "oe_batch_autotune": (C.c_int32, [P, C.c_int32]),
    "oe_scene_set_tile_ex": (C.c_int32, [P, C.POINTER(Tile)]),
    "oe_scene_get_tile_ex": (C.c_int32, [P, C.POINTER(Tile)]),
    "oe_scene_autotune": (C.c_int32, [P, C.c_int32, C.POINTER(Tile), f64p, i32p]),
    "oe_engine_set_autotune": (C.c_int32, [P, C.c_int32]),
    "oe_autotune_cache": (C.c_int32, [i32p, i32p, i32p, i32p, i32p, C.POINTER(Tile), f64p, i32p]),
    "oe_autotune_clear": (C.c_int32, []),
    "oe_batch_info_element": (C.c_int32, [P, C.c_int32, i32p, i64p, f64p]),
    "oe_server_create": (C.c_int32, [P, C.c_int32, C.c_int32, C.POINTER(P)]),
    "oe_server_destroy": (C.c_int32, [P]),
    "oe_server_add_scene": (C.c_int32, [P, P, i32p]),
    "oe_server_set_tile_ex": (C.c_int32, [P, C.c_int32, C.POINTER(Tile)]),
    "oe_server_get_tile_ex": (C.c_int32, [P, C.c_int32, C.POINTER(Tile)]),
    "oe_server_submit": (C.c_int32, [P, C.c_int32, C.c_float, C.c_int32, u64p]),
    "oe_server_drain": (C.c_int32, [P]),
    "oe_server_wait": (C.c_int32, [P, C.c_uint64]),
    "oe_server_poll": (C.c_int32, [P, C.c_uint64, i32p]),
    "oe_server_stats": (C.c_int32, [P, i64p, f64p, f64p, f64p, i64p, i64p, f64p]),
    "oe_server_latencies": (C.c_int32, [P, f64p, C.c_int64, i64p]),
    "oe_server_latency_slots": (C.c_int32, [P, i32p, C.c_int64, i64p]),
    "oe_host_beam_mesh": (C.c_int32, [C.c_int32, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_float,
                                      i32p, i32p, f32p, i32p]),
    "oe_host_shape": (C.c_int32, [C.c_int32, i32p, i32p, f64p, f64p, f64p]),
    "oe_host_color_primal": (C.c_int32, [C.c_int32, i32p, C.c_int32, C.c_int32, C.c_int32, i32p, i32p, i64p]),
    "oe_host_color_dual": (C.c_int32, [C.c_int32, i32p, C.c_int32, C.c_int32, i32p, i32p, i64p]),
    "oe_host_fem_constants": (C.c_int32, [C.c_int32, f32p, C.c_int32, i32p, C.c_int32, f32p]),
    "oe_host_masses": (C.c_int32, [C.c_int32, f32p, C.c_int32, i32p, C.c_int32, C.c_float, C.c_int32, f32p]),
    "oe_scene_element_diagnostics": (C.c_int32, [P, f32p, f32p, f32p, f32p]),
    "oe_scene_residual": (C.c_int32, [P, C.c_float, f32p]),
    "oe_scene_residual_scratch": (C.c_int32, [P, C.c_float, f32p]),
    "oe_mesh_load": (C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "oe_mesh_info": (C.c_int32, [P, i32p, i32p]),
    "oe_mesh_vertices": (C.c_int32, [P, f32p]),
    "oe_mesh_topology": (C.c_int32, [P, C.c_int32, i32p]),
    "oe_mesh_destroy": (C.c_int32, [P]),
    "oe_vtk_open": (C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "oe_vtk_mesh": (C.c_int32, [P, C.c_int32, f32p, C.c_int32, i32p, C.c_int32]),
    "oe_vtk_point_data": (C.c_int32, [P]),
    "oe_vtk_cell_data": (C.c_int32, [P]),
    "oe_vtk_scalars": (C.c_int32, [P, C.c_char_p, f32p, C.c_int32]),
    "oe_vtk_vectors": (C.c_int32, [P, C.c_char_p, f32p, C.c_int32]),
    "oe_vtk_close": (C.c_int32, [P]),
    "oe_scene_save_vtk": (C.c_int32, [P, C.c_char_p]),
    "oe_record_open": (C.c_int32, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "oe_record_time": (C.c_int32, [P, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float]),
    "oe_record_mesh": (C.c_int32, [P, i32p, C.c_int32]),
    "oe_record_material": (C.c_int32, [P, C.c_int32, C.c_int32, C.c_float, C.c_float, C.c_int32, C.c_int32,
                                       C.c_float]),
    "oe_record_string": (C.c_int32, [P, C.c_char_p, C.c_char_p]),
    "oe_record_flexion": (C.c_int32, [P, C.c_float, f32p, f32p]),
    "oe_record_rotation": (C.c_int32, [P, C.c_float, C.c_float, f32p, f32p, C.c_int32]),
    "oe_record_text": (C.c_int32, [P, C.c_char_p, C.c_int64, C.POINTER(C.c_int64)]),
    "oe_record_close": (C.c_int32, [P]),
    "oe_flexion_error": (C.c_int32, [f32p, f32p, C.c_int32, f32p, f32p, f32p, f32p, C.POINTER(C.c_int32)]),
    "oe_rotation_error": (C.c_int32, [C.c_int32, i32p, C.c_int32, f32p, f32p, C.c_int32, C.c_int32, C.c_float,
                                      f32p, f32p, C.POINTER(C.c_int32)]),
    "oe_host_affine_records": (C.c_int32, [C.c_int32, f32p, C.c_int32, i32p, C.c_int32, f32p, f32p, i32p]),
}

OPTIONAL_SIGS: dict = {}


def _declare(sigs: dict, required: bool) -> None:
    for name, (res, args) in sigs.items():
        try:
            fn = getattr(lib, name)
        except AttributeError:
            if required:
                raise ImportError(f"{LIB_PATH} does not export {name}")
            continue
        fn.restype = res
        fn.argtypes = args


_declare(_SIGS, True)


def check(status: int) -> None:
    if status != OE_OK:
        raise OEError(status, lib.oe_last_error().decode())


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(f32p)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(i32p)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(f64p)
