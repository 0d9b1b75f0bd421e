"""ctypes mirror of include/clothmi.h and the loader of the HIP library.

The library is built in-tree (licenta_amd/libclothmi.so, see licenta_amd/build.py).
There is no CPU fallback: if the shared object is missing or fails to load, every
entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libclothmi.so")


class CmParams(C.Structure):
    _fields_ = [
        ("timestep", C.c_float),
        ("gravity", C.c_float * 3),
        ("k_stretch", C.c_float),
        ("k_bend", C.c_float),
        ("thickness", C.c_float),
        ("solver_iterations", C.c_int32),
        ("ground_height", C.c_float),
        ("world_min", C.c_float * 3),
        ("world_max", C.c_float * 3),
        ("global_vel_damp", C.c_float),
        ("k_damp", C.c_float),
        ("strain_limit", C.c_float),
        ("k_friction", C.c_float),
        ("pinned_mass", C.c_float),
        ("sphere_center", C.c_float * 3),
        ("sphere_radius", C.c_float),
        ("morton_bits", C.c_int32),
        ("morton_normalize", C.c_int32),
        ("pair_cap", C.c_int32),
        ("flags", C.c_uint32),
        ("ref_compat", C.c_uint32),
        ("reserved", C.c_int32 * 7),
    ]


class CmContact(C.Structure):
    _fields_ = [("v", C.c_int32 * 4), ("w", C.c_float * 4), ("t", C.c_float), ("n", C.c_float * 3)]


CONTACT_DTYPE = np.dtype([("v", "<i4", (4,)), ("w", "<f4", (4,)), ("t", "<f4"), ("n", "<f4", (3,))])
assert CONTACT_DTYPE.itemsize == 48 == C.sizeof(CmContact)


def make_params(d: dict) -> CmParams:
    p = CmParams()
    for name, typ in CmParams._fields_:
        if name == "reserved":  # capacity overrides (see cm_params in clothmi.h)
            for i, x in enumerate(d.get("reserved", ())):
                p.reserved[i] = int(x)
            continue
        val = d[name]
        if isinstance(val, (tuple, list, np.ndarray)):
            arr = getattr(p, name)
            for i, x in enumerate(val):
                arr[i] = x
        else:
            setattr(p, name, val)
    return p


def ptr(a, ctype=C.c_void_p):
    """Pointer to a numpy array's data (None passes through as NULL)."""
    if a is None:
        return None
    return a.ctypes.data_as(ctype)


_lib = None


def lib():
    """Load libclothmi.so (raises if it is not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m licenta_amd.build` "
            "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64 = C.c_void_p, C.c_int32, C.c_int64
    L.cm_last_error.restype = C.c_char_p
    L.cm_default_params.argtypes = [C.POINTER(CmParams)]
    L.cm_create.argtypes = [C.POINTER(CmParams), vp, i32, vp, i32, vp, vp, vp, i32, vp, vp,
                            C.POINTER(vp)]
    L.cm_create_batch.argtypes = [C.POINTER(CmParams), i32, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                  vp, C.POINTER(vp)]
    L.cm_destroy.argtypes = [vp]
    L.cm_step.argtypes = [vp, i32]
    L.cm_sync.argtypes = [vp]
    for f in ("cm_read_positions", "cm_read_velocities", "cm_read_normals", "cm_read_vertices"):
        getattr(L, f).argtypes = [vp, vp]
    L.cm_set_sphere.argtypes = [vp, vp, C.c_float]
    L.cm_set_state.argtypes = [vp, vp, vp, vp, vp]
    L.cm_get_counts.argtypes = [vp, vp]
    L.cm_get_buffer.argtypes = [vp, i32, vp, i64, C.POINTER(i64)]
    L.cm_get_warnings.argtypes = [vp, C.POINTER(C.c_uint32)]
    L.cm_query_pairs.argtypes = [vp, vp, i32, C.POINTER(i32)]
    L.cm_query_contacts.argtypes = [vp, vp, i32, C.POINTER(i32), vp, i32, C.POINTER(i32)]
    L.cm_tri_update.argtypes = [i32, i32, vp, vp, vp, C.c_float, i32, i32, vp, vp, vp, vp]
    L.cm_tri_update_noalloc.argtypes = L.cm_tri_update.argtypes
    L.cm_lbvh_sort_temp_bytes.argtypes = [i32]
    L.cm_lbvh_sort_temp_bytes.restype = C.c_size_t
    L.cm_lbvh_sort.argtypes = [i32, i32, vp, vp, vp, C.c_size_t, vp]
    L.cm_lbvh_build_temp_bytes.argtypes = [i32]
    L.cm_lbvh_build_temp_bytes.restype = C.c_size_t
    L.cm_lbvh_build.argtypes = [i32, vp, i32, vp, vp, vp, vp, vp, C.c_size_t, vp]
    L.cm_lbvh_refit.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_size_t, vp]
    L.cm_lbvh_pairs_temp_bytes.argtypes = [i32, i32]
    L.cm_lbvh_pairs_temp_bytes.restype = C.c_size_t
    L.cm_lbvh_pairs.argtypes = [i32, vp, vp, vp, vp, vp, vp, i32, vp, vp, vp, vp, C.c_size_t, vp]
    L.cm_ms_grid_springs.argtypes = [i32, i32, vp, C.c_float, C.c_float, C.c_float, vp, vp, vp,
                                     vp, vp, vp, C.POINTER(i32)]
    L.cm_ms_step.argtypes = [i32, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, C.c_float,
                             C.c_float, C.c_float, i32, C.c_float, vp, vp]
    L.cm_impulses_temp_bytes.argtypes = [i32, i32]
    L.cm_impulses_temp_bytes.restype = C.c_size_t
    L.cm_impulses.argtypes = [i32, vp, vp, vp, vp, vp, i32, vp, i32, C.c_float, C.c_float,
                              C.c_float, C.c_float, vp, C.c_size_t, vp]
    L.cm_time_kernel.argtypes = [vp, i32, i32, C.POINTER(C.c_float)]
    L.cm_stage_times.argtypes = [vp, i32, vp]
    L.cm_set_stiffness.argtypes = [vp, vp, vp]
    L.cm_forward_ckpt.argtypes = [vp, i32, i32]
    L.cm_adjoint.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.cm_autotune.argtypes = [vp, i32, vp]
    _lib = L
    return L


def check(rc: int) -> None:
    if rc != 0:
        msg = lib().cm_last_error()
        raise RuntimeError(f"clothmi error {rc}: {msg.decode() if msg else ''}")


# exported names declared in include/clothmi.h (kept in sync by tests/test_abi.py)
EXPORTS = (
    "cm_default_params", "cm_last_error", "cm_create", "cm_create_batch", "cm_destroy", "cm_step",
    "cm_sync",
    "cm_read_positions", "cm_read_velocities", "cm_read_normals", "cm_set_sphere",
    "cm_set_state", "cm_get_counts", "cm_get_buffer", "cm_get_warnings", "cm_query_pairs",
    "cm_query_contacts", "cm_tri_update", "cm_tri_update_noalloc", "cm_lbvh_sort_temp_bytes",
    "cm_lbvh_sort",
    "cm_lbvh_build_temp_bytes", "cm_lbvh_build", "cm_lbvh_refit", "cm_lbvh_pairs_temp_bytes",
    "cm_lbvh_pairs", "cm_time_kernel", "cm_stage_times", "cm_set_stiffness", "cm_forward_ckpt",
    "cm_adjoint", "cm_autotune", "cm_read_vertices", "cm_ms_grid_springs", "cm_ms_step",
    "cm_impulses_temp_bytes", "cm_impulses",
)
# Rendering::VertexFormat as cm_read_vertices fills it (position, color, normal, uv)
VERTEX_DTYPE = np.dtype([("position", "<f4", (3,)), ("color", "<f4", (4,)),
                         ("normal", "<f4", (3,)), ("uv", "<f4", (2,))])
