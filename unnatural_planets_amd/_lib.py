"""ctypes binding of libupf.so (the C-ABI declared in include/upf.h).

The library is loaded from this package directory; there is no fallback: when it is
missing, every call raises.
"""
from __future__ import annotations

import ctypes
import os

from .spec import BakeSpec, BatchOpts, Chunk, FieldSpec, Grid, Layer, MaterialSpec, MeshBuffers, NbStats

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libupf.so")

_lib = None

# (plan, field, workspace, workspace bytes, positions, normals, indices, cap vertices,
#  cap triangles, chunk meshes, stream)
_BATCH_RUN_ARGS = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p,
                   ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]

# name -> (restype, argtypes)
SIGNATURES = {
    "upf_version": (ctypes.c_int, []),
    "upf_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "upf_layer_init": (ctypes.c_int, [ctypes.POINTER(Layer), ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                      ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                      ctypes.c_uint32, ctypes.c_float, ctypes.c_int32]),
    "upf_spec_set_elevation_mode": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.c_int32, ctypes.c_uint32]),
    "upf_spec_validate": (ctypes.c_int, [ctypes.POINTER(FieldSpec)]),
    "upf_flops_per_sample": (ctypes.c_double, [ctypes.POINTER(FieldSpec)]),
    "upf_adjoint_flops_per_sample": (ctypes.c_double, [ctypes.POINTER(FieldSpec)]),
    "upf_field_eval_grid": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_void_p,
                                           ctypes.c_void_p]),
    "upf_field_eval_grid_ex": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p]),
    "upf_field_eval_points": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.c_void_p, ctypes.c_int64,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "upf_field_adjoint": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "upf_field_adjoint_grid": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_void_p,
                                              ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "upf_mc_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Grid)]),
    "upf_mc_extract": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Grid), ctypes.c_float,
                                      ctypes.POINTER(MeshBuffers), ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.c_void_p]),
    "upf_mc_count": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Grid), ctypes.c_float, ctypes.c_void_p,
                                    ctypes.c_size_t, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                    ctypes.c_void_p]),
    "upf_field_mesh_count": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_float,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                            ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64),
                                            ctypes.c_void_p]),
    "upf_field_mesh_async": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_float,
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MeshBuffers),
                                            ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "upf_mc_extract_async": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Grid), ctypes.c_float,
                                            ctypes.POINTER(MeshBuffers), ctypes.c_void_p, ctypes.c_void_p,
                                            ctypes.c_size_t, ctypes.c_void_p]),
    "upf_static_scratch_bytes": (ctypes.c_size_t, []),
    "upf_field_mesh_async_static": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_float,
                                                   ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(MeshBuffers),
                                                   ctypes.c_void_p, ctypes.c_void_p]),
    "upf_mc_emit": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Grid), ctypes.c_float,
                                   ctypes.POINTER(MeshBuffers), ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "upf_batch_create": (ctypes.c_int, [ctypes.POINTER(Chunk), ctypes.c_int32, ctypes.POINTER(BatchOpts),
                                        ctypes.POINTER(ctypes.c_void_p)]),
    "upf_batch_destroy": (None, [ctypes.c_void_p]),
    "upf_batch_field_floats": (ctypes.c_int64, [ctypes.c_void_p]),
    "upf_batch_field_offset": (ctypes.c_int64, [ctypes.c_void_p, ctypes.c_int32]),
    "upf_batch_workspace_bytes": (ctypes.c_size_t, [ctypes.c_void_p]),
    "upf_batch_set_opts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchOpts)]),
    "upf_batch_get_opts": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(BatchOpts)]),
    "upf_batch_run": (ctypes.c_int, _BATCH_RUN_ARGS),
    "upf_batch_totals": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64),
                                        ctypes.POINTER(ctypes.c_int64), ctypes.c_void_p]),
    "upf_batch_capture": (ctypes.c_int, _BATCH_RUN_ARGS),
    "upf_batch_replay": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    "upf_batch_autotune": (ctypes.c_int, _BATCH_RUN_ARGS + [ctypes.POINTER(BatchOpts)]),
    "upf_mc_case_codes": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(Grid), ctypes.c_float, ctypes.c_void_p,
                                         ctypes.c_void_p]),
    "upf_nb_workspace_bytes": (ctypes.c_size_t, [ctypes.POINTER(Grid), ctypes.c_int64]),
    "upf_nb_extract": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.POINTER(Grid), ctypes.c_float, ctypes.c_float,
                                      ctypes.POINTER(MeshBuffers), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                      ctypes.POINTER(NbStats), ctypes.c_void_p]),
    "upf_image_dilate": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32,
                                        ctypes.c_int32, ctypes.c_void_p]),
    "upf_image_to_u8": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]),
    "upf_material_tiles": (ctypes.c_int, [ctypes.POINTER(MaterialSpec), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "upf_bake_material": (ctypes.c_int, [ctypes.POINTER(MaterialSpec), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                         ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]),
    "upf_image_height_to_normal": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_float,
                                                  ctypes.c_void_p, ctypes.c_void_p]),
    "upf_image_special_to_gltf_pbr": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p,
                                                     ctypes.c_void_p]),
    "upf_mesh_post_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int64]),
    "upf_mesh_remove_disconnected": (ctypes.c_int, [ctypes.POINTER(MeshBuffers), ctypes.POINTER(MeshBuffers),
                                                    ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "upf_mesh_water_mask": (ctypes.c_int, [ctypes.POINTER(FieldSpec), ctypes.c_float, ctypes.POINTER(MeshBuffers),
                                           ctypes.POINTER(MeshBuffers), ctypes.c_void_p, ctypes.c_size_t,
                                           ctypes.c_void_p]),
    "upf_atlas_per_triangle": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p,
                                              ctypes.c_void_p]),
    "upf_bake_rasterize": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                          ctypes.c_void_p, ctypes.c_void_p]),
    "upf_bake_rasterize_workspace_bytes": (ctypes.c_size_t, [ctypes.c_int64, ctypes.c_int32, ctypes.c_int32]),
    "upf_bake_rasterize_ws": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]),
    "upf_bake": (ctypes.c_int, [ctypes.POINTER(BakeSpec), ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
                                ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32,
                                ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
}


class UpfError(RuntimeError):
    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {status_string(status)} (status {status})")


def lib():
    """The loaded libupf.so.  Raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libupf.so is not built ({LIB_PATH}); run unnatural_planets_amd.build.build()")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def status_string(status: int) -> str:
    return lib().upf_status_string(status).decode()


def check(status: int, where: str) -> None:
    if status != 0:
        raise UpfError(status, where)
