"""ctypes binding of libbcubehip.so (the C ABI declared in include/bcubehip.h).

The product path has no CPU fallback: if the library is missing this module
raises, and every assembly call goes through the HIP kernels.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbcubehip.so")

BH_OK = 0
ERRORS = {-1: "BH_ERR_ARG", -2: "BH_ERR_UNSUPPORTED", -3: "BH_ERR_HIP", -4: "BH_ERR_OOM",
          -5: "BH_ERR_RANGE"}
CELLTYPE = {"Quad4": 0, "Quad9": 1, "Tri3": 2, "Hexa8": 3, "Hexa27": 4, "Bar2": 5}
MODE = {"fp64": 0, "fp32": 1, "bf16x3": 2}
HOST, DEVICE = 0, 1
FORM_STIFFNESS, FORM_MASS, FORM_SOURCE = 0, 1, 10
FORM_GRAD_SOURCE = (11, 12, 13)  # ∫ f ∂v/∂x_a dΩ, a = 0, 1, 2
FORM_ADVECTION = (2, 3, 4)  # ∫ η ∂u/∂x_a v dΩ, a = 0, 1, 2
COEF_CONST, COEF_CELL, COEF_QP = 0, 1, 2

EXPORTED = [
    "bh_ctx_create", "bh_ctx_destroy", "bh_last_error", "bh_ctx_set_stream", "bh_sync",
    "bh_ctx_set_timing", "bh_ctx_timings", "bh_ctx_timings_ex", "bh_ctx_set_autotune", "bh_ctx_set_tuning_cache", "bh_mesh_create", "bh_mesh_set_nodes",
    "bh_mesh_destroy", "bh_dofhandler", "bh_dofhandler_device", "bh_space_create", "bh_space_create_from_dofs", "bh_space_create_orders", "bh_space_edge_orders",
    "bh_space_info", "bh_space_dofs", "bh_space_destroy", "bh_pattern_build", "bh_pattern_build_device", "bh_pattern_info", "bh_pattern_tuning", "bh_pattern_set_tuning", "bh_pattern_onepass", "bh_pattern_set_offset_templates", "bh_pattern_offset_templates",
    "bh_dirichlet_create", "bh_dirichlet_info", "bh_dirichlet_positions", "bh_dirichlet_apply_matrix",
    "bh_dirichlet_apply_vector", "bh_dirichlet_destroy",
    "bh_blocks_create", "bh_blocks_info", "bh_blocks_export_csc", "bh_blocks_block_values",
    "bh_blocks_assemble", "bh_blocks_assemble_block", "bh_blocks_add_field", "bh_blocks_destroy",
    "bh_pattern_export_csc", "bh_pattern_values_device", "bh_pattern_destroy",
    "bh_assemble_bilinear", "bh_assemble_bilinear_noalloc", "bh_assemble_linear", "bh_quadrature_points", "bh_fe_eval_qp", "bh_dg_create",
    "bh_dg_info", "bh_dg_update_geometry", "bh_dg_transport_residual", "bh_dg_minv_axpy", "bh_dg_transport_stage",
    "bh_dg_destroy", "bh_dgb_create", "bh_dgb_info", "bh_dgb_export_csc", "bh_dgb_assemble",
    "bh_dgb_boundary_points", "bh_dgb_boundary_linear", "bh_dgb_destroy", "bh_euler_create",
    "bh_euler_residual", "bh_euler_step", "bh_euler_destroy", "bh_graph_begin",
    "bh_graph_end", "bh_graph_launch", "bh_graph_destroy",
]


class BhTerm(C.Structure):
    _fields_ = [("form", C.c_int32), ("coef_kind", C.c_int32), ("coef", C.c_void_p),
                ("scale", C.c_double)]


class BcubeHipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load libbcubehip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    try:
        # One HIP runtime per process: torch bundles its own libamdhip64 with the
        # same soname, and whichever loads first serves everyone.  Load torch's
        # first so that torch (device memory, streams) and this library agree.
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise BcubeHipError(
            f"HIP assembly library not built: {LIB_PATH} is missing "
            "(run `python bcube.jl_amd/build.py` or __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    P = C.POINTER
    sig = {
        "bh_ctx_create": (C.c_int, [C.c_int, P(vp)]),
        "bh_ctx_destroy": (C.c_int, [vp]),
        "bh_last_error": (C.c_char_p, [vp]),
        "bh_ctx_set_stream": (C.c_int, [vp, vp]),
        "bh_sync": (C.c_int, [vp]),
        "bh_ctx_set_timing": (C.c_int, [vp, C.c_int]),
        "bh_ctx_timings": (C.c_int, [vp, vp, C.c_int, P(C.c_int)]),
        "bh_ctx_timings_ex": (C.c_int, [vp, vp, vp, C.c_int, P(C.c_int)]),
        "bh_ctx_set_autotune": (C.c_int, [vp, C.c_int]),
        "bh_ctx_set_tuning_cache": (C.c_int, [vp, C.c_char_p]),
        "bh_pattern_tuning": (C.c_int, [vp, C.c_int, vp, vp]),
        "bh_pattern_set_tuning": (C.c_int, [vp, C.c_int, vp]),
        "bh_pattern_onepass": (C.c_int, [vp, vp]),
        "bh_pattern_set_offset_templates": (C.c_int, [vp, C.c_int]),
        "bh_pattern_offset_templates": (C.c_int, [vp]),
        "bh_dirichlet_create": (C.c_int, [vp, i64, vp, P(vp)]),
        "bh_dirichlet_info": (C.c_int, [vp, P(i64), P(i64)]),
        "bh_dirichlet_positions": (C.c_int, [vp, vp, vp, vp, vp, vp]),
        "bh_dirichlet_apply_matrix": (C.c_int, [vp, C.c_double, C.c_int, vp, C.c_int]),
        "bh_dirichlet_apply_vector": (C.c_int, [vp, vp, vp, C.c_int]),
        "bh_dirichlet_destroy": (C.c_int, [vp]),
        "bh_mesh_create": (C.c_int, [vp, C.c_int, i64, vp, i64, vp, vp, vp, P(vp)]),
        "bh_mesh_set_nodes": (C.c_int, [vp, vp, C.c_int]),
        "bh_mesh_destroy": (C.c_int, [vp]),
        "bh_dofhandler": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, P(i64)]),
        "bh_dofhandler_device": (C.c_int, [vp, C.c_int, C.c_int, vp, vp, P(i64)]),
        "bh_space_create": (C.c_int, [vp, C.c_int, C.c_int, P(vp)]),
        "bh_space_create_from_dofs": (C.c_int, [vp, C.c_int, C.c_int, i64, vp, vp, P(vp)]),
        "bh_space_create_orders": (C.c_int, [vp, vp, P(vp)]),
        "bh_space_edge_orders": (C.c_int, [vp, vp, P(i64)]),
        "bh_space_info": (C.c_int, [vp, P(i64), P(i32)]),
        "bh_space_dofs": (C.c_int, [vp, vp]),
        "bh_space_destroy": (C.c_int, [vp]),
        "bh_pattern_build": (C.c_int, [vp, vp, P(vp)]),
        "bh_pattern_build_device": (C.c_int, [vp, vp, P(vp)]),
        "bh_pattern_info": (C.c_int, [vp, P(i64), P(i64), P(i64), P(i32)]),
        "bh_pattern_export_csc": (C.c_int, [vp, vp, vp, C.c_int]),
        "bh_pattern_values_device": (C.c_int, [vp, P(vp)]),
        "bh_pattern_destroy": (C.c_int, [vp]),
        "bh_blocks_create": (C.c_int, [vp, C.c_int, vp, i64, C.c_int, vp, P(vp)]),
        "bh_blocks_info": (C.c_int, [vp, P(i64), P(i64)]),
        "bh_blocks_export_csc": (C.c_int, [vp, vp, vp, C.c_int]),
        "bh_blocks_block_values": (C.c_int, [vp, C.c_int, P(vp)]),
        "bh_blocks_assemble": (C.c_int, [vp, C.c_int, vp, vp, vp, C.c_int]),
        "bh_blocks_assemble_block": (C.c_int, [vp, C.c_int, C.c_int, P(BhTerm), C.c_int, C.c_int,
                                               C.c_int]),
        "bh_blocks_add_field": (C.c_int, [vp, C.c_int, vp, vp, C.c_int]),
        "bh_blocks_destroy": (C.c_int, [vp]),
        "bh_assemble_bilinear": (C.c_int, [vp, C.c_int, P(BhTerm), C.c_int, C.c_int, vp, C.c_int]),
        "bh_assemble_bilinear_noalloc": (C.c_int, [vp, C.c_int, P(BhTerm), C.c_int, C.c_int, vp]),
        "bh_assemble_linear": (C.c_int, [vp, C.c_int, P(BhTerm), C.c_int, C.c_int, vp, C.c_int]),
        "bh_quadrature_points": (C.c_int, [vp, C.c_int, vp, C.c_int, P(i32)]),
        "bh_fe_eval_qp": (C.c_int, [vp, C.c_int, vp, vp, C.c_int]),
        "bh_dg_create": (C.c_int, [vp, C.c_int, vp, C.c_int, P(vp)]),
        "bh_dg_info": (C.c_int, [vp, P(i64), P(i64), P(i32)]),
        "bh_dg_update_geometry": (C.c_int, [vp]),
        "bh_dg_transport_residual": (C.c_int, [vp, i32, vp, vp, C.c_int]),
        "bh_dg_minv_axpy": (C.c_int, [vp, i32, vp, vp, dbl, vp, vp, dbl, vp]),
        "bh_dg_transport_stage": (C.c_int, [vp, C.c_int32, vp, vp, C.c_double, vp, vp, C.c_double, vp, vp]),
        "bh_dg_destroy": (C.c_int, [vp]),
        "bh_dgb_create": (C.c_int, [vp, C.c_int, C.c_int, P(BhTerm), C.c_int, vp, vp, P(vp)]),
        "bh_dgb_info": (C.c_int, [vp, P(i64), P(i64), P(i64), P(i64), P(i32)]),
        "bh_dgb_export_csc": (C.c_int, [vp, vp, vp, C.c_int]),
        "bh_dgb_assemble": (C.c_int, [vp, vp, C.c_int]),
        "bh_dgb_boundary_points": (C.c_int, [vp, vp, C.c_int]),
        "bh_dgb_boundary_linear": (C.c_int, [vp, dbl, dbl, vp, vp, C.c_int]),
        "bh_dgb_destroy": (C.c_int, [vp]),
        "bh_euler_create": (C.c_int, [vp, C.c_int, dbl, i64, vp, vp, i64, P(vp)]),
        "bh_euler_residual": (C.c_int, [vp, C.c_int, dbl, vp, vp, C.c_int]),
        "bh_euler_step": (C.c_int, [vp, dbl, vp, vp, C.c_int]),
        "bh_euler_destroy": (C.c_int, [vp]),
        "bh_graph_begin": (C.c_int, [vp]),
        "bh_graph_end": (C.c_int, [vp, P(vp)]),
        "bh_graph_launch": (C.c_int, [vp]),
        "bh_graph_destroy": (C.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, ctx=None, what: str = ""):
    """Turn a non-zero return code into a BcubeHipError with the library's message."""
    if rc == BH_OK:
        return
    msg = lib().bh_last_error(ctx)
    msg = msg.decode() if msg else ""
    raise BcubeHipError(f"{what}: {ERRORS.get(rc, rc)}: {msg}")
