"""ctypes binding of libtangerine_gpu.so (include/tangerine_gpu.h).

The library is the product: there is no CPU fallback.  Loading fails loudly
when the shared object is missing; device calls fail with the library's own
error message when no GPU is present.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libtangerine_gpu.so"

TG_OK = 0
TG_FLAG_FAST = 1
TG_FLAG_NO_PRUNE = 2
TG_FLAG_CERTIFY = 4

_f = C.c_float
_u32 = C.c_uint32
_u64 = C.c_uint64
_p = C.c_void_p
_fp = C.POINTER(C.c_float)
_u32p = C.POINTER(C.c_uint32)
_u64p = C.POINTER(C.c_uint64)


class TgGrid(C.Structure):
    _fields_ = [("x", _f), ("y", _f), ("z", _f), ("dx", _f), ("dy", _f), ("dz", _f),
                ("sx", _u64), ("sy", _u64), ("sz", _u64)]


class TgProgramInfo(C.Structure):
    _fields_ = [("instructions", _u64), ("max_stack", _u32), ("brushes", _u32), ("set_ops", _u32),
                ("lipschitz", _u32), ("flops_per_eval", C.c_double), ("leaf_extent", C.c_float),
                ("reserved", C.c_float)]


class TgExportStats(C.Structure):
    _fields_ = [("eval_ms", C.c_double), ("mesh_ms", C.c_double), ("total_ms", C.c_double),
                ("samples", _u64), ("tiles_evaluated", _u64), ("exact_evals", _u64), ("flops", C.c_double),
                ("bricks_total", _u64), ("bricks_flagged", _u64), ("prune_ms", C.c_double),
                ("cull_threshold", C.c_double), ("bricks_certified", _u64), ("block_ms", C.c_double),
                ("points_per_lane", _u32), ("waves", _u32), ("block_max_len", _u32), ("block_max_binary", _u32),
                ("block_lds_bytes", _u32), ("block_workgroups", _u32)]


class TgSurrogateStats(C.Structure):
    _fields_ = [("sample_ms", C.c_double), ("train_ms", C.c_double), ("total_ms", C.c_double),
                ("samples", _u64), ("calib_samples", _u64), ("surface_vertices", _u64), ("steps", _u32),
                ("final_loss", _f), ("margin", _f), ("mean_residual", _f)]


class TgTuneResult(C.Structure):
    _fields_ = [("points_per_lane", _u32), ("waves", _u32), ("workgroups_per_cu", _u32), ("min_prune_len", _u32),
                ("eval_ms", C.c_double), ("candidates", _u32), ("candidate_points", _u32 * 24),
                ("candidate_waves", _u32 * 24), ("candidate_workgroups_per_cu", _u32 * 24),
                ("candidate_ms", C.c_double * 24), ("candidate_len_ms", C.c_double * 3)]


class TgBatchStats(C.Structure):
    _fields_ = [("prune_ms", C.c_double), ("eval_ms", C.c_double), ("mesh_ms", C.c_double), ("total_ms", C.c_double),
                ("trees", _u64), ("blocks", _u64), ("samples", _u64), ("exact_evals", _u64), ("flops", C.c_double),
                ("points_per_lane", _u32), ("waves", _u32), ("workgroups", _u32), ("lds_bytes", _u32),
                ("max_len", _u32), ("max_binary", _u32), ("max_depth", _u32)]


class TgBatchTuneResult(C.Structure):
    _fields_ = [("points_per_lane", _u32), ("waves", _u32), ("workgroups_per_cu", _u32), ("eval_ms", C.c_double),
                ("candidates", _u32), ("candidate_points", _u32 * 18), ("candidate_waves", _u32 * 18),
                ("candidate_workgroups_per_cu", _u32 * 18), ("candidate_ms", C.c_double * 18)]


class TgCamera(C.Structure):
    _fields_ = [("eye", _f * 3), ("forward", _f * 3), ("right", _f * 3), ("up", _f * 3),
                ("tan_half_x", _f), ("tan_half_y", _f), ("width", _u32), ("height", _u32)]


class TgTraceStats(C.Structure):
    _fields_ = [("trace_ms", C.c_double), ("prune_ms", C.c_double), ("rays", _u64), ("hits", _u64),
                ("evaluations", _u64), ("flops", C.c_double)]


# name -> (restype, argtypes)
SIGNATURES = {
    "tg_last_error": (C.c_char_p, []),
    "tg_device_init": (C.c_int, [C.c_int]),
    "tg_device_synchronize": (C.c_int, []),
    "tg_device_copy_probe": (C.c_int, [_u64, C.c_int, C.POINTER(C.c_double)]),
    "tg_check_sqrt": (C.c_int, [_u32, C.POINTER(_u64)]),
    "tg_check_quotient": (C.c_int, [C.POINTER(C.c_float), _u32, C.POINTER(_u64), C.POINTER(_u64)]),
    "tg_sphere": (_p, [_f]),
    "tg_ellipsoid": (_p, [_f, _f, _f]),
    "tg_box": (_p, [_f, _f, _f]),
    "tg_torus": (_p, [_f, _f]),
    "tg_cylinder": (_p, [_f, _f]),
    "tg_plane": (_p, [_f, _f, _f]),
    "tg_cone": (_p, [_f, _f]),
    "tg_coninder": (_p, [_f, _f, _f]),
    "tg_union": (_p, [_p, _p]),
    "tg_inter": (_p, [_p, _p]),
    "tg_diff": (_p, [_p, _p]),
    "tg_blend_union": (_p, [_f, _p, _p]),
    "tg_blend_inter": (_p, [_f, _p, _p]),
    "tg_blend_diff": (_p, [_f, _p, _p]),
    "tg_flate": (_p, [_p, _f]),
    "tg_move": (_p, [_p, _f, _f, _f]),
    "tg_rotate": (_p, [_p, _f, _f, _f, _f]),
    "tg_rotate_x": (_p, [_p, _f]),
    "tg_rotate_y": (_p, [_p, _f]),
    "tg_rotate_z": (_p, [_p, _f]),
    "tg_scale": (_p, [_p, _f]),
    "tg_node_release": (None, [_p]),
    "tg_node_bounds": (C.c_int, [_p, _fp]),
    "tg_node_has_finite_bounds": (C.c_int, [_p]),
    "tg_node_leaf_count": (C.c_int, [_p]),
    "tg_node_stack_size": (C.c_int, [_p]),
    "tg_node_compile": (C.c_int, [_p, _u32p, _u64, _u64p]),
    "tg_program_from_node": (C.c_int, [_p, C.POINTER(_p)]),
    "tg_program_from_words": (C.c_int, [_u32p, _u64, C.POINTER(_p)]),
    "tg_program_info_get": (C.c_int, [_p, C.POINTER(TgProgramInfo)]),
    "tg_program_destroy": (None, [_p]),
    "tg_words_inspect": (C.c_int, [_u32p, _u64, C.POINTER(TgProgramInfo)]),
    "tg_eval_points": (C.c_int, [_p, _fp, _u64, _fp, _u32]),
    "tg_eval_points_device": (C.c_int, [_p, _p, _u64, _p, _u32, _p]),
    "tg_gradient_points": (C.c_int, [_p, _fp, _u64, _f, C.c_int, _fp, _u32]),
    "tg_raymarch": (C.c_int, [_p, _fp, _fp, _u64, C.c_int, _f, _f, _u32, C.POINTER(C.c_uint8), _fp, _fp]),
    "tg_raymarch_static": (C.c_int, [_p, _fp, _fp, _u64, C.c_int, _f, _f, _u32, C.POINTER(C.c_uint8), _fp, _fp]),
    "tg_pivot_towards": (C.c_int, [_p, _u64, _fp, _fp, _fp, _f, _f, _f, C.c_int, _fp, _fp, C.POINTER(C.c_int32), _u32]),
    "tg_pick": (C.c_int, [C.POINTER(_p), _fp, _u32, _fp, _fp, _u64, _u32, C.POINTER(C.c_int32), _fp, _fp]),
    "tg_camera_look_at": (C.c_int, [C.POINTER(TgCamera), _fp, _fp, _fp, _f, _u32, _u32]),
    "tg_tracer_create": (C.c_int, [C.POINTER(_p)]),
    "tg_tracer_destroy": (None, [_p]),
    "tg_tracer_set_domain": (C.c_int, [_p, _fp, _f]),
    "tg_tracer_set_passes": (C.c_int, [_p, C.c_int, C.c_int, _u32]),
    "tg_tracer_render": (C.c_int, [_p, _p, C.POINTER(TgCamera), C.c_int, _f, _f, _f, _u32]),
    "tg_tracer_copy": (C.c_int, [_p, _fp, _fp, _u32p]),
    "tg_tracer_stats": (C.c_int, [_p, C.POINTER(TgTraceStats)]),
    "tg_mesher_create": (C.c_int, [C.POINTER(_p)]),
    "tg_mesher_destroy": (None, [_p]),
    "tg_mesher_export": (C.c_int, [_p, _p, C.POINTER(TgGrid), _u32]),
    "tg_mesher_mesh_values": (C.c_int, [_p, _fp, C.POINTER(TgGrid)]),
    "tg_mesher_counts": (C.c_int, [_p, _u64p, _u64p]),
    "tg_mesher_copy_mesh": (C.c_int, [_p, _fp, _u32p]),
    "tg_mesher_copy_grid": (C.c_int, [_p, _fp]),
    "tg_mesher_stats": (C.c_int, [_p, C.POINTER(TgExportStats)]),
    "tg_mesher_autotune": (C.c_int, [_p, _p, C.POINTER(TgGrid), _u32, C.POINTER(TgTuneResult)]),
    "tg_mesher_set_tuning": (C.c_int, [_p, _u32, _u32, _u32, _u32]),
    "tg_mesher_set_culled_tuning": (C.c_int, [_p, _u32, _u32, _u32]),
    "tg_mesher_set_emission": (C.c_int, [_p, _u32]),
    "tg_mesher_write_stl": (C.c_int, [_p, _p, C.c_char_p, _f, _u32]),
    "tg_mesher_write_ply": (C.c_int, [_p, _p, C.c_char_p, _f, _u32]),
    "tg_mesher_grid_get": (C.c_int, [_p, C.POINTER(TgGrid)]),
    "tg_mesher_export_sodapop": (C.c_int, [_p, _p, _fp, _f, C.c_int, _u32]),
    "tg_mesher_copy_normals": (C.c_int, [_p, _fp]),
    "tg_mesher_export_points": (C.c_int, [_p, _p, _fp, _fp, _fp, C.c_int, _u32]),
    "tg_mesher_export_vox": (C.c_int, [_p, _p, _fp, _f, C.c_int, C.c_char_p, _u32, C.POINTER(_u64)]),
    "tg_mesher_copy_vox_cells": (C.c_int, [_p, _u32p]),
    "tg_vox_write": (C.c_int, [C.c_char_p, _u32p, _u32p, _u64, C.c_int]),
    "tg_batch_create": (C.c_int, [C.POINTER(_p)]),
    "tg_batch_destroy": (None, [_p]),
    "tg_batch_export": (C.c_int, [_p, C.POINTER(_p), C.POINTER(TgGrid), _u32, _u32]),
    "tg_batch_replay": (C.c_int, [_p, C.POINTER(C.c_double)]),
    "tg_batch_autotune": (C.c_int, [_p, C.POINTER(TgBatchTuneResult)]),
    "tg_batch_set_tuning": (C.c_int, [_p, _u32, _u32, _u32]),
    "tg_batch_mesher": (_p, [_p, _u32]),
    "tg_batch_stats_get": (C.c_int, [_p, C.POINTER(TgBatchStats)]),
    "tg_mesher_export_culled": (C.c_int, [_p, _p, C.POINTER(TgGrid), _p, _u32]),
    "tg_mesher_brick_flags": (C.c_int, [_p, C.POINTER(C.c_uint8), _u64, _u64p]),
    "tg_mesher_active_bricks": (C.c_int, [_p, C.POINTER(C.c_uint8), _u64, _u64p]),
    "tg_surrogate_create": (C.c_int, [C.POINTER(_p), _fp, _f, _f, _f, _u64]),
    "tg_surrogate_destroy": (None, [_p]),
    "tg_surrogate_params": (C.c_int, [_p, _fp, _fp, C.c_int]),
    "tg_surrogate_predict": (C.c_int, [_p, _fp, _u64, _fp]),
    "tg_surrogate_gradients": (C.c_int, [_p, _fp, _fp, _u64, _fp, _fp]),
    "tg_surrogate_fit": (C.c_int, [_p, _p, _u64, _u64, _f, _u32, _u32, C.POINTER(TgSurrogateStats)]),
    "tg_surrogate_set_margin": (C.c_int, [_p, _f]),
    "tg_surrogate_get_stats": (C.c_int, [_p, C.POINTER(TgSurrogateStats)]),
    "tg_surrogate_loss_log": (C.c_int, [_p, _fp, _u64, _u64p]),
}


class TangerineError(RuntimeError):
    pass


_lib = None


def load(path: os.PathLike | str | None = None) -> C.CDLL:
    """Loads the shared library (once) and declares every signature."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    if not p.exists():
        raise TangerineError(
            f"{p} is missing: build the HIP library first (python -m tangerine_amd.build)")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


def last_error() -> str:
    msg = load().tg_last_error()
    return msg.decode() if msg else ""


def check(rc: int, what: str) -> None:
    if rc != TG_OK:
        raise TangerineError(f"{what} failed ({rc}): {last_error()}")


def grid_struct(origin, step, cells) -> TgGrid:
    g = TgGrid()
    g.x, g.y, g.z = (float(v) for v in origin)
    g.dx, g.dy, g.dz = (float(v) for v in step)
    g.sx, g.sy, g.sz = (int(v) for v in cells)
    return g
