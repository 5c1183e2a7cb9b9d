"""ctypes declarations of include/mupsi.h (the C ABI of libmupsi_hip.so).

The struct layouts here mirror the header field for field; tests check that
the library exports every symbol the header declares.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libmupsi_hip.so"

c_float_p = C.POINTER(C.c_float)
c_int32_p = C.POINTER(C.c_int32)

MUPSI_OK = 0
STATUS = {
    0: "ok", 1: "invalid argument", 2: "HIP runtime error", 3: "out of memory",
    4: "aborted", 5: "unsupported", 6: "I/O error", 7: "parse error",
}
GP_SINGLE_REALIZATION, GP_RENEWAL_PLUS = 0, 1
MEAN_SPHERE, MEAN_MESH = 0, 1
PRIM_SPHERE, PRIM_MESH = 0, 1
MEDIUM_GP, MEDIUM_SDF, MEDIUM_GRID = 0, 1, 2
BSDF_LAMBERTIAN, BSDF_NULL, BSDF_SPECULAR = 0, 1, 2
BSDF_BY_NAME = {"lambertian": BSDF_LAMBERTIAN, "null": BSDF_NULL, "specular": BSDF_SPECULAR}


class GpParams(C.Structure):
    _fields_ = [("sigma", C.c_float), ("kernel_radius", C.c_float), ("aniso", C.c_float * 3),
                ("points_per_cell", C.c_int32)]


class MeshDesc(C.Structure):
    _fields_ = [("vertices", c_float_p), ("normals", c_float_p), ("texcoords", c_float_p),
                ("faces", c_int32_p), ("face_tex", c_int32_p),
                ("n_vertices", C.c_int32), ("n_normals", C.c_int32), ("n_texcoords", C.c_int32),
                ("n_faces", C.c_int32), ("position", C.c_float * 3), ("scale", C.c_float)]


class TextureDesc(C.Structure):
    _fields_ = [("rgb", c_float_p), ("width", C.c_int32), ("height", C.c_int32)]


class PrimitiveDesc(C.Structure):
    _fields_ = [("type", C.c_int32), ("bsdf", C.c_int32), ("center", C.c_float * 3),
                ("radius", C.c_float), ("has_emission", C.c_int32), ("emission", C.c_float * 3),
                ("int_medium", C.c_int32), ("mesh", MeshDesc), ("albedo", TextureDesc)]


class GridMediumDesc(C.Structure):
    _fields_ = [("mean", c_float_p), ("variance", c_float_p), ("resolution", C.c_int32),
                ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3),
                ("noise_scale", C.c_float)]


class SceneDesc(C.Structure):
    _fields_ = [("width", C.c_int32), ("height", C.c_int32), ("cam_pos", C.c_float * 3),
                ("cam_dir", C.c_float * 3), ("cam_up", C.c_float * 3), ("fov", C.c_float),
                ("max_bounce", C.c_int32), ("gp_mode", C.c_int32), ("skip_medium", C.c_int32),
                ("has_gp_medium", C.c_int32), ("mean_type", C.c_int32),
                ("mean_center", C.c_float * 3), ("mean_radius", C.c_float),
                ("mean_mesh", MeshDesc), ("kernel", GpParams), ("sky", TextureDesc),
                ("primitives", C.POINTER(PrimitiveDesc)), ("n_primitives", C.c_int32),
                ("medium_kind", C.c_int32), ("grid", GridMediumDesc)]


class CondSoA(C.Structure):
    _fields_ = [("active", C.c_void_p), ("cx", C.c_void_p), ("cy", C.c_void_p),
                ("cz", C.c_void_p), ("u_tilde", C.c_void_p), ("gx", C.c_void_p),
                ("gy", C.c_void_p), ("gz", C.c_void_p)]


class Stats(C.Structure):
    _fields_ = [("paths", C.c_int64), ("march_calls", C.c_int64), ("march_steps", C.c_int64),
                ("field_evals", C.c_int64), ("impulse_evals", C.c_int64),
                ("march_kernel_launches", C.c_int64), ("march_kernel_ms", C.c_double),
                ("total_ms", C.c_double)]


class GuideConfig(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("learning_rate", C.c_float), ("alpha", C.c_float),
                ("train_passes", C.c_int32), ("batch_size", C.c_int32),
                ("min_steps_per_pass", C.c_int32), ("max_steps_per_pass", C.c_int32),
                ("max_records", C.c_int32),
                ("box_center", C.c_float * 3), ("box_radius", C.c_float),
                ("indirect_clamp", C.c_float), ("rr_boost", C.c_float),
                ("passes_per_update", C.c_int32)]


class GuideStats(C.Structure):
    _fields_ = [("train_passes_done", C.c_int64), ("train_steps", C.c_int64),
                ("train_samples", C.c_int64), ("last_loss", C.c_double),
                ("train_ms", C.c_double), ("infer_ms", C.c_double)]


class BatchTuning(C.Structure):
    _fields_ = [("paths_per_launch", C.c_int64), ("march_blocks_per_cu", C.c_int32),
                ("use_graph", C.c_int32)]


class BatchStats(C.Structure):
    _fields_ = [("paths", C.c_int64), ("launches", C.c_int64), ("total_ms", C.c_double),
                ("tune_ms", C.c_double), ("graph_builds", C.c_int64)]


GUIDE_NUM_PARAMS = 13097
MARCH_NOALLOC_MAX_RAYS = 262144

VP = C.c_void_p
I32, I64 = C.c_int32, C.c_int64

# name -> (restype, argtypes); pointers to device memory are passed as void*.
SIGNATURES = {
    "mupsi_open": (I32, [I32, C.POINTER(VP)]),
    "mupsi_close": (None, [VP]),
    "mupsi_status_string": (C.c_char_p, [I32]),
    "mupsi_last_error": (C.c_char_p, []),
    "mupsi_device_info": (I32, [VP, C.c_char_p, I32, C.POINTER(I32)]),
    "mupsi_scene_upload": (I32, [VP, C.POINTER(SceneDesc), C.POINTER(VP)]),
    "mupsi_scene_load_json": (I32, [VP, C.c_char_p, C.POINTER(VP), C.POINTER(I32)]),
    "mupsi_scene_free": (None, [VP]),
    "mupsi_scene_size": (I32, [VP, C.POINTER(I32), C.POINTER(I32)]),
    "mupsi_scene_set_gp_medium": (I32, [VP, I32, VP, C.c_float, C.POINTER(MeshDesc),
                                        C.POINTER(GpParams)]),
    "mupsi_scene_set_settings": (I32, [VP, I32, I32, I32]),
    "mupsi_render_passes": (I32, [VP, I32, I32, I32, VP]),
    "mupsi_scene_set_graph": (I32, [VP, I32]),
    "mupsi_scene_graph_builds": (I32, [VP, C.POINTER(I64)]),
    "mupsi_scene_set_fused": (I32, [VP, I32]),
    "mupsi_reset_accum": (I32, [VP]),
    "mupsi_read_accum": (I32, [VP, VP, C.POINTER(I64)]),
    "mupsi_write_accum": (I32, [VP, VP, I64]),
    "mupsi_trace_paths": (I32, [VP, VP, VP, VP, I64, VP]),
    "mupsi_last_stats": (I32, [VP, C.POINTER(Stats)]),
    "mupsi_psi_eval": (I32, [VP, C.POINTER(GpParams), VP, VP, VP, VP, I64, VP, VP]),
    "mupsi_cell_impulses": (I32, [VP, C.POINTER(GpParams), VP, VP, I64, VP, VP, VP]),
    "mupsi_field_eval": (I32, [VP, VP, VP, VP, VP, C.POINTER(CondSoA), I64, VP]),
    "mupsi_march": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, C.POINTER(CondSoA), I64,
                          VP, VP, VP, VP, VP]),
    "mupsi_march_noalloc": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, C.POINTER(CondSoA),
                                  I64, VP, VP, VP, VP, VP]),
    "mupsi_condition": (I32, [VP, VP, VP, VP, VP, VP, VP, VP, I64, VP, VP]),
    "mupsi_mesh_intersect": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP, VP, I64, VP, VP, VP, VP]),
    "mupsi_mesh_occluded": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP, VP, I64, VP]),
    "mupsi_mesh_closest": (I32, [VP, I32, VP, VP, VP, I64, VP, VP, VP]),
    "mupsi_mesh_inside": (I32, [VP, I32, VP, VP, VP, I64, VP]),
    "mupsi_mesh_query_cost": (I32, [VP, I32, I32, VP, VP, VP, I64, VP, VP]),
    "mupsi_last_kernel_ms": (I32, [VP, C.POINTER(C.c_float)]),
    "mupsi_bvh_layout": (I32, [VP, I32, VP, I32, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "mupsi_bvh_build": (I32, [VP, I32, VP, I32, VP, C.POINTER(I32), VP, C.POINTER(C.c_uint32)]),
    "mupsi_bvh_xbins": (I32, [VP, I32, VP, I32, C.POINTER(I32), C.POINTER(I32), VP,
                              C.POINTER(C.c_int64), VP, C.c_int64, VP, C.c_int64]),
    "mupsi_batch_create": (I32, [VP, VP, I32, C.POINTER(VP)]),
    "mupsi_batch_free": (None, [VP]),
    "mupsi_batch_set_tuning": (I32, [VP, C.POINTER(BatchTuning)]),
    "mupsi_batch_get_tuning": (I32, [VP, C.POINTER(BatchTuning)]),
    "mupsi_batch_render": (I32, [VP, VP, VP, VP]),
    "mupsi_batch_stats_get": (I32, [VP, C.POINTER(BatchStats)]),
    "mupsi_guide_create": (I32, [VP, C.POINTER(GuideConfig), C.POINTER(VP)]),
    "mupsi_guide_free": (None, [VP]),
    "mupsi_guide_get_params": (I32, [VP, VP]),
    "mupsi_guide_set_params": (I32, [VP, VP]),
    "mupsi_guide_stats_get": (I32, [VP, C.POINTER(GuideStats)]),
    "mupsi_scene_set_guide": (I32, [VP, VP]),
    "mupsi_scene_guide_records": (I32, [VP, VP, VP, VP, I64, C.POINTER(I64)]),
    "mupsi_guide_forward": (I32, [VP, VP, I64, VP, VP]),
    "mupsi_guide_gradient": (I32, [VP, VP, VP, VP, I64, VP, C.POINTER(C.c_double)]),
    "mupsi_guide_train_step": (I32, [VP, VP, VP, VP, I64]),
    "mupsi_guide_encode": (I32, [VP, VP, VP, VP, I64, VP]),
    "mupsi_guide_sample": (I32, [VP, VP, VP, VP, I64, VP, VP, VP]),
}

_lib = None


def load(path: Path | None = None) -> C.CDLL:
    """Load libmupsi_hip.so and declare its signatures; raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = Path(path) if path else LIB_PATH
    try:
        # torch ships its own libamdhip64: loading torch first makes the library
        # bind to that same HIP runtime (loaded the other way round, two runtimes
        # end up in the process and torch finds no device)
        import torch  # noqa: F401
    except ImportError:
        pass
    if not p.exists():
        raise RuntimeError(
            f"{p} is missing: build the HIP library first (python -c 'import __graft_entry__ as g; g.build()')")
    lib = C.CDLL(str(p))
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class MupsiError(RuntimeError):
    def __init__(self, status: int, where: str, detail: str = ""):
        self.status = status
        super().__init__(f"{where}: {STATUS.get(status, status)}" + (f" ({detail})" if detail else ""))


def check(status: int, where: str) -> None:
    if status != MUPSI_OK:
        lib = load()
        detail = lib.mupsi_last_error()
        raise MupsiError(status, where, detail.decode() if detail else "")
