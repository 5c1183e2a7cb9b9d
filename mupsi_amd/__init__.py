"""mupsi_amd — MI355X (gfx950) GPIS path tracer, a drop-in for mupsi's CPU trace loop.

Host-side mirror of the reference's render entry points over the C ABI of
libmupsi_hip.so (include/mupsi.h):

    ctx = Context()                       # one HIP stream on one GPU
    scene = Scene(ctx, scene_dict)        # Controller::load + Renderer::prepareRender
    r = Renderer(scene)                   # Renderer
    r.start_render_progressive(spp, on_spp)   # Renderer::startRenderProgressive
    img = r.framebuffer()                 # Renderer::getFramebuffer (mean radiance)

There is no CPU fallback: every call fails with MupsiError when the HIP
library or a gfx950 device is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import threading
import weakref
from pathlib import Path
from typing import Callable, Optional

import numpy as np

from . import _abi
from ._abi import MupsiError, check
from .scene import build_desc, write_scene_files

__all__ = ["Context", "Scene", "Renderer", "Guide", "Batch", "MupsiError", "build_desc",
           "write_scene_files", "library_path"]


def library_path() -> Path:
    return _abi.LIB_PATH


class Context:
    """mupsi_open / mupsi_close: one HIP stream on one device."""

    def __init__(self, device: int = 0):
        self.lib = _abi.load()
        h = C.c_void_p()
        check(self.lib.mupsi_open(device, C.byref(h)), "mupsi_open")
        self.handle = h
        self.device = device
        self._scenes: "weakref.WeakSet[Scene]" = weakref.WeakSet()
        self._guides: "weakref.WeakSet[Guide]" = weakref.WeakSet()
        self._batches: "weakref.WeakSet[Batch]" = weakref.WeakSet()

    def info(self) -> tuple[str, int]:
        name = C.create_string_buffer(256)
        cus = C.c_int32()
        check(self.lib.mupsi_device_info(self.handle, name, 256, C.byref(cus)), "mupsi_device_info")
        return name.value.decode(), cus.value

    def close(self) -> None:
        """Free every scene of this context, then the context (mupsi_close)."""
        if self.handle:
            for b in list(self._batches):
                b.close()
            for s in list(self._scenes):
                s.close()
            for g in list(self._guides):
                g.close()
            self.lib.mupsi_close(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """An uploaded scene (mupsi_scene_upload / mupsi_scene_load_json)."""

    def __init__(self, ctx: Context, scene: dict | str | os.PathLike):
        self.ctx = ctx
        self.lib = ctx.lib
        h = C.c_void_p()
        self.spp = 1
        if isinstance(scene, (str, os.PathLike)):
            spp = C.c_int32()
            check(self.lib.mupsi_scene_load_json(ctx.handle, str(scene).encode(), C.byref(h),
                                                 C.byref(spp)), "mupsi_scene_load_json")
            self.spp = spp.value
        else:
            desc, holder = build_desc(scene)
            check(self.lib.mupsi_scene_upload(ctx.handle, C.byref(desc), C.byref(h)),
                  "mupsi_scene_upload")
            del holder
            self.spp = int(scene.get("spp", 1))
        self.handle = h
        ctx._scenes.add(self)
        w, hh = C.c_int32(), C.c_int32()
        check(self.lib.mupsi_scene_size(h, C.byref(w), C.byref(hh)), "mupsi_scene_size")
        self.width, self.height = w.value, hh.value

    def render_passes(self, k_begin: int, k_count: int, passes_per_launch: int = 0,
                      abort_flag: Optional[C.c_int32] = None) -> None:
        ptr = C.cast(C.byref(abort_flag), C.c_void_p) if abort_flag is not None else None
        check(self.lib.mupsi_render_passes(self.handle, k_begin, k_count, passes_per_launch, ptr),
              "mupsi_render_passes")

    def reset(self) -> None:
        check(self.lib.mupsi_reset_accum(self.handle), "mupsi_reset_accum")

    def set_graph(self, use_graph: bool = True) -> None:
        """Replay a captured bounce loop for plain renders without an abort flag
        (mupsi_scene_set_graph)."""
        check(self.lib.mupsi_scene_set_graph(self.handle, 1 if use_graph else 0),
              "mupsi_scene_set_graph")

    def set_fused(self, fused: bool = True) -> None:
        """Shade the medium hits of a sphere-mean GP medium inside the march launch
        (mupsi_scene_set_fused; the default); the images are the same bits either way."""
        check(self.lib.mupsi_scene_set_fused(self.handle, 1 if fused else 0),
              "mupsi_scene_set_fused")

    def graph_builds(self) -> int:
        n = C.c_int64()
        check(self.lib.mupsi_scene_graph_builds(self.handle, C.byref(n)), "mupsi_scene_graph_builds")
        return int(n.value)

    def set_gp_medium(self, mean_type: str = "sphere", mean_center=(0.0, 0.0, 0.0),
                      mean_radius: float = 70.0, mesh=None, transform: dict | None = None,
                      kernel: dict | None = None) -> None:
        """A new GPMedium for the scene's int_medium primitives (mupsi_scene_set_gp_medium,
        Controller::applyGpMedium): mean "sphere" (mean_center, mean_radius) or "mesh"
        (an OBJ path or mesh dict placed by transform {position, scale}); kernel has the
        scene.json kernel keys (sigma, length_scale, aniso, points_per_cell)."""
        from .batch import gp_params
        from .scene import SceneHolder, _mesh_desc
        h = SceneHolder()
        mt = {"sphere": _abi.MEAN_SPHERE, "mesh": _abi.MEAN_MESH}.get(mean_type)
        if mt is None:
            raise ValueError(f"unknown mean_type: {mean_type}")
        md = _mesh_desc(h, mesh, transform) if mt == _abi.MEAN_MESH else None
        centre = (C.c_float * 3)(*[float(x) for x in mean_center])
        gp = gp_params(kernel or {})
        check(self.lib.mupsi_scene_set_gp_medium(self.handle, mt, C.cast(centre, C.c_void_p),
                                                 float(mean_radius),
                                                 C.byref(md) if md is not None else None,
                                                 C.byref(gp)), "mupsi_scene_set_gp_medium")
        del h

    def set_settings(self, max_bounce: int, gp_mode: str, skip_medium: bool) -> None:
        """The RenderContext snapshot of the path tracer and GP settings
        (mupsi_scene_set_settings): gp_mode "renewal_plus" or "single_realization"."""
        mode = _abi.GP_RENEWAL_PLUS if gp_mode == "renewal_plus" else _abi.GP_SINGLE_REALIZATION
        check(self.lib.mupsi_scene_set_settings(self.handle, int(max_bounce), mode,
                                                int(bool(skip_medium))), "mupsi_scene_set_settings")

    def read_accum(self) -> tuple[np.ndarray, int]:
        buf = np.zeros((self.height, self.width, 3), np.float64)
        cnt = C.c_int64()
        check(self.lib.mupsi_read_accum(self.handle, buf.ctypes.data_as(C.c_void_p), C.byref(cnt)),
              "mupsi_read_accum")
        return buf, cnt.value

    def write_accum(self, sums: np.ndarray, count: int) -> None:
        sums = np.ascontiguousarray(sums, dtype=np.float64)
        check(self.lib.mupsi_write_accum(self.handle, sums.ctypes.data_as(C.c_void_p), count),
              "mupsi_write_accum")

    def image(self) -> np.ndarray:
        s, n = self.read_accum()
        return s / max(n, 1)

    def trace_paths(self, x, y, k) -> np.ndarray:
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        out = np.zeros((x.size, 3), np.float32)
        check(self.lib.mupsi_trace_paths(self.handle, x.ctypes.data_as(C.c_void_p),
                                         y.ctypes.data_as(C.c_void_p),
                                         k.ctypes.data_as(C.c_void_p), x.size,
                                         out.ctypes.data_as(C.c_void_p)), "mupsi_trace_paths")
        return out

    def set_guide(self, guide: "Optional[Guide]") -> None:
        """Attach a learned guide (None: the plain tracer); resets its training count."""
        check(self.lib.mupsi_scene_set_guide(self.handle, guide.handle if guide else None),
              "mupsi_scene_set_guide")
        self._guide = guide

    def guide_records(self, capacity: int = 1 << 22):
        """Training samples recorded for the attached guide: (feats, dirs, weights)."""
        f = np.zeros((capacity, 32), np.float32)
        d = np.zeros((capacity, 3), np.float32)
        w = np.zeros(capacity, np.float32)
        n = C.c_int64()
        check(self.lib.mupsi_scene_guide_records(self.handle, f.ctypes.data_as(C.c_void_p),
                                                 d.ctypes.data_as(C.c_void_p),
                                                 w.ctypes.data_as(C.c_void_p), capacity,
                                                 C.byref(n)), "mupsi_scene_guide_records")
        return f[:n.value], d[:n.value], w[:n.value]

    def stats(self) -> dict:
        s = _abi.Stats()
        check(self.lib.mupsi_last_stats(self.handle, C.byref(s)), "mupsi_last_stats")
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            if self.ctx.handle:           # mupsi_close already freed it otherwise
                self.lib.mupsi_scene_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Guide:
    """Learned path-guiding network (mupsi_guide_*): an MLP on MFMA proposing
    directions at GP-medium hits, trained online during the first train_passes
    passes of a render it is attached to (Scene.set_guide)."""

    DEFAULTS = dict(seed=1, learning_rate=1e-3, alpha=0.9, train_passes=32, batch_size=16384,
                    min_steps_per_pass=4, max_steps_per_pass=128, max_records=3,
                    box_center=(0.0, 0.0, 0.0), box_radius=1.0, indirect_clamp=1.0,
                    rr_boost=4.0, passes_per_update=4)

    def __init__(self, ctx: Context, **cfg):
        self.ctx = ctx
        self.lib = ctx.lib
        c = dict(self.DEFAULTS)
        c.update(cfg)
        self.config = c
        k = _abi.GuideConfig()
        k.seed = int(c["seed"])
        k.learning_rate = float(c["learning_rate"])
        k.alpha = float(c["alpha"])
        k.train_passes = int(c["train_passes"])
        k.batch_size = int(c["batch_size"])
        k.min_steps_per_pass = int(c["min_steps_per_pass"])
        k.max_steps_per_pass = int(c["max_steps_per_pass"])
        k.max_records = int(c["max_records"])
        k.box_center[:] = [float(x) for x in c["box_center"]]
        k.box_radius = float(c["box_radius"])
        k.indirect_clamp = float(c["indirect_clamp"])
        k.rr_boost = float(c["rr_boost"])
        k.passes_per_update = int(c["passes_per_update"])
        h = C.c_void_p()
        check(self.lib.mupsi_guide_create(ctx.handle, C.byref(k), C.byref(h)), "mupsi_guide_create")
        self.handle = h
        ctx._guides.add(self)

    @classmethod
    def for_scene(cls, ctx: Context, scene: dict, **cfg) -> "Guide":
        """Encoding box around the GP medium of a scene dict (mean sphere or mesh)."""
        gm = scene.get("gp_medium", {})
        if gm.get("mean_type", "sphere") == "sphere":
            c = [float(x) for x in gm.get("mean_center", [0, 0, 0])]
            r = 2.0 * float(gm.get("mean_radius", 1.0))
        else:
            v = np.asarray(gm["mesh"]["vertices"], np.float64)
            tr = gm.get("transform", {})
            v = v * float(tr.get("scale", 1.0)) + np.asarray(tr.get("position", [0, 0, 0]))
            lo, hi = v.min(0), v.max(0)
            c = list((lo + hi) / 2)
            r = 1.5 * float(np.linalg.norm(hi - lo) / 2)
        cfg.setdefault("box_center", c)
        cfg.setdefault("box_radius", r)
        return cls(ctx, **cfg)

    def params(self) -> np.ndarray:
        out = np.zeros(_abi.GUIDE_NUM_PARAMS, np.float32)
        check(self.lib.mupsi_guide_get_params(self.handle, out.ctypes.data_as(C.c_void_p)),
              "mupsi_guide_get_params")
        return out

    def set_params(self, theta: np.ndarray) -> None:
        theta = np.ascontiguousarray(theta, np.float32)
        assert theta.size == _abi.GUIDE_NUM_PARAMS
        check(self.lib.mupsi_guide_set_params(self.handle, theta.ctypes.data_as(C.c_void_p)),
              "mupsi_guide_set_params")

    def stats(self) -> dict:
        s = _abi.GuideStats()
        check(self.lib.mupsi_guide_stats_get(self.handle, C.byref(s)), "mupsi_guide_stats_get")
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            if self.ctx.handle:
                self.lib.mupsi_guide_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Several scenes of one context traced in shared wavefront launches
    (mupsi_batch_*): render(k_begin, k_count) renders passes
    [k_begin[s], k_begin[s] + k_count[s]) of every scene s into its own
    accumulator (read with Scene.read_accum / Scene.image)."""

    def __init__(self, ctx: Context, scenes: list, tuning: dict | None = None):
        self.ctx = ctx
        self.lib = ctx.lib
        self.scenes = list(scenes)
        arr = (C.c_void_p * len(self.scenes))(*[sc.handle.value for sc in self.scenes])
        h = C.c_void_p()
        check(self.lib.mupsi_batch_create(ctx.handle, arr, len(self.scenes), C.byref(h)),
              "mupsi_batch_create")
        self.handle = h
        ctx._batches.add(self)
        if tuning:
            self.set_tuning(**tuning)

    def set_tuning(self, paths_per_launch: int = 0, march_blocks_per_cu: int = 0,
                   use_graph: bool = True) -> None:
        t = _abi.BatchTuning(paths_per_launch, march_blocks_per_cu, 1 if use_graph else 0)
        check(self.lib.mupsi_batch_set_tuning(self.handle, C.byref(t)), "mupsi_batch_set_tuning")

    def tuning(self) -> dict:
        t = _abi.BatchTuning()
        check(self.lib.mupsi_batch_get_tuning(self.handle, C.byref(t)), "mupsi_batch_get_tuning")
        return {name: getattr(t, name) for name, _ in t._fields_}

    def render(self, k_begin, k_count, abort_flag: Optional[C.c_int32] = None) -> None:
        n = len(self.scenes)
        kb = np.ascontiguousarray(np.broadcast_to(np.asarray(k_begin, np.int32), (n,)))
        kc = np.ascontiguousarray(np.broadcast_to(np.asarray(k_count, np.int32), (n,)))
        ptr = C.cast(C.byref(abort_flag), C.c_void_p) if abort_flag is not None else None
        check(self.lib.mupsi_batch_render(self.handle, kb.ctypes.data_as(C.c_void_p),
                                          kc.ctypes.data_as(C.c_void_p), ptr), "mupsi_batch_render")

    def stats(self) -> dict:
        s = _abi.BatchStats()
        check(self.lib.mupsi_batch_stats_get(self.handle, C.byref(s)), "mupsi_batch_stats_get")
        return {name: getattr(s, name) for name, _ in s._fields_}

    def close(self) -> None:
        if getattr(self, "handle", None):
            if self.ctx.handle:
                self.lib.mupsi_batch_free(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Renderer:
    """Renderer (src/rendering/renderer.h:14-40) on the GPU.

    start_render_progressive(target_spp, on_spp) runs passes k = 0..target_spp-1,
    calls on_spp(k+1) after each completed pass group and returns False if the
    cancel flag was raised (the sums keep the last completed group), True otherwise.
    """

    def __init__(self, scene: Scene, passes_per_call: int = 0):
        self.scene = scene
        self._cancel = C.c_int32(0)
        self.passes_per_call = passes_per_call
        self._done = 0
        self._target = 0

    def set_cancel_flag(self, flag: "C.c_int32 | threading.Event | None" = None) -> None:
        """Renderer::setCancelFlag: a shared C.c_int32, which the library also polls
        between bounce iterations, or a threading.Event polled between pass groups."""
        if isinstance(flag, C.c_int32):
            self._cancel = flag
            self._cancel_event = None
        else:
            self._cancel_event = flag

    def cancel(self) -> None:
        self._cancel.value = 1

    def prepare_render(self) -> None:
        self.scene.reset()
        self._done = 0

    def start_render(self, spp: int) -> bool:
        return self.start_render_progressive(spp)

    def start_render_progressive(self, target_spp: int,
                                 on_spp: Optional[Callable[[int], None]] = None) -> bool:
        self._target = target_spp
        self._done = 0
        step = self.passes_per_call or (target_spp if on_spp is None else 1)
        k = 0
        while k < target_spp:
            ev = getattr(self, "_cancel_event", None)
            if self._cancel.value or (ev is not None and ev.is_set()):
                return False
            n = min(step, target_spp - k)
            try:
                self.scene.render_passes(k, n, 0, self._cancel)
            except MupsiError as e:
                if e.status == 4:  # aborted
                    return False
                raise
            k += n
            self._done = k
            if on_spp:
                on_spp(k)
        return True

    def progress(self) -> int:
        return 0 if self._target == 0 else (100 * self._done) // self._target

    def framebuffer(self) -> np.ndarray:
        return self.scene.image()


from .controller import Controller  # noqa: E402  (needs the classes above)

__all__.append("Controller")
