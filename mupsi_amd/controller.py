"""Controller (src/controller/controller.h:18-110, controller.cpp) over the GPU renderer.

The reference's Controller owns the scene, a render thread and two framebuffers:
the render thread renders into the Renderer's framebuffer and, after each spp
pass of a progressive render, copies it into the back buffer and swaps it with
the front buffer under a mutex, so a display thread can copy a stable image at
any time.  Here the render thread drives the HIP library (each pass is one
`mupsi_render_passes` call polling the shared cancel flag between bounce
iterations), the framebuffers are float32 host arrays of the averaged radiance,
and the GP parameter setters are applied to the uploaded scene in place
(`mupsi_scene_set_gp_medium`, `mupsi_scene_set_settings`) when the next
progressive render starts, as Controller::startProgressive does.

    c = Controller()
    c.load("scene.json")
    c.start_progressive(64)
    while c.get_current_spp() < 64:
        if c.is_spp_pass_done():
            c.ack_spp_pass()
            rgb = c.copy_display_to()      # Reinhard-tonemapped front buffer
    c.stop()
"""
from __future__ import annotations

import ctypes as C
import json
import os
import sys
import threading
from pathlib import Path
from typing import Optional

import numpy as np

from . import Context, MupsiError, Renderer, Scene
from .framebuffer import save_framebuffer, tonemap


def load_settings(j: dict, spp: int = 1) -> dict:
    """What Controller::load keeps from a scene.json object besides the scene itself
    (controller.cpp:40-137): spp (the previous value when absent), max_bounce,
    output, gp_mode, and the gp_medium mean and kernel parameters that
    applyGpMedium rebuilds from (kernel at the top level, as the reference reads it)."""
    out = {"spp": int(j.get("spp", spp)), "max_bounce": int(j.get("max_bounce", 3)),
           "output": j.get("output", "test.png"),
           "gp_mode": j.get("gp_mode", "single_realization"),
           "has_gp_medium": "gp_medium" in j}
    if out["has_gp_medium"]:
        gpm = j["gp_medium"]
        out["mean_type"] = gpm.get("mean_type", "sphere")
        if out["mean_type"] == "sphere":
            out["mean_center"] = [float(x) for x in gpm["mean_center"]]
            out["mean_radius"] = float(gpm.get("mean_radius", 70.0))
        elif out["mean_type"] == "mesh":
            t = gpm.get("transform", {}) or {}
            out["mean_mesh"] = gpm.get("mesh", gpm.get("file"))
            out["mean_mesh_pos"] = [float(x) for x in t.get("position", [0.0, 0.0, 0.0])]
            out["mean_mesh_scale"] = float(t.get("scale", 1.0))
        else:
            raise ValueError(f"unknown mean_type: {out['mean_type']}")
        kern = j.get("kernel", {}) or {}
        out["sigma"] = float(kern.get("sigma", 1.0))
        out["length_scale"] = float(kern.get("length_scale", 1.0))
        out["points_per_cell"] = int(kern.get("points_per_cell", 3))
    return out


class Controller:
    def __init__(self, ctx: Optional[Context] = None, device: int = 0):
        self.ctx = ctx if ctx is not None else Context(device)
        self._scene: Optional[Scene] = None
        self._renderer: Optional[Renderer] = None
        self._front: Optional[np.ndarray] = None
        self._display_mtx = threading.Lock()
        self._thread: Optional[threading.Thread] = None
        self._cancel = C.c_int32(0)
        self._spp_pass_done = False
        self._current_spp = 0
        self._spp = 1
        self._config_path = ""
        self._output_path = "test.png"
        self._max_bounce = 3
        self._skip_medium = False
        self.last_error: Optional[BaseException] = None
        # GP medium parameters (controller.h:37-53)
        self._has_gp_medium = False
        self._gp_kernel_sigma = 1.0
        self._gp_kernel_length = 1.0
        self._gp_points_per_cell = 3
        self._gp_mode = "single_realization"
        self._gp_mean_type = "sphere"
        self._gp_mean_center = [0.0, 0.0, 0.0]
        self._gp_mean_radius = 70.0
        self._gp_mean_mesh = None          # OBJ path or mesh dict
        self._gp_mean_mesh_pos = [0.0, 0.0, 0.0]
        self._gp_mean_mesh_scale = 1.0

    # --- load (controller.cpp:40-264) ---

    def load(self, config: dict | str | os.PathLike) -> bool:
        """Load a scene.json path or a scene dict; False (with a message) on failure.
        A render still running on the previous scene is stopped first."""
        self.stop()
        try:
            if isinstance(config, (str, os.PathLike)):
                st = load_settings(json.loads(Path(config).read_text()), self._spp)
                scene = Scene(self.ctx, config)
                self._config_path = str(config)
            else:
                st = load_settings(config, self._spp)
                scene = Scene(self.ctx, config)
                self._config_path = ""
        except (OSError, ValueError, KeyError, MupsiError) as e:
            print(f"Controller::load failed to load {config}: {e}", file=sys.stderr)
            return False
        if self._scene is not None:
            self._scene.close()
        self._scene = scene
        self._spp, self._max_bounce = st["spp"], st["max_bounce"]
        self._output_path, self._gp_mode = st["output"], st["gp_mode"]
        self._has_gp_medium = st["has_gp_medium"]
        if self._has_gp_medium:
            self._gp_mean_type = st["mean_type"]
            if self._gp_mean_type == "sphere":
                self._gp_mean_center, self._gp_mean_radius = st["mean_center"], st["mean_radius"]
            else:
                self._gp_mean_mesh = st["mean_mesh"]
                self._gp_mean_mesh_pos = st["mean_mesh_pos"]
                self._gp_mean_mesh_scale = st["mean_mesh_scale"]
            self._gp_kernel_sigma = st["sigma"]
            self._gp_kernel_length = st["length_scale"]
            self._gp_points_per_cell = st["points_per_cell"]
        return True

    # --- render control (controller.cpp:268-345) ---

    def _run(self, body) -> None:
        try:
            body()
        except BaseException as e:  # reported, as the reference prints and keeps going
            self.last_error = e
            print(f"Controller: render thread failed: {e}", file=sys.stderr)

    def _snapshot_settings(self, scene: Scene) -> None:
        scene.set_settings(self._max_bounce, self._gp_mode, self._skip_medium)

    def start(self) -> None:
        """Full render of spp passes on the render thread; the front buffer is
        replaced when it completes uncancelled (controller.cpp:268-291)."""
        if self._scene is None:
            print("Controller::start: no scene loaded", file=sys.stderr)
            return
        self.stop()
        self._cancel.value = 0
        scene, spp = self._scene, self._spp

        def body():
            self._snapshot_settings(scene)
            r = Renderer(scene)
            r.set_cancel_flag(self._cancel)
            self._renderer = r
            r.prepare_render()
            r.start_render(spp)
            if not self._cancel.value:
                img = r.framebuffer().astype(np.float32)
                with self._display_mtx:
                    self._front = img
                self._spp_pass_done = True

        self._thread = threading.Thread(target=self._run, args=(body,), daemon=True)
        self._thread.start()

    def start_progressive(self, target_spp: int, skip_medium: bool = False) -> None:
        """Progressive render of target_spp passes: after each pass the averaged
        image is copied to the back buffer, swapped to the front, and
        current_spp / spp_pass_done are set (controller.cpp:293-345)."""
        if self._scene is None:
            return
        self.stop()
        self._cancel.value = 0
        self._spp_pass_done = False
        self._current_spp = 0
        self._skip_medium = bool(skip_medium)
        scene = self._scene
        w, h = scene.width, scene.height
        with self._display_mtx:
            if self._front is None or self._front.shape[:2] != (h, w):
                self._front = np.zeros((h, w, 3), np.float32)

        def body():
            self._snapshot_settings(scene)
            r = Renderer(scene)
            r.set_cancel_flag(self._cancel)
            self._renderer = r
            r.prepare_render()
            if self._has_gp_medium:
                self.apply_gp_medium()
            back = np.zeros((h, w, 3), np.float32)

            def on_spp(s: int) -> None:
                nonlocal back
                if self._cancel.value:
                    return
                back[...] = r.framebuffer()
                with self._display_mtx:
                    self._front, back = back, self._front
                self._current_spp = s
                self._spp_pass_done = True

            r.start_render_progressive(target_spp, on_spp)

        self._thread = threading.Thread(target=self._run, args=(body,), daemon=True)
        self._thread.start()

    def cancel(self) -> None:
        self._cancel.value = 1

    def stop(self) -> None:
        self.cancel()
        if self._thread is not None:
            self._thread.join()
            self._thread = None

    def join(self, timeout: Optional[float] = None) -> bool:
        """Wait for the render thread without cancelling it (no reference
        counterpart: its callers poll is_spp_pass_done); True once it has ended."""
        if self._thread is None:
            return True
        self._thread.join(timeout)
        return not self._thread.is_alive()

    def is_spp_pass_done(self) -> bool:
        return self._spp_pass_done

    def ack_spp_pass(self) -> None:
        self._spp_pass_done = False

    def get_current_spp(self) -> int:
        return self._current_spp

    # --- display (controller.cpp:347-387) ---

    def _copy(self, dst: Optional[np.ndarray], tone: bool) -> Optional[np.ndarray]:
        with self._display_mtx:
            if self._front is None:
                return None
            img = tonemap(self._front) if tone else self._front.copy()
        if dst is None:
            return img
        np.copyto(dst, img.reshape(dst.shape))  # dst: (H, W, 3) or H*W*3 floats, any strides
        return dst

    def copy_display_to(self, dst: Optional[np.ndarray] = None) -> Optional[np.ndarray]:
        """Reinhard-tonemapped copy of the front buffer, (H, W, 3) float32 [y, x]
        (or into dst of H*W*3 floats); None before anything was displayed."""
        return self._copy(dst, True)

    def copy_display_raw_to(self, dst: Optional[np.ndarray] = None) -> Optional[np.ndarray]:
        return self._copy(dst, False)

    def save_display(self, path: str | os.PathLike) -> None:
        with self._display_mtx:
            if self._front is not None:
                save_framebuffer(path, self._front)

    # --- accessors (controller.h:85-108) ---

    def get_scene(self) -> Optional[Scene]:
        return self._scene

    def get_renderer(self) -> Optional[Renderer]:
        return self._renderer

    def get_spp(self) -> int:
        return self._spp

    def set_output_path(self, path: str) -> None:
        self._output_path = path

    def output_path(self) -> str:
        return self._output_path

    def config_path(self) -> str:
        return self._config_path

    def has_gp_medium(self) -> bool:
        return self._has_gp_medium

    def gp_kernel_sigma(self) -> float:
        return self._gp_kernel_sigma

    def set_gp_kernel_sigma(self, v: float) -> None:
        self._gp_kernel_sigma = float(v)

    def gp_kernel_length(self) -> float:
        return self._gp_kernel_length

    def set_gp_kernel_length(self, v: float) -> None:
        self._gp_kernel_length = float(v)

    def gp_points_per_cell(self) -> int:
        return self._gp_points_per_cell

    def set_gp_points_per_cell(self, v: int) -> None:
        self._gp_points_per_cell = int(v)

    def gp_mode(self) -> str:
        return self._gp_mode

    def set_gp_mode(self, v: str) -> None:
        """Any value other than "renewal_plus" means single realization
        (controller.cpp:389-395); applied when the next render starts."""
        self._gp_mode = v

    def apply_gp_medium(self) -> None:
        """Controller::applyGpMedium (controller.cpp:397-426): a new GPMedium from
        the current parameters with the kernel's aniso reset to (1, 1, 1).  A mesh
        mean that cannot be read is reported and the old medium kept."""
        if self._scene is None:
            return
        kernel = dict(sigma=self._gp_kernel_sigma, length_scale=self._gp_kernel_length,
                      aniso=[1.0, 1.0, 1.0], points_per_cell=self._gp_points_per_cell)
        if self._gp_mean_type == "sphere":
            self._scene.set_gp_medium("sphere", self._gp_mean_center, self._gp_mean_radius,
                                      kernel=kernel)
        elif self._gp_mean_type == "mesh":
            try:
                self._scene.set_gp_medium("mesh", mesh=self._gp_mean_mesh,
                                          transform=dict(position=self._gp_mean_mesh_pos,
                                                         scale=self._gp_mean_mesh_scale),
                                          kernel=kernel)
            except (OSError, ValueError) as e:
                print(f"Controller::applyGpMedium failed to reload mesh: {e}", file=sys.stderr)

    def close(self) -> None:
        self.stop()
        if self._scene is not None:
            self._scene.close()
            self._scene = None
