"""Scene descriptions in the reference's scene.json vocabulary, lowered to mupsi_scene_desc.

A scene is a dict with the keys of mupsi's scene.json (Controller::load,
src/controller/controller.cpp:40-264).  Where the JSON names a file (an OBJ
mesh, an HDR image), the dict may instead carry arrays:
  mesh:    {"vertices": (V,3) f32, "faces": (F,3) i32, "normals": (V,3) f32 | None,
            "texcoords": (T,2) f32 | None, "face_tex": (F,3) i32 | None}
  image:   (H, W, 3) float32, row 0 at the top.
`write_scene_files` turns such a dict into scene.json + OBJ + PFM files.
"""
from __future__ import annotations

import ctypes as C
import json
import os
from pathlib import Path
from typing import Any

import numpy as np

from . import _abi as A


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _ptr(arr: np.ndarray | None, ctype):
    if arr is None:
        return C.POINTER(ctype)()
    return arr.ctypes.data_as(C.POINTER(ctype))


def read_obj(path: str | os.PathLike) -> dict:
    """OBJ reader with Mesh::fetchFrom's rules (src/primitives/mesh.cpp:20-136)."""
    v, vn, vt, f, ft = [], [], [], [], []

    def norm_idx(i: int, count: int) -> int:
        return i - 1 if i > 0 else (count + i if i < 0 else 0)

    with open(path) as fh:
        for line in fh:
            line = line.rstrip("\r\n")
            if not line or line.startswith("#"):
                continue
            parts = line.split()
            if not parts:
                continue
            t = parts[0]
            if t == "v":
                v.append([float(x) for x in parts[1:4]])
            elif t == "vn":
                vn.append([float(x) for x in parts[1:4]])
            elif t == "vt":
                vt.append([float(x) for x in parts[1:3]])
            elif t == "f":
                vi, ti = [], []
                for spec in parts[1:]:
                    s = spec.split("/")
                    a = int(s[0])
                    b = int(s[1]) if len(s) > 1 and s[1] else 0
                    vi.append(norm_idx(a, len(v)))
                    ti.append(norm_idx(b, len(vt)))
                if len(vi) == 3:
                    f.append(vi)
                    ft.append(ti)
    if not v or not f:
        raise ValueError(f"no valid geometry in {path}")
    return {
        "vertices": _f32(v), "faces": _i32(f),
        "normals": _f32(vn) if vn else None,
        "texcoords": _f32(vt) if vt else None,
        "face_tex": _i32(ft) if vt else None,
    }


def read_pfm(path: str | os.PathLike) -> np.ndarray:
    """PFM -> (H, W, 3) float32, row 0 at the top (PFM stores rows bottom-up)."""
    with open(path, "rb") as fh:
        magic = fh.readline().strip()
        w, h = (int(x) for x in fh.readline().split())
        scale = float(fh.readline())
        ch = 3 if magic == b"PF" else 1
        data = np.frombuffer(fh.read(w * h * ch * 4), dtype=">f4" if scale > 0 else "<f4")
    img = data.reshape(h, w, ch)[::-1].astype(np.float32)
    if ch == 1:
        img = np.repeat(img, 3, axis=2)
    return np.ascontiguousarray(img)


def write_pfm(path: str | os.PathLike, img: np.ndarray) -> None:
    img = np.asarray(img, dtype=np.float32)
    h, w, _ = img.shape
    with open(path, "wb") as fh:
        fh.write(b"PF\n%d %d\n-1.0\n" % (w, h))
        fh.write(np.ascontiguousarray(img[::-1]).astype("<f4").tobytes())


def write_obj(path: str | os.PathLike, mesh: dict) -> None:
    v, f = np.asarray(mesh["vertices"]), np.asarray(mesh["faces"])
    vn, vt, ft = mesh.get("normals"), mesh.get("texcoords"), mesh.get("face_tex")
    lines = [f"v {x:.9g} {y:.9g} {z:.9g}" for x, y, z in v.tolist()]
    if vn is not None:
        lines += [f"vn {x:.9g} {y:.9g} {z:.9g}" for x, y, z in np.asarray(vn).tolist()]
    if vt is not None:
        lines += [f"vt {x:.9g} {y:.9g}" for x, y in np.asarray(vt).tolist()]
    if vt is not None and ft is not None:
        lines += [f"f {a+1}/{ta+1} {b+1}/{tb+1} {c+1}/{tc+1}"
                  for (a, b, c), (ta, tb, tc) in zip(f.tolist(), np.asarray(ft).tolist())]
    else:
        lines += [f"f {a+1} {b+1} {c+1}" for a, b, c in f.tolist()]
    Path(path).write_text("\n".join(lines) + "\n")


class SceneHolder:
    """Keeps every array a SceneDesc points at alive."""

    def __init__(self):
        self.keep: list[Any] = []

    def arr(self, a, kind):
        if a is None:
            return None
        a = _f32(a) if kind == "f" else _i32(a)
        self.keep.append(a)
        return a


def _mesh_desc(h: SceneHolder, mesh: dict | str, transform: dict | None) -> A.MeshDesc:
    if isinstance(mesh, (str, os.PathLike)):
        mesh = read_obj(mesh)
    d = A.MeshDesc()
    v = h.arr(mesh["vertices"], "f").reshape(-1, 3)
    f = h.arr(mesh["faces"], "i").reshape(-1, 3)
    vn = h.arr(mesh.get("normals"), "f")
    vt = h.arr(mesh.get("texcoords"), "f")
    ft = h.arr(mesh.get("face_tex"), "i") if vt is not None else None
    d.vertices = _ptr(v, C.c_float)
    d.faces = _ptr(f, C.c_int32)
    d.normals = _ptr(vn, C.c_float)
    d.texcoords = _ptr(vt, C.c_float)
    d.face_tex = _ptr(ft, C.c_int32)
    d.n_vertices = v.shape[0]
    d.n_faces = f.shape[0]
    d.n_normals = 0 if vn is None else vn.size // 3
    d.n_texcoords = 0 if vt is None else vt.size // 2
    t = transform or {}
    d.position[:] = [float(x) for x in t.get("position", [0.0, 0.0, 0.0])]
    d.scale = float(t.get("scale", 1.0))
    return d


def _texture(h: SceneHolder, img) -> A.TextureDesc:
    t = A.TextureDesc()
    if img is None:
        return t
    if isinstance(img, (str, os.PathLike)):
        img = read_pfm(img)
    a = h.arr(img, "f")
    t.rgb = _ptr(a, C.c_float)
    t.height, t.width = a.shape[0], a.shape[1]
    return t


def read_grid(path: str | os.PathLike, res: int) -> np.ndarray:
    """A GridMeanFunction file (meanfunction.cpp:41-54): res^3 raw float32, x-major.
    As the reference, a missing or short file leaves zeros where data is absent."""
    out = np.zeros(res ** 3, np.float32)
    try:
        data = np.fromfile(path, dtype="<f4", count=res ** 3)
    except (OSError, ValueError):
        return out
    out[: data.size] = data
    return out


def _mean_fields(d: A.SceneDesc, h: SceneHolder, block: dict) -> None:
    """mean_type / mean_center / mean_radius / file+transform of a gp_medium or
    sdf_medium block (controller.cpp:85-109, 140-168)."""
    mt = block.get("mean_type", "sphere")
    if mt == "sphere":
        d.mean_type = A.MEAN_SPHERE
        d.mean_center[:] = [float(x) for x in block["mean_center"]]
        d.mean_radius = float(block.get("mean_radius", 70.0))
    elif mt == "mesh":
        d.mean_type = A.MEAN_MESH
        d.mean_mesh = _mesh_desc(h, block.get("mesh", block.get("file")), block.get("transform"))
    else:
        raise ValueError(f"unknown mean_type: {mt}")


def build_desc(scene: dict) -> tuple[A.SceneDesc, SceneHolder]:
    """Lower a scene dict to a mupsi_scene_desc (same keys and defaults as Controller::load)."""
    h = SceneHolder()
    d = A.SceneDesc()
    d.width = int(scene.get("width", 256))
    d.height = int(scene.get("height", 256))
    cam = scene["camera"]
    d.cam_pos[:] = [float(x) for x in cam["pos"]]
    d.cam_dir[:] = [float(x) for x in cam["dir"]]
    d.cam_up[:] = [float(x) for x in cam.get("up", [0, 1, 0])]
    d.fov = float(cam.get("fov", 45.0))
    d.max_bounce = int(scene.get("max_bounce", 3))
    d.gp_mode = A.GP_RENEWAL_PLUS if scene.get("gp_mode") == "renewal_plus" else A.GP_SINGLE_REALIZATION
    d.skip_medium = int(bool(scene.get("skip_medium", False)))
    gpm = scene.get("gp_medium")
    sdf = scene.get("sdf_medium")
    ext = scene.get("external_gp_medium")
    # int_medium spheres take the GP medium, else the SDF medium, else the
    # external GP medium (controller.cpp:230-234)
    if gpm is None and sdf is not None:
        d.has_gp_medium = 1
        d.medium_kind = A.MEDIUM_SDF
        _mean_fields(d, h, sdf)
    elif gpm is None and ext is not None:
        d.has_gp_medium = 1
        d.medium_kind = A.MEDIUM_GRID
        res = int(ext.get("resolution", 48))
        base = Path(scene.get("config_dir", "."))

        def grid(key_arr, key_file):
            if ext.get(key_arr) is not None:
                return h.arr(np.asarray(ext[key_arr], np.float32).reshape(-1), "f")
            if ext.get(key_file):
                return h.arr(read_grid(base / ext[key_file], res), "f")
            return None
        mean = grid("mean", "mean_file")
        if mean is None:
            mean = h.arr(np.zeros(res ** 3, np.float32), "f")
        var = grid("variance", "variance_file")
        g = d.grid
        g.mean = _ptr(mean, C.c_float)
        g.variance = _ptr(var, C.c_float) if var is not None else None
        g.resolution = res
        g.bounds_min[:] = [float(x) for x in ext["bounds_min"]]
        g.bounds_max[:] = [float(x) for x in ext["bounds_max"]]
        g.noise_scale = float(ext.get("noise_scale", 1.0))
    if gpm is not None:
        d.has_gp_medium = 1
        d.medium_kind = A.MEDIUM_GP
        _mean_fields(d, h, gpm)
        k = scene.get("kernel", {})
        d.kernel.sigma = float(k.get("sigma", 1.0))
        d.kernel.kernel_radius = float(k.get("length_scale", 1.0))
        d.kernel.aniso[:] = [float(x) for x in k.get("aniso", [1, 1, 1])]
        d.kernel.points_per_cell = int(k.get("points_per_cell", 3))
    if scene.get("skydrome") is not None:
        d.sky = _texture(h, scene["skydrome"])
    prims = []
    for pj in scene.get("primitives", []):
        p = A.PrimitiveDesc()
        bsdf_name = pj.get("bsdf", "lambertian")
        p.bsdf = A.BSDF_BY_NAME.get(bsdf_name, A.BSDF_LAMBERTIAN)
        if pj.get("type", "sphere") == "sphere":
            p.type = A.PRIM_SPHERE
            p.center[:] = [float(x) for x in pj["center"]]
            p.radius = float(pj.get("radius", 100.0))
            if "emission" in pj:
                p.has_emission = 1
                p.emission[:] = [float(x) for x in pj["emission"]]
            p.int_medium = int("int_medium" in pj)
        elif pj["type"] == "mesh":
            p.type = A.PRIM_MESH
            p.mesh = _mesh_desc(h, pj.get("mesh", pj.get("file")), pj.get("transform"))
            # the reference ignores "emission" on meshes (controller.cpp:225-228); the
            # "extensions": {"mesh_emission": true} scene key turns it on
            if "emission" in pj and scene.get("extensions", {}).get("mesh_emission"):
                p.has_emission = 1
                p.emission[:] = [float(x) for x in pj["emission"]]
            if "texture" in pj and bsdf_name == "lambertian":
                p.albedo = _texture(h, pj["texture"])
        else:
            continue
        prims.append(p)
    arr = (A.PrimitiveDesc * max(1, len(prims)))(*prims)
    h.keep.append(arr)
    d.primitives = C.cast(arr, C.POINTER(A.PrimitiveDesc))
    d.n_primitives = len(prims)
    return d, h


def write_scene_files(scene: dict, directory: str | os.PathLike, name: str = "scene.json") -> Path:
    """Write scene.json (+ OBJ/PFM files for array payloads) in the reference's schema."""
    directory = Path(directory)
    directory.mkdir(parents=True, exist_ok=True)
    out = {k: v for k, v in scene.items()
           if k not in ("skydrome", "primitives", "gp_medium", "sdf_medium", "external_gp_medium")}
    if scene.get("skydrome") is not None:
        sky = scene["skydrome"]
        if not isinstance(sky, (str, os.PathLike)):
            write_pfm(directory / "sky.pfm", sky)
            sky = str(directory / "sky.pfm")
        out["skydrome"] = str(sky)
    for key in ("gp_medium", "sdf_medium"):
        if scene.get(key) is not None:
            g = dict(scene[key])
            if "mesh" in g:
                write_obj(directory / f"{key}_mesh.obj", g.pop("mesh"))
                g["file"] = str(directory / f"{key}_mesh.obj")
            out[key] = g
    if scene.get("external_gp_medium") is not None:
        # grid files are named relative to the scene file's directory
        # (controller.cpp:177-181), as raw little-endian float32
        g = dict(scene["external_gp_medium"])
        for arr_key, file_key in (("mean", "mean_file"), ("variance", "variance_file")):
            a = g.pop(arr_key, None)
            if a is not None:
                np.asarray(a, "<f4").reshape(-1).tofile(directory / f"{arr_key}.raw")
                g[file_key] = f"{arr_key}.raw"
        out["external_gp_medium"] = g
    prims = []
    for i, pj in enumerate(scene.get("primitives", [])):
        pj = dict(pj)
        if "mesh" in pj:
            fn = directory / f"prim{i}.obj"
            write_obj(fn, pj.pop("mesh"))
            pj["file"] = str(fn)
        if "texture" in pj and not isinstance(pj["texture"], (str, os.PathLike)):
            fn = directory / f"prim{i}_tex.pfm"
            write_pfm(fn, pj["texture"])
            pj["texture"] = str(fn)
        prims.append(pj)
    out["primitives"] = prims
    path = directory / name
    path.write_text(json.dumps(out, indent=2))
    return path
