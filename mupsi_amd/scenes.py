"""Procedural scenes of the benchmark workloads (seeded, bit-reproducible).

- `env_map`: the 64x32 FP32 environment map: a sky gradient plus one Gaussian
  sun lobe of peak 50.
- `icosphere`: a radius-1 icosphere subdivided n times, optionally displaced
  along the vertex normals by seeded 3-octave value noise.
- `generated_batch(seed)`: 64 uneven generated scenes for the batched renderer.
- `preset_scene(name)`: the named scene dicts (PRESETS): sphere_single_128 (sphere
  mean, single realization, 128x128, 16 spp), sphere_renewal_256 (sphere mean,
  renewal_plus, 256x256, 64 spp), icosphere_renewal_512 (displaced-icosphere mesh
  mean, 512x512) and guided_sphere_quad_1024 (the renewal_plus sphere scene with a
  mirror sphere and a small emissive quad at 1024x1024, the learned-guiding
  workload).
"""
from __future__ import annotations

import numpy as np

ENV_SEED = 20240611
MESH_SEED = 7


def env_map(width: int = 64, height: int = 32, seed: int = ENV_SEED) -> np.ndarray:
    """Sky gradient + Gaussian sun lobe of peak 50; (H, W, 3) float32, row 0 = up (+y)."""
    rng = np.random.default_rng(seed)
    sun_theta = rng.uniform(0.35, 0.6)     # angle from the zenith
    sun_phi = rng.uniform(0.0, 2 * np.pi)
    sun = np.array([np.sin(sun_theta) * np.cos(sun_phi), np.cos(sun_theta),
                    np.sin(sun_theta) * np.sin(sun_phi)])
    # texel centres -> directions with the skydrome mapping inverted:
    # u = (atan2(z, x) + pi) / 2pi, v_img = 1 - acos(-y)/pi, row = v_img * h
    u = (np.arange(width) + 0.5) / width
    v_img = (np.arange(height) + 0.5) / height
    phi = u * 2 * np.pi - np.pi
    theta = np.pi * (1.0 - v_img)          # acos(-y)
    y = -np.cos(theta)[:, None] * np.ones((1, width))
    s = np.sin(theta)[:, None]
    x = s * np.cos(phi)[None, :]
    z = s * np.sin(phi)[None, :]
    up = np.clip(y, -1, 1)
    zenith = np.array([0.25, 0.45, 0.95])
    horizon = np.array([0.9, 0.9, 0.85])
    ground = np.array([0.18, 0.16, 0.14])
    t = np.clip(up, 0, 1)[..., None] ** 0.5
    sky = horizon * (1 - t) + zenith * t
    sky = np.where(up[..., None] >= 0, sky, ground * (0.6 + 0.4 * (1 + up[..., None])))
    cosang = np.clip(x * sun[0] + y * sun[1] + z * sun[2], -1, 1)
    ang = np.arccos(cosang)
    lobe = 50.0 * np.exp(-0.5 * (ang / 0.06) ** 2)
    img = sky + lobe[..., None] * np.array([1.0, 0.95, 0.85])
    return np.ascontiguousarray(img.astype(np.float32))


def _icosahedron():
    t = (1.0 + 5.0 ** 0.5) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]],
                 dtype=np.float64)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9],
                  [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2],
                  [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10],
                  [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    return v / np.linalg.norm(v, axis=1, keepdims=True), f


def _subdivide(v: np.ndarray, f: np.ndarray):
    e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    e.sort(axis=1)
    uniq, inv = np.unique(e, axis=0, return_inverse=True)
    inv = inv.reshape(-1)
    mid = v[uniq[:, 0]] + v[uniq[:, 1]]
    mid /= np.linalg.norm(mid, axis=1, keepdims=True)
    base = v.shape[0]
    nf = f.shape[0]
    a = base + inv[:nf]
    b = base + inv[nf:2 * nf]
    c = base + inv[2 * nf:]
    f0, f1, f2 = f[:, 0], f[:, 1], f[:, 2]
    nfaces = np.concatenate([np.stack([f0, a, c], 1), np.stack([f1, b, a], 1),
                             np.stack([f2, c, b], 1), np.stack([a, b, c], 1)])
    return np.concatenate([v, mid]), nfaces


def _lattice(ix, iy, iz, seed):
    h = (ix.astype(np.uint64) * np.uint64(73856093)) ^ (iy.astype(np.uint64) * np.uint64(19349663)) \
        ^ (iz.astype(np.uint64) * np.uint64(83492791)) ^ np.uint64(seed * 2654435761 % (1 << 32))
    h = (h ^ (h >> np.uint64(13))) * np.uint64(0x5bd1e995)
    h = h ^ (h >> np.uint64(15))
    return (h & np.uint64(0xFFFFFF)).astype(np.float64) / float(0xFFFFFF) * 2.0 - 1.0


def value_noise(p: np.ndarray, seed: int, octaves: int = 3) -> np.ndarray:
    """Seeded 3-octave trilinear value noise, amplitude sum 1, range [-1, 1]."""
    out = np.zeros(p.shape[0])
    amp, freq, total = 1.0, 4.0, 0.0
    for o in range(octaves):
        q = p * freq + 1000.0
        i = np.floor(q).astype(np.int64)
        fr = q - i
        s = fr * fr * (3 - 2 * fr)
        acc = np.zeros(p.shape[0])
        for dx in (0, 1):
            for dy in (0, 1):
                for dz in (0, 1):
                    w = (s[:, 0] if dx else 1 - s[:, 0]) * (s[:, 1] if dy else 1 - s[:, 1]) \
                        * (s[:, 2] if dz else 1 - s[:, 2])
                    acc += w * _lattice(i[:, 0] + dx, i[:, 1] + dy, i[:, 2] + dz, seed + o)
        out += amp * acc
        total += amp
        amp *= 0.5
        freq *= 2.0
    return out / total


def icosphere(subdivisions: int, displacement: float = 0.0, seed: int = MESH_SEED) -> dict:
    """Radius-1 icosphere (20 * 4^n faces), displaced along normals by value noise."""
    v, f = _icosahedron()
    for _ in range(subdivisions):
        v, f = _subdivide(v, f)
    if displacement:
        v = v * (1.0 + displacement * value_noise(v, seed))[:, None]
    return {"vertices": v.astype(np.float32), "faces": f.astype(np.int32),
            "normals": None, "texcoords": None, "face_tex": None}


def ground_quad(y: float, half: float = 40.0, z0: float = -60.0, z1: float = 10.0) -> dict:
    v = np.array([[-half, y, z0], [half, y, z0], [half, y, z1], [-half, y, z1]], np.float32)
    f = np.array([[0, 2, 1], [0, 3, 2]], np.int32)   # normals point up (+y)
    return {"vertices": v, "faces": f, "normals": None, "texcoords": None, "face_tex": None}


CAMERA = {"pos": [0.0, 0.0, 0.0], "dir": [0.0, 0.0, -1.0], "up": [0.0, 1.0, 0.0], "fov": 45.0}


PRESETS = ("sphere_single_128", "sphere_renewal_256", "icosphere_renewal_512",
           "guided_sphere_quad_1024")


def preset_scene(name: str, *, width: int | None = None, height: int | None = None,
                 mesh_subdivisions: int = 7) -> dict:
    """Scene dict of a named benchmark workload (spp is stored under "spp")."""
    sky = env_map()
    if name in ("sphere_single_128", "sphere_renewal_256"):
        single = name == "sphere_single_128"
        res = 128 if single else 256
        return {
            "width": width or res, "height": height or res,
            "spp": 16 if single else 64, "max_bounce": 10, "max_medium_bounce": 10,
            "gp_mode": "single_realization" if single else "renewal_plus",
            "camera": dict(CAMERA),
            "kernel": {"sigma": 1.0 if single else 1.5, "length_scale": 1.0 if single else 2.0,
                       "aniso": [1, 1, 1] if single else [1, 1, 4], "points_per_cell": 3},
            "gp_medium": {"mean_type": "sphere", "mean_center": [0, 0, -500], "mean_radius": 70},
            "skydrome": sky,
            "primitives": [{"type": "sphere", "center": [0, 0, -500], "radius": 130,
                            "bsdf": "null", "int_medium": True}],
        }
    if name == "icosphere_renewal_512":
        return {
            "width": width or 512, "height": height or 512, "spp": 64, "max_bounce": 10,
            "max_medium_bounce": 10, "gp_mode": "renewal_plus", "camera": dict(CAMERA),
            "kernel": {"sigma": 0.3, "length_scale": 0.5, "aniso": [1, 1, 1], "points_per_cell": 4},
            "gp_medium": {"mean_type": "mesh",
                          "mesh": icosphere(mesh_subdivisions, displacement=0.05),
                          "transform": {"scale": 2.0, "position": [0.0, 0.0, -5.0]}},
            "skydrome": sky,
            "primitives": [
                {"type": "sphere", "center": [0, 0, -5], "radius": 3, "bsdf": "null",
                 "int_medium": True},
                {"type": "mesh", "mesh": ground_quad(-2.2), "bsdf": "lambertian",
                 "texture": np.full((1, 1, 3), 0.7, np.float32)},
            ],
        }
    if name == "guided_sphere_quad_1024":
        sc = preset_scene("sphere_renewal_256", width=width or 1024, height=height or 1024)
        sc["spp"] = 256
        # mesh emission is an extension of the reference schema (it ignores the key)
        sc["extensions"] = {"mesh_emission": True}
        sc["primitives"] = sc["primitives"] + [
            # mirror sphere of radius 0.5 in view, 5 units in front of the camera
            {"type": "sphere", "center": [0.9, -0.6, -5.0], "radius": 0.5, "bsdf": "specular"},
            # emissive quad of area 0.01 (0.1 x 0.1), radiance 200, facing down (-y),
            # 25 units above the top of the mean sphere, inside the medium sphere
            {"type": "mesh", "mesh": emissive_quad([0.0, 95.0, -500.0], 0.1), "bsdf": "lambertian",
             "emission": [200.0, 200.0, 200.0]},
        ]
        return sc
    raise ValueError(f"no preset scene named {name!r} (one of {PRESETS})")


def sphere_light_scene(width: int = 256, height: int = 256) -> dict:
    """The sphere_renewal_256 scene lit also by an emissive sphere, given by the
    schema's own sphere "emission" key (the only emitter the reference's scene
    parser creates): radius 40 at (0, 200, -500), above the medium and in view,
    radiance 10.  Every GP-medium hit and Lambertian surface hit takes a light
    sample of it (Sphere::sampleDirect), the medium's through a transmittance
    march."""
    sc = preset_scene("sphere_renewal_256", width=width, height=height)
    sc["primitives"] = sc["primitives"] + [
        {"type": "sphere", "center": [0.0, 200.0, -500.0], "radius": 40.0, "bsdf": "lambertian",
         "emission": [10.0, 10.0, 10.0]}]
    return sc


def sphere_sdf_grid(res: int, bmin, bmax, center, radius: float) -> np.ndarray:
    """A GridMeanFunction payload: |p - center| - radius at the cell centres of a
    res^3 grid over [bmin, bmax], x-major (index (i * res + j) * res + k)."""
    bmin = np.asarray(bmin, np.float64)
    bmax = np.asarray(bmax, np.float64)
    c = [bmin[a] + (np.arange(res) + 0.5) * (bmax[a] - bmin[a]) / res for a in range(3)]
    X, Y, Z = np.meshgrid(c[0], c[1], c[2], indexing="ij")
    d = np.sqrt((X - center[0]) ** 2 + (Y - center[1]) ** 2 + (Z - center[2]) ** 2) - radius
    return d.astype(np.float32).reshape(-1)


def medium_scene(kind: str, *, width: int = 128, height: int = 128, mean: str = "sphere",
                 noise_variance: float = 25.0, res: int = 48) -> dict:
    """The sphere_renewal_256 layout (enclosing null sphere r130 at z = -500, env
    map) with an SDFMedium ("sdf") or ExternalGPMedium ("grid") inside instead of
    the GP medium.  sdf: the mean sphere r70 (or, mean="mesh", a displaced
    icosphere of radius 70); grid: the sphere's signed distance on a res^3 grid
    with a constant variance grid (noise_scale 1)."""
    sc = preset_scene("sphere_renewal_256", width=width, height=height)
    sc.pop("gp_medium")
    sc.pop("kernel")
    sc["spp"] = 16
    if kind == "sdf":
        if mean == "sphere":
            sc["sdf_medium"] = {"mean_type": "sphere", "mean_center": [0, 0, -500], "mean_radius": 70}
        else:
            sc["sdf_medium"] = {"mean_type": "mesh", "mesh": icosphere(3, displacement=0.05),
                                "transform": {"scale": 70.0, "position": [0.0, 0.0, -500.0]}}
    elif kind == "grid":
        bmin, bmax = [-130.0, -130.0, -630.0], [130.0, 130.0, -370.0]
        sc["external_gp_medium"] = {
            "resolution": res, "bounds_min": bmin, "bounds_max": bmax, "noise_scale": 1.0,
            "mean": sphere_sdf_grid(res, bmin, bmax, [0, 0, -500], 70.0),
            "variance": np.full(res ** 3, noise_variance, np.float32)}
    else:
        raise ValueError(kind)
    return sc


def emissive_quad(center, side: float) -> dict:
    """Horizontal square of the given side centred at `center`, face normals -y."""
    cx, cy, cz = (float(c) for c in center)
    h = side / 2
    v = np.array([[cx - h, cy, cz - h], [cx + h, cy, cz - h], [cx + h, cy, cz + h],
                  [cx - h, cy, cz + h]], np.float32)
    f = np.array([[0, 1, 2], [0, 2, 3]], np.int32)
    return {"vertices": v, "faces": f, "normals": None, "texcoords": None, "face_tex": None}


BATCH_SEED = 4


def generated_batch(seed: int = BATCH_SEED, n: int = 64, max_res: int = 512,
                    max_subdivisions: int = 7) -> list[dict]:
    """The batched workload: n seeded renewal_plus scenes, sphere means (the
    sphere_renewal_256 geometry) and displaced-icosphere mesh means (the
    icosphere_renewal_512 geometry) alternating at random; square resolution uniform
    in 64..max_res, spp uniform in 16..128, kernel length_scale (the cell size)
    log-uniform in [0.05, 5], sigma uniform in [0.1, 3], points_per_cell uniform
    in 1..8, icosphere subdivisions uniform in 2..max_subdivisions."""
    rng = np.random.default_rng(seed)
    sky = env_map()
    meshes: dict[int, dict] = {}
    out = []
    for i in range(n):
        res = int(rng.integers(min(64, max_res), max_res + 1))
        spp = int(rng.integers(16, 129))
        kernel = {"sigma": float(rng.uniform(0.1, 3.0)),
                  "length_scale": float(np.exp(rng.uniform(np.log(0.05), np.log(5.0)))),
                  "aniso": [1, 1, 1], "points_per_cell": int(rng.integers(1, 9))}
        if rng.uniform() < 0.5:
            sc = preset_scene("sphere_renewal_256", width=res, height=res)
            sc["kernel"] = kernel
        else:
            sub = int(rng.integers(2, max_subdivisions + 1))
            if sub not in meshes:
                meshes[sub] = icosphere(sub, displacement=0.05)
            sc = {
                "width": res, "height": res, "max_bounce": 10, "max_medium_bounce": 10,
                "gp_mode": "renewal_plus", "camera": dict(CAMERA), "kernel": kernel,
                "gp_medium": {"mean_type": "mesh", "mesh": meshes[sub],
                              "transform": {"scale": 2.0, "position": [0.0, 0.0, -5.0]}},
                "skydrome": sky,
                "primitives": [
                    {"type": "sphere", "center": [0, 0, -5], "radius": 3, "bsdf": "null",
                     "int_medium": True},
                    {"type": "mesh", "mesh": ground_quad(-2.2), "bsdf": "lambertian",
                     "texture": np.full((1, 1, 3), 0.7, np.float32)},
                ],
            }
        sc["spp"] = spp
        sc["name"] = f"batch{i:02d}"
        out.append(sc)
    return out



def chain_mesh(n: int = 70, ratio: float = 2.05, x0: float = 1e-9) -> dict:
    """n flat triangles facing +x at x_i = x0 ratio^i, each spanning 0.1 x_i in y
    and z: every centroid-midpoint split of the reference's BVH build peels off
    the farthest triangle, so the tree is about n levels deep (a pathological
    case for the 64-entry traversal stack)."""
    x = x0 * ratio ** np.arange(n, dtype=np.float64)
    v = np.zeros((3 * n, 3), np.float64)
    v[0::3, 0] = x
    v[1::3, 0] = x
    v[1::3, 1] = 0.1 * x
    v[2::3, 0] = x
    v[2::3, 2] = 0.1 * x
    f = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    return {"vertices": v.astype(np.float32), "faces": f, "normals": None, "texcoords": None,
            "face_tex": None}
