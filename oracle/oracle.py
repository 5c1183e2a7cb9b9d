"""ctypes wrapper of liboracle.so — TEST INFRASTRUCTURE ONLY.

The CPU restatement of the mupsi hot path (oracle.cpp) used as the checker by
tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  Scene
descriptions are the same plain structs the product takes (include/mupsi.h).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_random.so"

VP, I32, I64 = C.c_void_p, C.c_int32, C.c_int64

_SIGS = {
    "orc_make_seed": (C.c_uint32, [VP, I32]),
    "orc_pcg_draws": (None, [C.c_uint64, I32, I32, VP]),
    "orc_normalized_uint": (C.c_float, [C.c_uint32]),
    "orc_constant_sampler_i": (C.c_uint32, [C.c_uint32]),
    "orc_cell_impulses": (None, [VP, VP, VP, I64, VP, VP, VP]),
    "orc_psi_eval": (None, [VP, VP, VP, I64, I32, VP, VP]),
    "orc_psi_var": (C.c_float, [VP]),
    "orc_scene_create": (VP, [VP]),
    "orc_scene_free": (None, [VP]),
    "orc_field_eval": (None, [VP, VP, VP, VP, I64, I32, VP]),
    "orc_march": (None, [VP, I32, VP, VP, VP, I64, I32, VP, VP, VP, VP, VP, VP, VP]),
    "orc_condition": (None, [VP, VP, VP, VP, I64, I32, VP, VP]),
    "orc_mean_eval": (None, [VP, VP, I64, VP]),
    "orc_bvh_build": (I32, [VP, I32, VP, I32, VP, VP, VP]),
    "orc_mesh_intersect": (None, [VP, I32, VP, I64, VP, VP, VP, VP]),
    "orc_mesh_intersect_brute": (None, [VP, I32, VP, I64, VP, VP]),
    "orc_mesh_occluded": (None, [VP, I32, VP, I64, VP]),
    "orc_mesh_closest": (None, [VP, I32, VP, I64, VP, VP, VP]),
    "orc_mesh_inside": (None, [VP, I32, VP, I64, VP]),
    "orc_mesh_query_cost": (None, [VP, I32, I32, VP, I64, VP, VP]),
    "orc_hash11": (None, [VP, VP, I64, VP]),
    "orc_camera_rays": (None, [VP, VP, VP, I64, VP]),
    "orc_sky_lookup": (None, [VP, VP, I64, VP, VP]),
    "orc_trace_paths": (None, [VP, VP, VP, VP, I64, VP, I32]),
    "orc_trace_paths_amb": (None, [VP, VP, VP, VP, I64, VP, VP, I32]),
    "orc_choose_light": (None, [VP, VP, VP, I64, VP, VP, VP, VP, VP]),
    "orc_render": (C.c_double, [VP, I32, I32, I32, I32, I32, VP, VP]),
    "orc_max_threads": (I32, []),
}

_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        _lib = C.CDLL(str(LIB))
        for n, (r, a) in _SIGS.items():
            f = getattr(_lib, n)
            f.restype = r
            f.argtypes = a
    return _lib


def ref_lib():
    """The reference's own random.h compiled into oracle/_ref (None if not built)."""
    if not REF_LIB.exists():
        return None
    L = C.CDLL(str(REF_LIB))
    L.ref_make_seed4.restype = C.c_uint32
    L.ref_make_seed4.argtypes = [I32, I32, I32, C.c_uint32]
    L.ref_make_seed5.restype = C.c_uint32
    L.ref_make_seed5.argtypes = [I32, I32, I32, C.c_uint32, I32]
    L.ref_random_ctor_draws.argtypes = [C.c_uint64, I32, VP]
    L.ref_random_state_draws.argtypes = [C.c_uint64, I32, VP]
    L.ref_normalized_uint.restype = C.c_float
    L.ref_normalized_uint.argtypes = [C.c_uint32]
    L.ref_bernoulli.restype = C.c_float
    L.ref_bernoulli.argtypes = [C.c_float]
    return L


def _p(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(VP)


def make_seed(*vals) -> int:
    a = np.array([v & 0xFFFFFFFF for v in vals], np.uint32)
    return int(lib().orc_make_seed(_p(a), len(vals)))


def pcg_draws(seed: int, n: int, ctor: bool = True) -> np.ndarray:
    out = np.zeros(n, np.uint32)
    lib().orc_pcg_draws(seed, int(ctor), n, _p(out))
    return out


def constant_sampler_i(seed: int) -> int:
    return int(lib().orc_constant_sampler_i(seed))


def gp_params(kernel: dict):
    from mupsi_amd import _abi as A
    g = A.GpParams()
    g.sigma = float(kernel.get("sigma", 1.0))
    g.kernel_radius = float(kernel.get("length_scale", 1.0))
    g.aniso[:] = [float(x) for x in kernel.get("aniso", [1, 1, 1])]
    g.points_per_cell = int(kernel.get("points_per_cell", 3))
    return g


def psi_var(kernel: dict) -> float:
    g = gp_params(kernel)
    return float(lib().orc_psi_var(C.byref(g)))


def cell_impulses(kernel: dict, cells: np.ndarray, seeds: np.ndarray):
    g = gp_params(kernel)
    cells = np.ascontiguousarray(cells, np.int32).reshape(-1, 3)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    n, ppc = cells.shape[0], g.points_per_cell
    cs = np.zeros(n, np.uint32)
    pos = np.zeros((n, ppc, 3), np.float32)
    w = np.zeros((n, ppc), np.int8)
    lib().orc_cell_impulses(C.byref(g), _p(cells), _p(seeds), n, _p(cs), _p(pos), _p(w))
    return cs, pos, w


def psi_eval(kernel: dict, p: np.ndarray, seeds: np.ndarray, fp64: bool = True):
    g = gp_params(kernel)
    p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros((p.shape[0], 4), np.float64)
    l1 = np.zeros(p.shape[0], np.float64)
    lib().orc_psi_eval(C.byref(g), _p(p), _p(seeds), p.shape[0], int(fp64), _p(out), _p(l1))
    return out, l1


class OracleScene:
    """orc_scene_create over the same scene dict the product takes."""

    def __init__(self, scene: dict):
        from mupsi_amd.scene import build_desc
        desc, self._holder = build_desc(scene)
        self.h = lib().orc_scene_create(C.byref(desc))
        self.width, self.height = desc.width, desc.height

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_scene_free(self.h)
            self.h = None

    @staticmethod
    def _cond(cond: np.ndarray | None, n: int):
        if cond is None:
            return None
        c = np.ascontiguousarray(cond, np.float32).reshape(n, 8)
        return c

    def field_eval(self, p, seeds, cond=None, fp64=True):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        n = p.shape[0]
        seeds = np.ascontiguousarray(seeds, np.uint32)
        c = self._cond(cond, n)
        out = np.zeros((n, 4), np.float64)
        lib().orc_field_eval(self.h, _p(p), _p(seeds), _p(c), n, int(fp64), _p(out))
        return out

    def mean_eval(self, p):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        out = np.zeros((p.shape[0], 4), np.float64)
        lib().orc_mean_eval(self.h, _p(p), p.shape[0], _p(out))
        return out

    def march(self, rays, seeds, cond=None, mode=0, fp64=False, track=True):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        seeds = np.ascontiguousarray(seeds, np.uint32)
        c = self._cond(cond, n)
        hit = np.zeros(n, np.uint8)
        t = np.zeros(n, np.float32)
        p3 = np.zeros((n, 3), np.float32)
        n3 = np.zeros((n, 3), np.float32)
        steps = np.zeros(n, np.int32)
        amb = np.zeros(n, np.float64) if track else None
        cell = np.zeros(n, np.float64) if track else None
        lib().orc_march(self.h, mode, _p(rays), _p(seeds), _p(c), n, int(fp64), _p(hit), _p(t),
                        _p(p3), _p(n3), _p(steps), _p(amb), _p(cell))
        return {"hit": hit, "t": t, "p": p3, "n": n3, "steps": steps, "amb": amb, "cell": cell}

    def condition(self, C3, N3, seed_next, fp64=True):
        C3 = np.ascontiguousarray(C3, np.float32).reshape(-1, 3)
        N3 = np.ascontiguousarray(N3, np.float32).reshape(-1, 3)
        n = C3.shape[0]
        seed_next = np.ascontiguousarray(seed_next, np.uint32)
        u = np.zeros(n, np.float64)
        g = np.zeros((n, 3), np.float64)
        lib().orc_condition(self.h, _p(C3), _p(N3), _p(seed_next), n, int(fp64), _p(u), _p(g))
        return u, g

    def mesh_intersect(self, prim, rays, brute=False):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        n = rays.shape[0]
        tri = np.zeros(n, np.int32)
        t = np.zeros(n, np.float32)
        if brute:
            lib().orc_mesh_intersect_brute(self.h, prim, _p(rays), n, _p(tri), _p(t))
            return tri, t
        u = np.zeros(n, np.float32)
        v = np.zeros(n, np.float32)
        lib().orc_mesh_intersect(self.h, prim, _p(rays), n, _p(tri), _p(t), _p(u), _p(v))
        return tri, t, u, v

    def mesh_occluded(self, prim, rays):
        rays = np.ascontiguousarray(rays, np.float32).reshape(-1, 8)
        occ = np.zeros(rays.shape[0], np.uint8)
        lib().orc_mesh_occluded(self.h, prim, _p(rays), rays.shape[0], _p(occ))
        return occ

    def mesh_closest(self, prim, p):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        n = p.shape[0]
        d = np.zeros(n, np.float32)
        tri = np.zeros(n, np.int32)
        cl = np.zeros((n, 3), np.float32)
        lib().orc_mesh_closest(self.h, prim, _p(p), n, _p(d), _p(tri), _p(cl))
        return d, tri, cl

    def mesh_query_cost(self, prim, kind, p):
        """Per point (n, 3): BVH nodes and triangles tested by distToPoint (kind 0)
        or inside (kind 1)."""
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        n = p.shape[0]
        nodes = np.zeros(n, dtype=np.uint32)
        tris = np.zeros(n, dtype=np.uint32)
        lib().orc_mesh_query_cost(self.h, prim, kind, _p(p), n, _p(nodes), _p(tris))
        return nodes.astype(np.int64), tris.astype(np.int64)

    def mesh_inside(self, prim, p):
        p = np.ascontiguousarray(p, np.float32).reshape(-1, 3)
        out = np.zeros(p.shape[0], np.uint8)
        lib().orc_mesh_inside(self.h, prim, _p(p), p.shape[0], _p(out))
        return out

    def camera_rays(self, x, y):
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        out = np.zeros((x.size, 3), np.float32)
        lib().orc_camera_rays(self.h, _p(x), _p(y), x.size, _p(out))
        return out

    def sky_lookup(self, d):
        d = np.ascontiguousarray(d, np.float32).reshape(-1, 3)
        rgb = np.zeros_like(d)
        tx = np.zeros(d.shape[0], np.int32)
        lib().orc_sky_lookup(self.h, _p(d), d.shape[0], _p(rgb), _p(tx))
        return rgb, tx

    def trace_paths(self, x, y, k, nthreads=0):
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        out = np.zeros((x.size, 3), np.float32)
        lib().orc_trace_paths(self.h, _p(x), _p(y), _p(k), x.size, _p(out), nthreads)
        return out

    def trace_paths_amb(self, x, y, k, nthreads=0):
        """trace_paths plus each path's smallest GP-march decision margin
        (FP64 |f| / (1 + |mu| + l1) at its visited grid points; 0 where the
        FP64 march decides otherwise; inf without a GP march)."""
        x = np.ascontiguousarray(x, np.int32)
        y = np.ascontiguousarray(y, np.int32)
        k = np.ascontiguousarray(k, np.int32)
        out = np.zeros((x.size, 3), np.float32)
        amb = np.zeros(x.size, np.float64)
        lib().orc_trace_paths_amb(self.h, _p(x), _p(y), _p(k), x.size, _p(out), _p(amb), nthreads)
        return out, amb

    def choose_light(self, p3, states):
        """Scene::chooseLight at points p3 (n, 3) from PCG states (n,) uint64
        (setState semantics).  Returns a dict: light (primitive index; -1 no
        draw, -2 sampleDirect failed), w (n, 3), d (n, 3), dist, pdf, state
        (the PCG state after the draws)."""
        p3 = np.ascontiguousarray(p3, np.float32).reshape(-1, 3)
        n = p3.shape[0]
        st = np.array(np.broadcast_to(np.asarray(states, np.uint64), (n,)))
        light = np.zeros(n, np.int32)
        w = np.zeros((n, 3), np.float32)
        d = np.zeros((n, 3), np.float32)
        dist = np.zeros(n, np.float32)
        pdf = np.zeros(n, np.float32)
        lib().orc_choose_light(self.h, _p(p3), _p(st), n, _p(light), _p(w), _p(d), _p(dist), _p(pdf))
        return {"light": light, "w": w, "d": d, "dist": dist, "pdf": pdf, "state": st}

    def render(self, k_begin, k_count, pixel_stride=1, pixel_offset=0, nthreads=0):
        sums = np.zeros((self.height, self.width, 3), np.float64)
        paths = C.c_int64()
        secs = lib().orc_render(self.h, k_begin, k_count, pixel_stride, pixel_offset, nthreads,
                                _p(sums), C.byref(paths))
        return sums, paths.value, secs


def hash11(p3: np.ndarray, seeds: np.ndarray) -> np.ndarray:
    """hash11 of the external GP medium at points p3 (n, 3) with seeds (n,)."""
    p3 = np.ascontiguousarray(p3, np.float32).reshape(-1, 3)
    seeds = np.ascontiguousarray(seeds, np.uint32)
    out = np.zeros(p3.shape[0], np.float32)
    lib().orc_hash11(_p(p3), _p(seeds), p3.shape[0], _p(out))
    return out


def bvh_build(vertices: np.ndarray, faces: np.ndarray):
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1, 3)
    f = np.ascontiguousarray(faces, np.int32).reshape(-1, 3)
    nodes = np.zeros((2 * max(1, f.shape[0]), 8), np.float32)
    tri = np.zeros(max(1, f.shape[0]), np.uint32)
    root = np.zeros(1, np.uint32)
    nn = lib().orc_bvh_build(_p(v), v.shape[0], _p(f), f.shape[0], _p(nodes), _p(tri), _p(root))
    return nodes[:nn], tri[:f.shape[0]], int(root[0])
