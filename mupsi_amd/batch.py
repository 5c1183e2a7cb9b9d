"""Batch entry points of the C ABI on HBM-resident torch tensors.

Each function takes/returns CUDA(HIP) tensors and passes their device
pointers straight to libmupsi_hip.so (no host staging).  torch is only the
device-memory allocator here.
"""
from __future__ import annotations

import ctypes as C

from . import _abi as A
from ._abi import check


def _torch():
    import torch
    return torch


def _dev(t):
    """Device pointer of a contiguous device tensor.  The library runs on its own
    stream, so the work torch has queued on the tensor's stream (the copy that
    made it contiguous, a fill) is waited for first."""
    torch = _torch()
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise ValueError("expected a contiguous device tensor")
    torch.cuda.current_stream(t.device).synchronize()
    return C.c_void_p(t.data_ptr())


_KINDS = {"f32": ("float32",), "i32": ("int32", "uint32"), "u8": ("uint8", "bool")}


def _check(t, kind: str, n: int, cols: int | None, device, name: str):
    """Refuses a tensor the kernels would misread: a dtype of another width or
    kind (f32: float32; i32: 32-bit integers; u8: uint8), fewer or more rows
    than n, a row width other than cols, or another device.  The entry points
    read n (x cols) elements through raw pointers, so a short or narrower
    array would be read past its end."""
    torch = _torch()
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name}: expected a torch tensor, got {type(t).__name__}")
    dt = str(t.dtype).replace("torch.", "")
    if dt not in _KINDS[kind]:
        raise TypeError(f"{name}: dtype {dt} where {'/'.join(_KINDS[kind])} is required")
    shape = (n,) if cols is None else (n, cols)
    if tuple(t.shape) != shape:
        raise ValueError(f"{name}: shape {tuple(t.shape)} where {shape} is required")
    if device is not None and t.device != device:
        raise ValueError(f"{name}: on {t.device}, the other arrays are on {device}")
    return t


def gp_params(kernel: dict) -> A.GpParams:
    g = A.GpParams()
    g.sigma = float(kernel.get("sigma", 1.0))
    g.kernel_radius = float(kernel.get("length_scale", 1.0))
    g.aniso[:] = [float(x) for x in kernel.get("aniso", [1, 1, 1])]
    g.points_per_cell = int(kernel.get("points_per_cell", 3))
    return g


def _cond(cond, n: int, device):
    """cond: None or dict of device tensors active(u8), C(n,3), u(n), g(n,3)."""
    if cond is None:
        return None, []
    torch = _torch()
    _check(cond["active"], "u8", n, None, device, "cond.active")
    _check(cond["C"], "f32", n, 3, device, "cond.C")
    _check(cond["u"], "f32", n, None, device, "cond.u")
    _check(cond["g"], "f32", n, 3, device, "cond.g")
    keep = [cond["active"].to(torch.uint8).contiguous()]
    Cc = cond["C"].float()
    g = cond["g"].float()
    cols = [Cc[:, 0].contiguous(), Cc[:, 1].contiguous(), Cc[:, 2].contiguous(),
            cond["u"].float().contiguous(), g[:, 0].contiguous(), g[:, 1].contiguous(),
            g[:, 2].contiguous()]
    keep += cols
    s = A.CondSoA(*[_dev(x) for x in keep])
    return s, keep


def psi_eval(ctx, kernel: dict, p, seed):
    torch = _torch()
    n = p.shape[0]
    _check(p, "f32", n, 3, None, "p")
    _check(seed, "i32", n, None, p.device, "seed")
    px, py, pz = (p[:, i].contiguous() for i in range(3))
    out = torch.empty((n, 4), dtype=torch.float32, device=p.device)
    l1 = torch.empty(n, dtype=torch.float32, device=p.device)
    g = gp_params(kernel)
    check(ctx.lib.mupsi_psi_eval(ctx.handle, C.byref(g), _dev(px), _dev(py), _dev(pz),
                                 _dev(seed), n, _dev(out), _dev(l1)), "mupsi_psi_eval")
    return out, l1


def cell_impulses(ctx, kernel: dict, cells, seed):
    torch = _torch()
    n = cells.shape[0]
    _check(cells, "i32", n, 3, None, "cells")
    _check(seed, "i32", n, None, cells.device, "seed")
    ppc = int(kernel.get("points_per_cell", 3))
    cs = torch.empty(n, dtype=torch.int32, device=cells.device)
    pos = torch.empty((n, ppc, 3), dtype=torch.float32, device=cells.device)
    w = torch.empty((n, ppc), dtype=torch.int8, device=cells.device)
    g = gp_params(kernel)
    check(ctx.lib.mupsi_cell_impulses(ctx.handle, C.byref(g), _dev(cells.contiguous()), _dev(seed),
                                      n, _dev(cs), _dev(pos), _dev(w)), "mupsi_cell_impulses")
    return cs, pos, w


def field_eval(scene, p, seed, cond=None):
    torch = _torch()
    n = p.shape[0]
    _check(p, "f32", n, 3, None, "p")
    _check(seed, "i32", n, None, p.device, "seed")
    px, py, pz = (p[:, i].contiguous() for i in range(3))
    out = torch.empty((n, 4), dtype=torch.float32, device=p.device)
    cs, keep = _cond(cond, n, p.device)
    check(scene.lib.mupsi_field_eval(scene.handle, _dev(px), _dev(py), _dev(pz), _dev(seed),
                                     C.byref(cs) if cs is not None else None, n, _dev(out)),
          "mupsi_field_eval")
    return out


def march(scene, rays, seed, cond=None, mode: int = 0, noalloc: bool = False):
    """rays: (n, 8) device tensor {ox, oy, oz, dx, dy, dz, near_t, far_t}.
    noalloc: run mupsi_march_noalloc (the library's fixed scratch, at most
    MARCH_NOALLOC_MAX_RAYS rays) instead of mupsi_march."""
    torch = _torch()
    n = rays.shape[0]
    _check(rays, "f32", n, 8, None, "rays")
    _check(seed, "i32", n, None, rays.device, "seed")
    cols = [rays[:, i].contiguous() for i in range(8)]
    dev = rays.device
    hit = torch.empty(n, dtype=torch.uint8, device=dev)
    t = torch.empty(n, dtype=torch.float32, device=dev)
    p = torch.empty((n, 3), dtype=torch.float32, device=dev)
    nrm = torch.empty((n, 3), dtype=torch.float32, device=dev)
    steps = torch.empty(n, dtype=torch.int32, device=dev)
    cs, keep = _cond(cond, n, dev)
    name = "mupsi_march_noalloc" if noalloc else "mupsi_march"
    check(getattr(scene.lib, name)(scene.handle, mode, *[_dev(c) for c in cols], _dev(seed),
                                   C.byref(cs) if cs is not None else None, n, _dev(hit), _dev(t),
                                   _dev(p), _dev(nrm), _dev(steps)), name)
    return {"hit": hit, "t": t, "p": p, "n": nrm, "steps": steps}


def condition(scene, C3, N3, seed_next):
    torch = _torch()
    n = C3.shape[0]
    _check(C3, "f32", n, 3, None, "C")
    _check(N3, "f32", n, 3, C3.device, "n")
    _check(seed_next, "i32", n, None, C3.device, "seed_next")
    cols = [C3[:, i].contiguous() for i in range(3)] + [N3[:, i].contiguous() for i in range(3)]
    u = torch.empty(n, dtype=torch.float64, device=C3.device)
    g = torch.empty((n, 3), dtype=torch.float64, device=C3.device)
    check(scene.lib.mupsi_condition(scene.handle, *[_dev(c) for c in cols], _dev(seed_next), n,
                                    _dev(u), _dev(g)), "mupsi_condition")
    return u, g


def mesh_intersect(scene, prim: int, rays):
    torch = _torch()
    n = rays.shape[0]
    _check(rays, "f32", n, 8, None, "rays")
    cols = [rays[:, i].contiguous() for i in range(8)]
    tri = torch.empty(n, dtype=torch.int32, device=rays.device)
    t, u, v = (torch.empty(n, dtype=torch.float32, device=rays.device) for _ in range(3))
    check(scene.lib.mupsi_mesh_intersect(scene.handle, prim, *[_dev(c) for c in cols], n, _dev(tri),
                                         _dev(t), _dev(u), _dev(v)), "mupsi_mesh_intersect")
    return tri, t, u, v


def mesh_occluded(scene, prim: int, rays):
    torch = _torch()
    n = rays.shape[0]
    _check(rays, "f32", n, 8, None, "rays")
    cols = [rays[:, i].contiguous() for i in range(8)]
    occ = torch.empty(n, dtype=torch.uint8, device=rays.device)
    check(scene.lib.mupsi_mesh_occluded(scene.handle, prim, *[_dev(c) for c in cols], n, _dev(occ)),
          "mupsi_mesh_occluded")
    return occ


def mesh_closest(scene, prim: int, p):
    torch = _torch()
    n = p.shape[0]
    _check(p, "f32", n, 3, None, "p")
    px, py, pz = (p[:, i].contiguous() for i in range(3))
    d = torch.empty(n, dtype=torch.float32, device=p.device)
    tri = torch.empty(n, dtype=torch.int32, device=p.device)
    cl = torch.empty((n, 3), dtype=torch.float32, device=p.device)
    check(scene.lib.mupsi_mesh_closest(scene.handle, prim, _dev(px), _dev(py), _dev(pz), n, _dev(d),
                                       _dev(tri), _dev(cl)), "mupsi_mesh_closest")
    return d, tri, cl


def mesh_inside(scene, prim: int, p):
    torch = _torch()
    n = p.shape[0]
    _check(p, "f32", n, 3, None, "p")
    px, py, pz = (p[:, i].contiguous() for i in range(3))
    out = torch.empty(n, dtype=torch.uint8, device=p.device)
    check(scene.lib.mupsi_mesh_inside(scene.handle, prim, _dev(px), _dev(py), _dev(pz), n, _dev(out)),
          "mupsi_mesh_inside")
    return out


def mesh_query_cost(scene, prim: int, kind: int, p):
    """Per point: BVH nodes tested and triangles tested by distToPoint (kind 0),
    inside as the reference's loop of casts (kind 1), distToPoint from the
    distance-grid bound (kind 2), inside as the library runs it (kind 3: over
    the point's leaf bin where the mesh has leaf bins, a node count being a bin
    entry read) or inside as the stackless walk over the tree (kind 4);
    int64 tensors (n,)."""
    torch = _torch()
    n = p.shape[0]
    _check(p, "f32", n, 3, None, "p")
    px, py, pz = (p[:, i].contiguous() for i in range(3))
    nodes = torch.empty(n, dtype=torch.int32, device=p.device)
    tris = torch.empty(n, dtype=torch.int32, device=p.device)
    check(scene.lib.mupsi_mesh_query_cost(scene.handle, prim, kind, _dev(px), _dev(py), _dev(pz), n,
                                          _dev(nodes), _dev(tris)), "mupsi_mesh_query_cost")
    return nodes.long(), tris.long()


def last_kernel_ms(ctx) -> float:
    """Device time of the kernel run by the last mesh query on this context."""
    ms = C.c_float(0.0)
    check(ctx.lib.mupsi_last_kernel_ms(ctx.handle, C.byref(ms)), "mupsi_last_kernel_ms")
    return float(ms.value)


# ---------------------------------------------------------------- guiding network

def guide_forward(guide, feats):
    """feats (n, 32) f32 -> (raw (n, 41), mixture (n, 44)) from the MFMA forward kernel."""
    torch = _torch()
    n = feats.shape[0]
    raw = torch.empty((n, 41), dtype=torch.float32, device=feats.device)
    mix = torch.empty((n, 44), dtype=torch.float32, device=feats.device)
    check(guide.lib.mupsi_guide_forward(guide.handle, _dev(feats), n, _dev(raw), _dev(mix)),
          "mupsi_guide_forward")
    return raw, mix


def guide_gradient(guide, feats, dirs, weights):
    """Mean loss and its gradient (numpy, 13032) over the batch."""
    import numpy as np
    grad = np.zeros(A.GUIDE_NUM_PARAMS, np.float32)
    loss = C.c_double()
    check(guide.lib.mupsi_guide_gradient(guide.handle, _dev(feats), _dev(dirs), _dev(weights),
                                         feats.shape[0], grad.ctypes.data_as(C.c_void_p),
                                         C.byref(loss)), "mupsi_guide_gradient")
    return loss.value, grad


def guide_train_step(guide, feats, dirs, weights):
    check(guide.lib.mupsi_guide_train_step(guide.handle, _dev(feats), _dev(dirs), _dev(weights),
                                           feats.shape[0]), "mupsi_guide_train_step")


def guide_encode(guide, p, nrm, wo):
    torch = _torch()
    out = torch.empty((p.shape[0], 32), dtype=torch.float32, device=p.device)
    check(guide.lib.mupsi_guide_encode(guide.handle, _dev(p), _dev(nrm), _dev(wo), p.shape[0],
                                       _dev(out)), "mupsi_guide_encode")
    return out


def guide_sample(guide, mixture, nrm, u):
    """The tracer's guided Lambertian sample on explicit uniforms u (n, 4)."""
    torch = _torch()
    n = mixture.shape[0]
    d = torch.empty((n, 3), dtype=torch.float32, device=mixture.device)
    q = torch.empty(n, dtype=torch.float32, device=mixture.device)
    w = torch.empty(n, dtype=torch.float32, device=mixture.device)
    check(guide.lib.mupsi_guide_sample(guide.handle, _dev(mixture), _dev(nrm), _dev(u), n,
                                       _dev(d), _dev(q), _dev(w)), "mupsi_guide_sample")
    return d, q, w
