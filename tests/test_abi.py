"""The C-ABI library loads, exports every symbol include/mupsi.h declares, and its
host-side entry points behave without a GPU."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from mupsi_amd import _abi, scenes
from oracle import oracle as O

HEADER = Path(__file__).resolve().parents[1] / "include" / "mupsi.h"


def header_functions() -> list[str]:
    text = HEADER.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mupsi_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    lib = _abi.load()
    names = header_functions()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), f"{n} missing from {_abi.LIB_PATH}"
        assert n in _abi.SIGNATURES, f"{n} has no ctypes signature"


def test_struct_sizes_match_header_layout():
    # 64-bit layout of the descriptor structs (pointers 8-byte aligned)
    assert C.sizeof(_abi.GpParams) == 24
    assert C.sizeof(_abi.MeshDesc) == 5 * 8 + 4 * 4 + 3 * 4 + 4
    assert C.sizeof(_abi.TextureDesc) == 16
    assert C.sizeof(_abi.Stats) == 64


def test_status_strings():
    lib = _abi.load()
    assert lib.mupsi_status_string(0) == b"ok"
    assert lib.mupsi_status_string(4) == b"aborted"


def test_open_without_gpu_fails_cleanly():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = _abi.load()
    h = C.c_void_p()
    st = lib.mupsi_open(0, C.byref(h))
    assert st == 2 and not h.value
    assert lib.mupsi_last_error()


def _bvh_product(v, f):
    lib = _abi.load()
    v = np.ascontiguousarray(v, np.float32)
    f = np.ascontiguousarray(f, np.int32)
    nodes = np.zeros((2 * len(f), 8), np.float32)
    tri = np.zeros(len(f), np.uint32)
    nn = C.c_int32()
    root = C.c_uint32()
    st = lib.mupsi_bvh_build(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p),
                             len(f), nodes.ctypes.data_as(C.c_void_p), C.byref(nn),
                             tri.ctypes.data_as(C.c_void_p), C.byref(root))
    assert st == 0
    return nodes[: nn.value], tri, root.value


@pytest.mark.parametrize("sub", [0, 2, 4])
def test_bvh_topology_equals_oracle(sub):
    m = scenes.icosphere(sub, displacement=0.05 if sub else 0.0)
    a = _bvh_product(m["vertices"], m["faces"])
    b = O.bvh_build(m["vertices"], m["faces"])
    assert a[2] == b[2] == len(a[0]) - 1          # post-order: the root is the last node
    assert np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32))
    assert np.array_equal(a[1], b[1])


def test_bvh_build_rejects_bad_faces():
    lib = _abi.load()
    v = np.zeros((3, 3), np.float32)
    f = np.array([[0, 1, 7]], np.int32)
    nn = C.c_int32()
    st = lib.mupsi_bvh_build(v.ctypes.data_as(C.c_void_p), 3, f.ctypes.data_as(C.c_void_p), 1,
                             None, C.byref(nn), None, None)
    assert st == 1


def _layout(m):
    lib = _abi.load()
    v = np.ascontiguousarray(m["vertices"], np.float32)
    f = np.ascontiguousarray(m["faces"], np.int32)
    nr, depth, sl = C.c_int32(), C.c_int32(), C.c_int32()
    st = lib.mupsi_bvh_layout(v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p),
                              len(f), C.byref(nr), C.byref(depth), C.byref(sl))
    assert st == 0
    return nr.value, depth.value, sl.value


def _ref_depth(nodes, root):
    """Levels of internal nodes on the longest root-to-leaf path of a post-order tree."""
    u = nodes.view(np.uint32)
    best, st = 0, [(root, 0)]
    while st:
        i, d = st.pop()
        if u[i, 7] & 0x80000000:
            best = max(best, d)
        else:
            st += [(u[i, 6], d + 1), (u[i, 7], d + 1)]
    return best


@pytest.mark.parametrize("which", ["ico4", "ico7", "chain"])
def test_bvh_device_layout(which):
    """One 64-byte record per internal node of the reference tree; the walk
    without a stack only for trees of at most 60 levels (a 70-triangle chain
    built by the midpoint split is deeper and keeps the reference's casts)."""
    m = scenes.chain_mesh() if which == "chain" else scenes.icosphere(int(which[3:]), displacement=0.05)
    nodes, _, root = O.bvh_build(m["vertices"], m["faces"])
    internal = int(np.count_nonzero((nodes.view(np.uint32)[:, 7] & 0x80000000) == 0))
    nr, depth, sl = _layout(m)
    assert nr == internal
    assert depth == _ref_depth(nodes, root)
    assert sl == (1 if depth <= 60 else 0)
    assert (which == "chain") == (sl == 0)
    if which == "chain":
        assert depth > 60


def _xbins(m):
    lib = _abi.load()
    v = np.ascontiguousarray(m["vertices"], np.float32)
    f = np.ascontiguousarray(m["faces"], np.int32)
    ny, nz, ne = C.c_int32(), C.c_int32(), C.c_int64()
    oi = np.zeros(3, np.float32)
    args = (v.ctypes.data_as(C.c_void_p), len(v), f.ctypes.data_as(C.c_void_p), len(f),
            C.byref(ny), C.byref(nz), oi.ctypes.data_as(C.c_void_p), C.byref(ne))
    assert lib.mupsi_bvh_xbins(*args, None, 0, None, 0) == 0
    ent = np.zeros((max(ne.value, 1), 8), np.float32)
    off = np.zeros(ny.value * nz.value + 1, np.uint32)
    assert lib.mupsi_bvh_xbins(*args, ent.ctypes.data_as(C.c_void_p), ne.value,
                               off.ctypes.data_as(C.c_void_p), len(off)) == 0
    return ny.value, nz.value, oi, ent[: ne.value], off


@pytest.mark.parametrize("sub", [2, 5])
def test_inside_leaf_bins(sub):
    """The leaf bins of Mesh::inside: the bin of every point (y, z) computed in
    float32 as the kernels compute it, (y - y0) * 2^-e truncated, lists every
    leaf whose y-z bounds hold the point, the points including each leaf's
    corners and values one ulp either side of them; a bin's leaves are sorted by
    max x, largest first; every entry is a leaf of the reference tree with its
    bounds; a tree deeper than the stackless walk allows gets no bins."""
    m = scenes.icosphere(sub, displacement=0.05)
    nodes, _, root = O.bvh_build(m["vertices"], m["faces"])
    ny, nz, oi, ent, off = _xbins(m)
    assert ny > 0 and nz > 0 and off[-1] == len(ent)
    y0, z0, inv = (np.float32(x) for x in oi)
    assert inv > 0 and np.frexp(inv)[0] == 0.5          # a power of two
    u = nodes.view(np.uint32)
    leaf = (u[:, 7] & 0x80000000) != 0
    lb = nodes[leaf, :6]
    lref = (0x80000000 | ((u[leaf, 7] & 0x7FFFFFFF) << 28) | u[leaf, 6]).astype(np.uint32)
    key = {(int(r), tuple(b)) for r, b in zip(lref, lb.tolist())}
    eref = ent[:, 6].view(np.uint32)
    ebox = np.stack([ent[:, 5], ent[:, 1], ent[:, 3], ent[:, 0], ent[:, 2], ent[:, 4]], 1)
    assert {(int(r), tuple(b)) for r, b in zip(eref, ebox.tolist())} == key
    for b in range(len(off) - 1):
        xs = ent[off[b]:off[b + 1], 0]
        assert np.all(xs[:-1] >= xs[1:])
    rng = np.random.default_rng(sub)
    pick = rng.choice(len(lb), min(len(lb), 400), replace=False)
    ys, zs = [], []
    for i in pick:
        for yv in (lb[i, 1], lb[i, 4]):
            for zv in (lb[i, 2], lb[i, 5]):
                for dy in (-np.inf, 0, np.inf):
                    for dz in (-np.inf, 0, np.inf):
                        ys.append(np.nextafter(yv, np.float32(dy)) if dy else yv)
                        zs.append(np.nextafter(zv, np.float32(dz)) if dz else zv)
    lo, hi = nodes[root, :3], nodes[root, 3:6]
    ys = np.concatenate([np.array(ys, np.float32), rng.uniform(lo[1], hi[1], 4000).astype(np.float32)])
    zs = np.concatenate([np.array(zs, np.float32), rng.uniform(lo[2], hi[2], 4000).astype(np.float32)])
    fy = (ys - y0) * inv
    fz = (zs - z0) * inv
    assert fy.dtype == np.float32 and fz.dtype == np.float32
    for y, z, a, c in zip(ys, zs, fy, fz):
        holds = (lb[:, 1] <= y) & (y <= lb[:, 4]) & (lb[:, 2] <= z) & (z <= lb[:, 5])
        if not (a >= 0 and c >= 0 and a < ny and c < nz):
            assert not holds.any()
            continue
        b = int(c) * ny + int(a)
        listed = set(eref[off[b]:off[b + 1]].tolist())
        assert set(lref[holds].tolist()) <= listed
    assert _xbins(scenes.chain_mesh())[:2] == (0, 0)


def test_inside_leaf_bins_flat_mesh():
    """A mesh flat in y (a 12 x 12 grid of quads in the x-z plane): one row of
    bins, every leaf listed in the bins its z bounds reach, and a point on the
    plane finds every leaf holding it in its bin."""
    n = 12
    xs, zs = np.meshgrid(np.linspace(-1, 1, n + 1), np.linspace(-1, 1, n + 1), indexing="ij")
    v = np.stack([xs.ravel(), np.zeros(xs.size), zs.ravel()], 1).astype(np.float32)
    f = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1, i * (n + 1) + j + 1
            f += [[a, b, c], [a, c, d]]
    m = {"vertices": v, "faces": np.array(f, np.int32)}
    ny, nz, oi, ent, off = _xbins(m)
    assert ny == 1 and nz > 1 and off[-1] == len(ent) > 0
    y0, z0, inv = (np.float32(x) for x in oi)
    nodes, _, _ = O.bvh_build(m["vertices"], m["faces"])
    u = nodes.view(np.uint32)
    leaf = (u[:, 7] & 0x80000000) != 0
    lb = nodes[leaf, :6]
    lref = (0x80000000 | ((u[leaf, 7] & 0x7FFFFFFF) << 28) | u[leaf, 6]).astype(np.uint32)
    eref = ent[:, 6].view(np.uint32)
    rng = np.random.default_rng(5)
    for z in np.concatenate([lb[:, 2], lb[:, 5], rng.uniform(-1, 1, 200)]).astype(np.float32):
        y = np.float32(0.0)
        c = (z - z0) * inv
        assert 0 <= c < nz
        b = int(c) * ny + int((y - y0) * inv)
        holds = (lb[:, 1] <= y) & (y <= lb[:, 4]) & (lb[:, 2] <= z) & (z <= lb[:, 5])
        assert set(lref[holds].tolist()) <= set(eref[off[b]:off[b + 1]].tolist())
