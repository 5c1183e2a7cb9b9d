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
