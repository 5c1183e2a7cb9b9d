"""Framebuffer display and file output (src/rendering/framebuffer.cpp:29-78) on the
host: Reinhard values, 8-bit truncation and the row flip, PNG / RGBE / PFM files
read back independently of the writers' own arithmetic."""
import numpy as np

from mupsi_amd.framebuffer import read_png, save_framebuffer, to_8bit, tonemap
from mupsi_amd.scene import read_pfm


def test_reinhard_known_values():
    v = np.array([0.0, 1.0, 3.0, -0.5, 1e30, np.inf], np.float32)
    t = tonemap(v)
    assert t[0] == 0.0 and t[1] == 0.5 and t[2] == 0.75 and t[3] == -1.0
    assert t[4] == 1.0 and np.isnan(t[5])    # inf / inf, as v / (1 + v) in float


def test_8bit_truncates_clamps_and_flips():
    img = np.zeros((2, 3, 3), np.float32)
    img[0, 0] = [1.0, 3.0, 0.0]         # camera row 0 -> bottom picture row
    img[1, 2] = [-0.5, 1e9, 0.9999]     # tone -1 clamps to 0; 1 - 1e-9 rounds to 1
    e = to_8bit(img)
    assert e.dtype == np.uint8 and e.shape == (2, 3, 3)
    assert list(e[1, 0]) == [127, 191, 0]           # 127.5 and 191.25 truncate
    assert list(e[0, 2]) == [0, 255, 127]           # 0.9999/1.9999*255 = 127.49...
    assert not e[0, :2].any() and not e[1, 1:].any()


def test_png_file_round_trip(tmp_path):
    rng = np.random.default_rng(1)
    img = rng.gamma(1.0, 1.0, (17, 23, 3)).astype(np.float32)
    save_framebuffer(tmp_path / "a.png", img)
    got = read_png(tmp_path / "a.png")
    assert got.shape == (17, 23, 3)
    assert np.array_equal(got, to_8bit(img))
    # the picture is upright: its bottom row is camera row 0
    assert np.array_equal(got[-1], to_8bit(img[:1])[0])


def _read_rgbe(path):
    data = open(path, "rb").read()
    head, body = data.split(b"\n\n", 1)
    assert head.startswith(b"#?RADIANCE")
    dims, body = body.split(b"\n", 1)
    _, h, _, w = dims.split()
    e = np.frombuffer(body, np.uint8).reshape(int(h), int(w), 4).astype(np.float64)
    scale = np.where(e[..., 3:] > 0, np.ldexp(1.0, (e[..., 3:] - 136).astype(int)), 0.0)
    return e[..., :3] * scale


def test_rgbe_file_is_upright_and_within_a_mantissa_step(tmp_path):
    rng = np.random.default_rng(2)
    img = (rng.gamma(0.5, 2.0, (9, 11, 3)) * rng.choice([1e-3, 1.0, 1e3], (9, 11, 1))).astype(np.float32)
    img[0, 0] = 0.0
    save_framebuffer(tmp_path / "a.hdr", img)
    dec = _read_rgbe(tmp_path / "a.hdr")[::-1]        # back to camera rows
    m = img.max(axis=2, keepdims=True).astype(np.float64)
    assert np.all(np.abs(dec - img) <= m / 128.0 + 1e-30)
    assert np.all(dec <= img * (1 + 1e-6) + 1e-30)   # truncation never rounds up
    assert not dec[0, 0].any()


def test_pfm_file_keeps_float_radiance(tmp_path):
    rng = np.random.default_rng(3)
    img = rng.normal(size=(5, 7, 3)).astype(np.float32)
    save_framebuffer(tmp_path / "a.pfm", img)
    # read_pfm returns the picture top-down; camera row 0 is its last row
    assert np.array_equal(read_pfm(tmp_path / "a.pfm")[::-1], img)
    raw = open(tmp_path / "a.pfm", "rb").read()
    assert np.array_equal(np.frombuffer(raw[-img.nbytes:], "<f4")[:3], img[0, 0])
