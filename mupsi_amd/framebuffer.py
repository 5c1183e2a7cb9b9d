"""Display and file output of a rendered image (Framebuffer, src/rendering/framebuffer.cpp).

An image here is the averaged radiance, (H, W, 3) float32 indexed [y, x] with
y the camera row (row 0 is the bottom of the picture, Camera::generateRay).
`tonemap` is Framebuffer::tonemapped (framebuffer.cpp:29-36);
`save_framebuffer` is Framebuffer::save (framebuffer.cpp:37-78): 8-bit
Reinhard-tonemapped PNG with rows flipped so the picture is upright, float
output for .hdr/.exr (written as Radiance RGBE, like the `mupsi` CLI) and
.pfm (rows bottom-up, camera row 0 first).
"""
from __future__ import annotations

import os
import struct
import zlib

import numpy as np

from .scene import write_pfm


def tonemap(img: np.ndarray) -> np.ndarray:
    """Reinhard v / (1 + v) per channel, in float32 as the reference."""
    v = np.asarray(img, dtype=np.float32)
    with np.errstate(invalid="ignore", divide="ignore"):
        return v / (np.float32(1.0) + v)


def to_8bit(img: np.ndarray) -> np.ndarray:
    """uint8(clamp(tone(v), 0, 1) * 255) with truncation, rows flipped to top-down."""
    t = np.clip(tonemap(img), np.float32(0.0), np.float32(1.0)) * np.float32(255.0)
    t = np.nan_to_num(t, nan=0.0)
    return np.ascontiguousarray(t.astype(np.uint8)[::-1])


def _chunk(tag: bytes, data: bytes) -> bytes:
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data))


def write_png(path: str | os.PathLike, rgb8: np.ndarray) -> None:
    """(H, W, 3) uint8, row 0 at the top, as an 8-bit RGB PNG."""
    rgb8 = np.ascontiguousarray(rgb8, dtype=np.uint8)
    h, w, _ = rgb8.shape
    raw = np.zeros((h, 1 + 3 * w), np.uint8)
    raw[:, 1:] = rgb8.reshape(h, 3 * w)
    with open(path, "wb") as fh:
        fh.write(b"\x89PNG\r\n\x1a\n")
        fh.write(_chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)))
        fh.write(_chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)))
        fh.write(_chunk(b"IEND", b""))


def read_png(path: str | os.PathLike) -> np.ndarray:
    """An 8-bit RGB PNG as `write_png` and the `mupsi` CLI write it (unfiltered
    scanlines) -> (H, W, 3) uint8, row 0 at the top."""
    with open(path, "rb") as fh:
        data = fh.read()
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError(f"not a PNG file: {path}")
    pos, idat, hdr = 8, [], None
    while pos < len(data):
        n, tag = struct.unpack(">I4s", data[pos:pos + 8])
        body = data[pos + 8:pos + 8 + n]
        if tag == b"IHDR":
            hdr = struct.unpack(">IIBBBBB", body)
        elif tag == b"IDAT":
            idat.append(body)
        pos += 12 + n
    w, h, depth, ctype, _, _, interlace = hdr
    if depth != 8 or ctype != 2 or interlace:
        raise ValueError("only 8-bit RGB non-interlaced PNG is read")
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), np.uint8).reshape(h, 1 + 3 * w)
    if raw[:, 0].any():
        raise ValueError("filtered PNG scanlines are not read")
    return np.ascontiguousarray(raw[:, 1:].reshape(h, w, 3))


def write_hdr(path: str | os.PathLike, img: np.ndarray) -> None:
    """Radiance RGBE (uncompressed scanlines), top picture row first."""
    v = np.asarray(img, dtype=np.float32)[::-1]
    h, w, _ = v.shape
    m = v.max(axis=2)
    mant, ex = np.frexp(m)
    ok = m > np.float32(1e-32)
    with np.errstate(divide="ignore", invalid="ignore"):
        s = np.where(ok, mant.astype(np.float32) * np.float32(256.0) / m, np.float32(0.0))
    e = np.zeros((h, w, 4), np.uint8)
    e[..., :3] = np.where(ok[..., None], (v * s[..., None]).astype(np.float32), 0).astype(np.uint8)
    e[..., 3] = np.where(ok, ex + 128, 0).astype(np.uint8)
    with open(path, "wb") as fh:
        fh.write(b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y %d +X %d\n" % (h, w))
        fh.write(e.tobytes())


def save_framebuffer(path: str | os.PathLike, img: np.ndarray) -> None:
    """Framebuffer::save: .pfm / .hdr / .exr keep the radiance, anything else is an 8-bit PNG."""
    p = str(path)
    img = np.asarray(img, dtype=np.float32)
    if p.endswith(".pfm"):
        write_pfm(p, img[::-1])        # write_pfm takes row 0 at the top
    elif p.endswith(".hdr") or p.endswith(".exr"):
        write_hdr(p, img)
    else:
        write_png(p, to_8bit(img))
