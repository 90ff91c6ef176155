"""Image files through the library's host-side codec (no GPU needed).

`load` / `decode_png` / `decode_jpeg` stand in for the reference's texture input,
cv::imread(path, IMREAD_COLOR) (config.h:35-37): the result is rows x cols x 3
bytes in B, G, R order with alpha dropped, the layout Material::SetTexture
takes (material.h:29-35).  `write` saves a frame (B, G, R bytes, as the
framebuffer holds them for such textures) as .ppm, .png or .jpg; `Video`
writes a Motion-JPEG AVI in place of the reference's cv::VideoWriter
(blackhole_solution_test.cc:69-72,330).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import abi


class ImageError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"image error {code}: {msg}")
        self.code = code


def _err(rc: int):
    raise ImageError(rc, abi.load().bh_io_last_error().decode())


def decode_png(data: bytes) -> np.ndarray:
    lib = abi.load()
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    rows, cols = C.c_int(), C.c_int()
    rc = lib.bh_png_decode(C.cast(buf, C.c_void_p), len(data), None, 0, C.byref(rows),
                           C.byref(cols))
    if rc:
        _err(rc)
    out = np.empty((rows.value, cols.value, 3), np.uint8)
    rc = lib.bh_png_decode(C.cast(buf, C.c_void_p), len(data), abi.ptr(out), out.size,
                           C.byref(rows), C.byref(cols))
    if rc:
        _err(rc)
    return out


def decode_jpeg(data: bytes) -> np.ndarray:
    lib = abi.load()
    buf = (C.c_uint8 * len(data)).from_buffer_copy(data)
    rows, cols = C.c_int(), C.c_int()
    rc = lib.bh_jpeg_decode(C.cast(buf, C.c_void_p), len(data), None, 0, C.byref(rows),
                            C.byref(cols))
    if rc:
        raise ImageError(rc, lib.bh_jpeg_last_error().decode())
    out = np.empty((rows.value, cols.value, 3), np.uint8)
    rc = lib.bh_jpeg_decode(C.cast(buf, C.c_void_p), len(data), abi.ptr(out), out.size,
                            C.byref(rows), C.byref(cols))
    if rc:
        raise ImageError(rc, lib.bh_jpeg_last_error().decode())
    return out


def load(path: os.PathLike | str) -> np.ndarray:
    lib = abi.load()
    p = os.fsencode(path)
    rows, cols = C.c_int(), C.c_int()
    rc = lib.bh_image_load(p, None, 0, C.byref(rows), C.byref(cols))
    if rc:
        _err(rc)
    out = np.empty((rows.value, cols.value, 3), np.uint8)
    rc = lib.bh_image_load(p, abi.ptr(out), out.size, C.byref(rows), C.byref(cols))
    if rc:
        _err(rc)
    return out


def write(path: os.PathLike | str, bgr: np.ndarray) -> None:
    a = np.ascontiguousarray(bgr, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("expected rows x cols x 3 bytes")
    rc = abi.load().bh_image_write(os.fsencode(path), abi.ptr(a), a.shape[0], a.shape[1])
    if rc:
        _err(rc)


def encode_jpeg(bgr: np.ndarray, quality: int = 95) -> bytes:
    """Baseline JFIF JPEG (4:2:0, optimal Huffman tables) of B, G, R bytes."""
    a = np.ascontiguousarray(bgr, dtype=np.uint8)
    if a.ndim != 3 or a.shape[2] != 3:
        raise ValueError("expected rows x cols x 3 bytes")
    lib = abi.load()
    out, n = C.POINTER(C.c_uint8)(), C.c_size_t()
    rc = lib.bh_jpeg_encode(abi.ptr(a), a.shape[0], a.shape[1], quality, C.byref(out),
                            C.byref(n))
    if rc:
        raise ImageError(rc, lib.bh_jpeg_last_error().decode())
    try:
        return C.string_at(out, n.value)
    finally:
        lib.bh_free(C.cast(out, C.c_void_p))


class Video:
    """Motion-JPEG AVI writer: `Video(path, rows, cols, fps=29)`, then
    `write(frame)` per B, G, R frame and `close()` (or use it as a context
    manager) to write the index."""

    def __init__(self, path: os.PathLike | str, rows: int, cols: int, fps: float = 29.0,
                 quality: int = 95):
        self._lib = abi.load()
        self.shape = (rows, cols, 3)
        h = C.c_void_p()
        rc = self._lib.bh_video_open(os.fsencode(path), rows, cols, fps, quality, C.byref(h))
        if rc:
            _err(rc)
        self._h = h

    def write(self, bgr: np.ndarray) -> None:
        if self._h is None:
            raise ValueError("video is closed")
        a = np.ascontiguousarray(bgr, dtype=np.uint8)
        if a.shape != self.shape:
            raise ValueError(f"expected a frame of shape {self.shape}, got {a.shape}")
        rc = self._lib.bh_video_write(self._h, abi.ptr(a))
        if rc:
            _err(rc)

    def close(self) -> None:
        if self._h is not None:
            h, self._h = self._h, None
            rc = self._lib.bh_video_close(h)
            if rc:
                _err(rc)

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        if getattr(self, "_h", None) is not None:
            self._lib.bh_video_close(self._h)
            self._h = None
