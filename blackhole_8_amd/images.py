"""Image files through the library's host-side codec (no GPU needed).

`load` / `decode_png` / `decode_jpeg` stand in for the reference's texture input,
cv::imread(path, IMREAD_COLOR) (config.h:35-37): the result is rows x cols x 3
bytes in B, G, R order with alpha dropped, the layout Material::SetTexture
takes (material.h:29-35).  `write` saves a frame (B, G, R bytes, as the
framebuffer holds them for such textures) as .ppm or .png.
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
