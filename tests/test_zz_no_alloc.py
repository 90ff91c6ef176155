"""bh_render_device_noalloc: the allocation-free variant of bh_render_device
(scene record and FP64 queue in the library's file-scope device array).  Its
planes must equal bh_render_device's bit for bit (the same kernels on the
same inputs), on consecutive calls of different sizes, and it must leave the
device's free memory unchanged."""
import ctypes as C

import numpy as np
import pytest

from blackhole_8_amd import abi, scenes

pytestmark = pytest.mark.gpu

PLANES = (("rgb", 3, "uint8"), ("hit", 1, "uint8"), ("phi_end", 1, "float64"),
          ("hit_point", 3, "float64"), ("steps", 1, "int32"))


def _planes(torch, w, h):
    dt = {"uint8": torch.uint8, "float64": torch.float64, "int32": torch.int32}
    t = {n: torch.full((h * w * c,), 7, dtype=dt[d], device="cuda") for n, c, d in PLANES}
    o = abi.Outputs()
    for n, _, _ in PLANES:
        setattr(o, n, t[n].data_ptr())
    return t, o


def _render(renderer, fn, p, torch):
    t, o = _planes(torch, p.width, p.height)
    torch.cuda.synchronize()
    rc = getattr(renderer.lib, fn)(renderer.ctx, C.byref(p.scene), C.byref(p.options),
                                   C.byref(o))
    assert rc == abi.BH_OK, renderer.lib.bh_last_error(renderer.ctx).decode()
    assert renderer.lib.bh_synchronize(renderer.ctx) == abi.BH_OK
    return {n: v.cpu().numpy() for n, v in t.items()}


def _same(a, b):
    for n, _, _ in PLANES:
        assert np.array_equal(a[n], b[n], equal_nan=True), n


def test_noalloc_equals_render_device(renderer):
    import torch

    big = scenes.disc_orbit(1920, 1080, frames=1)
    small = scenes.disc_orbit(640, 360, frames=1).with_camera(
        scenes.disc_orbit_camera(40, 640, 360))
    renderer.set_textures(big.textures)
    for p in (big, small):  # two calls in a row, the second smaller
        print(f"[case] noalloc {p.width}x{p.height}", flush=True)
        got = _render(renderer, "bh_render_device_noalloc", p, torch)
        want = _render(renderer, "bh_render_device", p, torch)
        _same(got, want)
        assert (got["hit"] == abi.HIT_DISC).sum() > 10000
    # the FP64 reference march (scene record only, no queue)
    p = scenes.plain_hole(320, 180)
    renderer.set_textures(p.textures)
    print("[case] noalloc fp64 plain hole 320x180", flush=True)
    _same(_render(renderer, "bh_render_device_noalloc", p, torch),
          _render(renderer, "bh_render_device", p, torch))


def test_noalloc_allocates_nothing():
    # a context of its own, whose queue buffers were never allocated: the
    # entry point must not allocate them (or anything else) on the device
    import torch
    from blackhole_8_amd import Renderer

    p = scenes.disc_orbit(1920, 1080, frames=1)
    with Renderer(0) as r:
        r.set_textures(p.textures)
        t, o = _planes(torch, p.width, p.height)
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        print("[case] noalloc free memory", flush=True)
        for _ in range(2):
            assert r.lib.bh_render_device_noalloc(r.ctx, C.byref(p.scene), C.byref(p.options),
                                                  C.byref(o)) == abi.BH_OK
        assert r.lib.bh_synchronize(r.ctx) == abi.BH_OK
        free1, _ = torch.cuda.mem_get_info()
        # bh_render_device would allocate the 17.8 MB queue of a 1080p frame
        assert abs(free0 - free1) < (1 << 20), (free0, free1)
        assert (t["hit"].cpu().numpy() == abi.HIT_DISC).sum() > 100000


def test_noalloc_rejects_frames_beyond_its_queue(renderer):
    import torch

    p = scenes.disc_orbit(3840, 2160, frames=1)
    renderer.set_textures(p.textures)
    o = abi.Outputs()  # never written: the call fails before any launch
    print("[case] noalloc 3840x2160 rejected", flush=True)
    rc = renderer.lib.bh_render_device_noalloc(renderer.ctx, C.byref(p.scene),
                                               C.byref(p.options), C.byref(o))
    assert rc == abi.BH_ERR_INVALID_ARG
    assert "too large" in renderer.lib.bh_last_error(renderer.ctx).decode()
