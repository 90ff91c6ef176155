"""Host-side API over the C ABI: a Renderer owns one bh_ctx (one HIP device).

Mirrors the reference's render entry point (the inline march of
blackhole_solution_test.cc:161-298) as `Renderer.render(scene, options)`.
There is no CPU fallback: without the built HIP library or a GPU every call
raises.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import abi


class BHError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"bh error {code}: {msg}")
        self.code = code


OUTPUT_NAMES = ("rgb", "hit", "phi", "hitpt", "steps")


def alloc_outputs(width: int, height: int, want=("rgb", "hit")):
    shapes = {"rgb": ((height, width, 3), np.uint8), "hit": ((height, width), np.uint8),
              "phi": ((height, width), np.float64), "hitpt": ((height, width, 3), np.float64),
              "steps": ((height, width), np.int32)}
    arrs = {k: np.zeros(*shapes[k]) for k in want}
    o = abi.Outputs(abi.ptr(arrs.get("rgb")), abi.ptr(arrs.get("hit")), abi.ptr(arrs.get("phi")),
                    abi.ptr(arrs.get("hitpt")), abi.ptr(arrs.get("steps")))
    return arrs, o


class Renderer:
    def __init__(self, device: int = 0):
        self.lib = abi.load()
        self.ctx = C.c_void_p()
        rc = self.lib.bh_ctx_create(device, C.byref(self.ctx))
        if rc != abi.BH_OK:
            raise BHError(rc, self.lib.bh_last_error(None).decode())
        self._tex_keep = {}

    # -- lifetime --------------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.bh_ctx_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc: int):
        if rc != abi.BH_OK:
            raise BHError(rc, self.lib.bh_last_error(self.ctx).decode())

    def device_info(self):
        name = C.create_string_buffer(256)
        cu, clk = C.c_int(), C.c_int()
        self._check(self.lib.bh_device_info(self.ctx, name, 256, C.byref(cu), C.byref(clk)))
        return {"name": name.value.decode(), "cu": cu.value, "clock_khz": clk.value}

    # -- textures (Material::SetTexture) ----------------------------------
    def set_texture(self, slot: int, img: np.ndarray):
        img = np.ascontiguousarray(img, np.uint8)
        self._check(self.lib.bh_texture_set(self.ctx, slot, C.byref(abi.image(img))))

    def set_textures(self, textures: dict):
        for slot, img in textures.items():
            self.set_texture(slot, img)

    # -- rendering ---------------------------------------------------------
    def render(self, scene: abi.Scene, opts: abi.Options, want=("rgb", "hit")) -> dict:
        w, h = scene.camera.width, scene.camera.height
        arrs, o = alloc_outputs(w, h, want)
        self._check(self.lib.bh_render(self.ctx, C.byref(scene), C.byref(opts), C.byref(o)))
        return arrs

    def render_preset(self, preset, want=("rgb", "hit")) -> dict:
        self.set_textures(preset.textures)
        return self.render(preset.scene, preset.options, want)

    def render_flat(self, scene: abi.FlatScene, want=("rgb", "hit")) -> dict:
        """Flat-space march (RayTracer::Prograde) of a FlatScene."""
        w, h = scene.camera.width, scene.camera.height
        arrs, o = alloc_outputs(w, h, want)
        self._check(self.lib.bh_render_flat(self.ctx, C.byref(scene), C.byref(o)))
        return arrs

    def time_flat(self, scene: abi.FlatScene, iters: int = 20) -> float:
        """Average ms of the flat march kernel (bh_time_flat)."""
        ms = C.c_double()
        self._check(self.lib.bh_time_flat(self.ctx, C.byref(scene), iters, C.byref(ms)))
        return ms.value

    def check_scene(self, scene: abi.Scene, opts: abi.Options) -> int:
        return self.lib.bh_scene_check(self.ctx, C.byref(scene), C.byref(opts))

    def time_render(self, scene: abi.Scene, opts: abi.Options, iters: int = 10):
        m, t, n = C.c_double(), C.c_double(), C.c_int64()
        self._check(self.lib.bh_time_render(self.ctx, C.byref(scene), C.byref(opts), iters,
                                            C.byref(m), C.byref(t), C.byref(n)))
        return m.value, t.value, n.value

    # -- batches ------------------------------------------------------------
    def render_batch(self, items, want=("rgb", "hit"), tuning: abi.BatchTuning | None = None,
                     timing: bool = False):
        """Render many (scene, options) pairs in one persistent launch
        (bh_render_batch_device).  Returns a list of dicts of numpy planes;
        with timing, also (scene_ms array, stats)."""
        import torch

        n = len(items)
        scenes = (abi.Scene * n)(*[s for s, _ in items])
        opts = (abi.Options * n)(*[o for _, o in items])
        outs = (abi.Outputs * n)()
        px = [s.camera.width * s.camera.height for s, _ in items]
        off = np.concatenate([[0], np.cumsum(px)]).astype(np.int64)
        total = int(off[-1])
        dev = torch.device("cuda", 0)
        bufs = {}
        if "rgb" in want:
            bufs["rgb"] = torch.empty(total * 3, dtype=torch.uint8, device=dev)
        if "hit" in want:
            bufs["hit"] = torch.empty(total, dtype=torch.uint8, device=dev)
        if "steps" in want:
            bufs["steps"] = torch.empty(total, dtype=torch.int32, device=dev)
        torch.cuda.synchronize(dev)
        for i in range(n):
            o = int(off[i])
            outs[i].rgb = bufs["rgb"].data_ptr() + 3 * o if "rgb" in bufs else None
            outs[i].hit = bufs["hit"].data_ptr() + o if "hit" in bufs else None
            outs[i].steps = bufs["steps"].data_ptr() + 4 * o if "steps" in bufs else None
        ms = (C.c_double * n)()
        st = abi.BatchStats()
        self._check(self.lib.bh_render_batch_device(
            self.ctx, n, scenes, opts, outs, C.byref(tuning) if tuning is not None else None,
            ms, C.byref(st)))
        host = {k: v.cpu().numpy() for k, v in bufs.items()}
        res = []
        for i, (s, _) in enumerate(items):
            w, h = s.camera.width, s.camera.height
            d = {}
            a, b = int(off[i]), int(off[i + 1])
            if "rgb" in host:
                d["rgb"] = host["rgb"][3 * a:3 * b].reshape(h, w, 3)
            if "hit" in host:
                d["hit"] = host["hit"][a:b].reshape(h, w)
            if "steps" in host:
                d["steps"] = host["steps"][a:b].reshape(h, w)
            res.append(d)
        if timing:
            return res, np.array(ms[:], dtype=np.float64), st
        return res

    def path(self, scene: abi.Scene, opts: abi.Options, cameras) -> "Path":
        return Path(self, scene, opts, cameras)


class Path:
    """A camera path captured as one HIP graph (bh_path_*)."""

    def __init__(self, r: Renderer, scene: abi.Scene, opts: abi.Options, cameras):
        self.r = r
        self.n = len(cameras)
        self.w, self.h = cameras[0].width, cameras[0].height
        cams = (abi.Camera * self.n)(*cameras)
        self.handle = C.c_void_p()
        r._check(r.lib.bh_path_create(r.ctx, C.byref(scene), C.byref(opts), cams, self.n,
                                      C.byref(self.handle)))

    def launch(self) -> float:
        ms = C.c_float()
        self.r._check(self.r.lib.bh_path_launch(self.handle, C.byref(ms)))
        return ms.value

    def time(self, iters: int = 3):
        """(march_ms, band_ms): the fused launches of a path whose frames
        all take the FP32 march (bh_path_time)."""
        m, b = C.c_double(), C.c_double()
        self.r._check(self.r.lib.bh_path_time(self.handle, iters, C.byref(m), C.byref(b)))
        return m.value, b.value

    def last_ms(self):
        """(march_ms, band_ms) of the last graph launch, from timing events
        the graph records around its two kernels (bh_path_last_ms)."""
        m, b = C.c_double(), C.c_double()
        self.r._check(self.r.lib.bh_path_last_ms(self.handle, C.byref(m), C.byref(b)))
        return m.value, b.value

    def read(self, frame: int):
        rgb = np.zeros((self.h, self.w, 3), np.uint8)
        hit = np.zeros((self.h, self.w), np.uint8)
        self.r._check(self.r.lib.bh_path_read(self.handle, frame, abi.ptr(rgb), abi.ptr(hit)))
        return rgb, hit

    def close(self):
        if self.handle:
            self.r.lib.bh_path_destroy(self.handle)
            self.handle = C.c_void_p()


class Ring:
    """Frame ring (bh_ring_*): frames render into device slots and are copied
    to pinned host memory while the host consumes earlier slots."""

    def __init__(self, r: Renderer, width: int, height: int, depth: int = 2):
        self.r = r
        self.w, self.h = width, height
        self.handle = C.c_void_p()
        r._check(r.lib.bh_ring_create(r.ctx, width, height, depth, C.byref(self.handle)))

    def render(self, scene) -> int:
        slot = C.c_int()
        if isinstance(scene, abi.FlatScene):
            rc = self.r.lib.bh_ring_render_flat(self.handle, C.byref(scene), C.byref(slot))
        else:
            s, o = scene
            rc = self.r.lib.bh_ring_render(self.handle, C.byref(s), C.byref(o), C.byref(slot))
        self.r._check(rc)
        return slot.value

    def wait(self, slot: int, copy: bool = True):
        """(rgb, hit, gpu_ms) of slot's frame; copies out of the pinned slot
        unless copy is False (then the arrays alias it until release)."""
        rgb, hit, ms = C.c_void_p(), C.c_void_p(), C.c_float()
        rc = self.r.lib.bh_ring_wait(self.handle, slot, C.byref(rgb), C.byref(hit), C.byref(ms))
        if rc != abi.BH_OK:
            raise BHError(rc, self.r.lib.bh_last_error(None).decode())
        n = self.w * self.h
        a = np.ctypeslib.as_array(C.cast(rgb, C.POINTER(C.c_uint8)), (n * 3,))
        b = np.ctypeslib.as_array(C.cast(hit, C.POINTER(C.c_uint8)), (n,))
        a = a.reshape(self.h, self.w, 3)
        b = b.reshape(self.h, self.w)
        if copy:
            a, b = a.copy(), b.copy()
        return a, b, ms.value

    def release(self, slot: int):
        rc = self.r.lib.bh_ring_release(self.handle, slot)
        if rc != abi.BH_OK:
            raise BHError(rc, self.r.lib.bh_last_error(None).decode())

    def close(self):
        if self.handle:
            self.r.lib.bh_ring_destroy(self.handle)
            self.handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
