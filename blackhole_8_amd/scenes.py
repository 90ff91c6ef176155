"""Scene presets for the measured workloads and seeded synthetic inputs.

Every texture comes from the library's SplitMix64 generators, so the GPU and
the oracle see byte-identical inputs.  Texture slot convention used by the
presets: 0 background, 1 disc, 2.. spheres.

Workloads (lengths in units of r_s = 2M = 1):
  plain_hole    hole alone, camera at 15 r_s, 60 deg, 320x180, FP64, escape at
                100 r_s (cap 10000), 2048x1024 equirect noise+checker background
  disc_orbit    hole + thin equatorial disc 3..12 r_s (1024x1024 RGBA radial
                bands + noise), camera at 20 r_s and inclination 80 deg,
                1920x1080, FP32 with the FP64 photon-sphere band, 120-frame orbit
                at 3 deg/frame
  spheres       hole + 16 textured spheres (radii U[0.3,1.5], centres uniform in
                the 5..30 r_s shell), plain_hole's background and camera,
                1280x720 FP32
  uneven_batch  512 scenes of uneven size, mass, camera and disc in one batch
  fit_plain     the reference-rule render of the 512x512 inverse-fit scene
  shipped       the reference demo's own scene (blackhole_solution_test.cc:54-62,
                103-112,148): 960x540, fov pi/2, camera (-2000,0,400), M=10, disc
                100..1000 with p3.y = -0100 (octal -64), synthetic disc texture
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import abi

SEED_BASE = 0xB1AC0000
SLOT_BG, SLOT_DISC, SLOT_SPHERE0 = 0, 1, 2
SLOT_FLAT0 = 40  # rectangles of the flat-space scene


# --------------------------------------------------------------------------
# generators
# --------------------------------------------------------------------------
def gen_background(rows: int = 1024, cols: int = 2048, cells=(16, 8),
                   seed: int = SEED_BASE) -> np.ndarray:
    out = np.empty((rows, cols, 3), np.uint8)
    abi.load().bh_gen_background(abi.ptr(out), rows, cols, cells[0], cells[1], seed)
    return out


def gen_disc_texture(rows: int = 1024, cols: int = 1024, band_px: int = 32,
                     seed: int = SEED_BASE + 1) -> np.ndarray:
    out = np.empty((rows, cols, 4), np.uint8)
    abi.load().bh_gen_disc_texture(abi.ptr(out), rows, cols, band_px, seed)
    return out


def gen_noise(shape, seed: int) -> np.ndarray:
    out = np.empty(shape, np.uint8)
    abi.load().bh_gen_noise(abi.ptr(out), out.size, seed)
    return out


class Rng:
    """SplitMix64 stream shared with the C generators."""

    def __init__(self, seed: int):
        self.state = C.c_uint64(seed)

    def next(self) -> int:
        return abi.load().bh_splitmix64(C.byref(self.state))

    def uniform(self, lo: float, hi: float) -> float:
        return abi.load().bh_uniform(C.byref(self.state), lo, hi)


# --------------------------------------------------------------------------
# cameras and objects
# --------------------------------------------------------------------------
def camera(width: int, height: int, fov: float, position=(0.0, 0.0, 0.0),
           axes=((1, 0, 0), (0, -1, 0), (0, 0, -1))) -> abi.Camera:
    """Camera<T>(w, h, fov) with the default frame (camera.h:30-35) moved to
    position (Object::MoveTo)."""
    c = abi.Camera()
    c.position[:] = position
    c.axis_x[:], c.axis_y[:], c.axis_z[:] = axes
    c.width, c.height, c.fov = width, height, fov
    return c


def look_at(width: int, height: int, fov: float, position, target=(0.0, 0.0, 0.0),
            up=(0.0, 0.0, 1.0)) -> abi.Camera:
    c = camera(width, height, fov, position)
    t = (C.c_double * 3)(*target)
    u = (C.c_double * 3)(*up)
    abi.load().bh_camera_look_at(C.byref(c), t, u)
    return c


def orbit_camera(r: float, incl_deg: float, azim_deg: float, width: int, height: int,
                 fov: float, target=(0.0, 0.0, 0.0)) -> abi.Camera:
    i, a = math.radians(incl_deg), math.radians(azim_deg)
    pos = (target[0] + r * math.sin(i) * math.cos(a), target[1] + r * math.sin(i) * math.sin(a),
           target[2] + r * math.cos(i))
    return look_at(width, height, fov, pos, target)


def annulus(p1, p2, p3, p4, r_outer: float, r_inner: float, texture: int) -> abi.Annulus:
    a = abi.Annulus()
    arr = [np.ascontiguousarray(p, dtype=np.float64) for p in (p1, p2, p3, p4)]
    abi.load().bh_annulus_init(C.byref(a), *[abi.dptr(v) for v in arr], r_outer, r_inner, texture)
    return a


def options(precision=abi.PREC_FP64, nstep=20, escape=False, r_escape=0.0, step_cap=10000,
            filter=abi.FILTER_NEAREST, root=abi.ROOT_BISECT20, band=1e-3,
            edge_width=0.0, edge_guard=0.0) -> abi.Options:
    o = abi.Options()
    o.precision, o.nstep, o.escape, o.step_cap = precision, nstep, int(escape), step_cap
    o.r_escape, o.filter, o.root, o.band, o.edge_width = r_escape, filter, root, band, edge_width
    o.edge_guard = edge_guard
    return o


@dataclass
class Preset:
    name: str
    scene: abi.Scene
    options: abi.Options
    textures: dict = field(default_factory=dict)
    cameras: list = field(default_factory=list)

    @property
    def width(self):
        return self.scene.camera.width

    @property
    def height(self):
        return self.scene.camera.height

    def with_camera(self, cam: abi.Camera) -> "Preset":
        s = abi.Scene.from_buffer_copy(self.scene)
        s.camera = cam
        return Preset(self.name, s, self.options, self.textures, self.cameras)

    def resized(self, width: int, height: int) -> "Preset":
        """Same scene at another resolution (fov kept)."""
        cam = abi.Camera.from_buffer_copy(self.scene.camera)
        cam.width, cam.height = width, height
        cams = []
        for c0 in self.cameras:
            c1 = abi.Camera.from_buffer_copy(c0)
            c1.width, c1.height = width, height
            cams.append(c1)
        p = self.with_camera(cam)
        p.cameras = cams
        return p


def _scene(cam: abi.Camera, mass: float = 0.5, hole_pos=(0.0, 0.0, 0.0)) -> abi.Scene:
    s = abi.Scene()
    s.camera = cam
    s.hole.position[:] = hole_pos
    s.hole.mass = mass
    s.has_disc = 0
    s.n_spheres = 0
    s.background = -1
    return s


# --------------------------------------------------------------------------
# the workloads
# --------------------------------------------------------------------------
def plain_hole(width: int = 320, height: int = 180) -> Preset:
    cam = camera(width, height, math.pi / 3, position=(-15.0, 0.0, 0.0))
    s = _scene(cam)
    s.background = SLOT_BG
    o = options(precision=abi.PREC_FP64, escape=True, r_escape=100.0, step_cap=10000)
    return Preset("plain_hole", s, o, {SLOT_BG: gen_background()})


def disc_orbit_camera(frame: int, width: int = 1920, height: int = 1080) -> abi.Camera:
    return orbit_camera(20.0, 80.0, 3.0 * frame, width, height, math.pi / 3)


def disc_orbit(width: int = 1920, height: int = 1080, frames: int = 120,
            precision=abi.PREC_FP32) -> Preset:
    s = _scene(disc_orbit_camera(0, width, height))
    r_out, r_in = 12.0, 3.0
    s.has_disc = 1
    s.disc = annulus((r_out, r_out, 0.0), (-r_out, r_out, 0.0), (-r_out, -r_out, 0.0),
                     (r_out, -r_out, 0.0), r_out, r_in, SLOT_DISC)
    o = options(precision=precision, escape=False, band=1e-3)
    cams = [disc_orbit_camera(k, width, height) for k in range(frames)]
    return Preset("disc_orbit", s, o, {SLOT_DISC: gen_disc_texture()}, cams)


def spheres(width: int = 1280, height: int = 720, precision=abi.PREC_FP32,
            n_spheres: int = 16) -> Preset:
    cam = camera(width, height, math.pi / 3, position=(-15.0, 0.0, 0.0))
    s = _scene(cam)
    s.background = SLOT_BG
    rng = Rng(SEED_BASE + 2)
    tex = {SLOT_BG: gen_background()}
    placed = 0
    while placed < n_spheres:
        rad = rng.uniform(0.3, 1.5)
        # uniform in the volume of the 5..30 shell
        rr = (rng.uniform(5.0 ** 3, 30.0 ** 3)) ** (1.0 / 3.0)
        z = rng.uniform(-1.0, 1.0)
        ph = rng.uniform(0.0, 2.0 * math.pi)
        q = math.sqrt(max(0.0, 1.0 - z * z))
        c = (rr * q * math.cos(ph), rr * q * math.sin(ph), rr * z)
        # keep the camera outside every sphere (with a margin)
        dc = math.dist(c, tuple(cam.position))
        if dc < rad + 0.5:
            continue
        sp = s.spheres[placed]
        sp.center[:] = c
        sp.radius = rad
        sp.texture = SLOT_SPHERE0 + placed
        tex[SLOT_SPHERE0 + placed] = gen_noise((256, 256, 3), SEED_BASE + 0x100 + placed)
        placed += 1
    s.n_spheres = n_spheres
    o = options(precision=precision, escape=True, r_escape=100.0, step_cap=10000)
    return Preset("spheres", s, o, tex)


def uneven_batch(n: int = 512, seed: int = SEED_BASE + 0x300):
    """Batched uneven renders: n independent hole (+ disc with probability
    1/2) scenes, square side log-uniform in [32, 1024], r_s ~ U[0.5, 2],
    camera at U[6, 50] r_s, inclination U[0, 90] deg, azimuth U[0, 360) deg,
    fov 60 deg, disc 3..12 r_s with the 1080p workload's texture, FP32.
    Returns (list of (scene, options), textures)."""
    rng = Rng(seed)
    items = []
    for _ in range(n):
        side = int(round(math.exp(rng.uniform(math.log(32.0), math.log(1024.0)))))
        r_s = rng.uniform(0.5, 2.0)
        dist = rng.uniform(6.0, 50.0) * r_s
        incl = max(1e-3, rng.uniform(0.0, 90.0))
        azim = rng.uniform(0.0, 360.0)
        disc = rng.uniform(0.0, 1.0) < 0.5
        cam = orbit_camera(dist, incl, azim, side, side, math.pi / 3)
        s = _scene(cam, mass=r_s / 2.0)
        if disc:
            r_out, r_in = 12.0 * r_s, 3.0 * r_s
            s.has_disc = 1
            s.disc = annulus((r_out, r_out, 0.0), (-r_out, r_out, 0.0), (-r_out, -r_out, 0.0),
                             (r_out, -r_out, 0.0), r_out, r_in, SLOT_DISC)
        items.append((s, options(precision=abi.PREC_FP32, escape=False, band=1e-3)))
    return items, {SLOT_DISC: gen_disc_texture()}


def fit_plain(theta, width: int = 512, height: int = 512, fov: float = math.radians(70.0),
                  r_outer: float = 12.0, azim: float = 0.0,
                  precision=abi.PREC_FP32) -> Preset:
    """The inverse-fit scene at theta = (r_s, incl, dist, r_inner) for the
    plain march (reference rule: 20-step bisection, nearest texel, hard disc
    edge): the same hole, disc and look-at camera as the differentiable
    render of those parameters."""
    r_s, incl, dist, r_in = (float(v) for v in theta)
    cam = orbit_camera(dist, math.degrees(incl), math.degrees(azim), width, height, fov)
    s = _scene(cam, mass=r_s / 2.0)
    s.has_disc = 1
    s.disc = annulus((r_outer, r_outer, 0.0), (-r_outer, r_outer, 0.0), (-r_outer, -r_outer, 0.0),
                     (r_outer, -r_outer, 0.0), r_outer, r_in, SLOT_DISC)
    o = options(precision=precision, escape=False, band=1e-3)
    return Preset("fit_plain", s, o, {SLOT_DISC: gen_disc_texture()})


def flat_test(width: int = 800, height: int = 450):
    """The flat-space scene of ray_tracer_test.cc:50-95: camera (fov pi/2)
    moved to (-400, 120, 100), turned by RotateY(-pi/24) then RotateZ(pi/12);
    three textured rectangles (the reference's jpeg photos replaced by seeded
    noise images); the ground plane z = 0 with ChessPattern2D(10); 10 chords
    per pixel.  Returns (FlatScene, textures)."""
    lib = abi.load()
    cam = camera(width, height, math.pi / 2, position=(-400.0, 120.0, 100.0))
    axes = (C.c_double * 9)(*cam.axis_x, *cam.axis_y, *cam.axis_z)
    lib.bh_object_rotate(axes, None, 0, 1, -math.pi / 24)
    lib.bh_object_rotate(axes, None, 0, 2, math.pi / 12)
    cam.axis_x[:], cam.axis_y[:], cam.axis_z[:] = axes[0:3], axes[3:6], axes[6:9]
    s = abi.FlatScene()
    s.camera = cam
    rects = [((100, 120, 100), (100, 0, 100), (100, 0, 0), (100, 120, 0)),
             ((100, 0, 150), (100, -100, 150), (100, -100, 0), (100, 0, 0)),
             ((-100, 50, 100), (-100, -50, 100), (-100, -50, 0), (-100, 50, 0))]
    sizes = [(360, 480), (400, 300), (320, 320)]
    tex = {}
    for i, (pts, (h, w)) in enumerate(zip(rects, sizes)):
        r = s.rects[i]
        r.vertex[0][:] = (0.0, 0.0, 0.0)  # Object's default pivot
        for k in range(4):
            r.vertex[k + 1][:] = pts[k]
        r.texture = SLOT_FLAT0 + i
        tex[SLOT_FLAT0 + i] = gen_noise((h, w, 3), SEED_BASE + 0x200 + i)
    s.n_rects = 3
    pl = s.planes[0]
    pl.position[:] = (0.0, 0.0, 0.0)
    pl.axis_x[:], pl.axis_y[:], pl.axis_z[:] = (1, 0, 0), (0, 1, 0), (0, 0, 1)
    pl.pattern_size = 10.0
    s.n_planes = 1
    s.steps = 10
    return s, tex


def shipped(width: int = 960, height: int = 540, precision=abi.PREC_FP64) -> Preset:
    cam = camera(width, height, math.pi / 2, position=(-2000.0, 0.0, 400.0))
    s = _scene(cam, mass=10.0)
    s.has_disc = 1
    s.disc = annulus((1000.0, 1000.0, 0.0), (-1000.0, 1000.0, 0.0), (-1000.0, -64.0, 0.0),
                     (1000.0, -1000.0, 0.0), 1000.0, 100.0, SLOT_DISC)
    o = options(precision=precision, escape=False, band=1e-3)
    return Preset("shipped", s, o, {SLOT_DISC: gen_disc_texture(598, 578, 16, SEED_BASE + 3)})
