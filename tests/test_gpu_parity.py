"""GPU parity: the HIP march against the FP64 oracle on identical seeded inputs.

Bars (from the workload definitions):
  FP64 march  hit class and RGB identical except a stated fraction <= 1e-4 of
              pixels; phi_end within 1e-9 rad; step counts identical.
  FP32 march  hit class identical except <= 1e-4 of pixels; RGB within 1/255
              except <= 1e-4 of pixels (no exemption: rays whose FP32 texel
              coordinate lies near a texel edge are re-marched in FP64);
              phi_end within 1e-5 rad on pixels with the same class and steps;
              hit points within 1e-4 (world units, scenes of r_s = 1) on
              pixels with the same class.
The FP64 march and the oracle both follow the reference's operation order, but
the device's sin/cos/acos may differ from the host libm's in the last bit, which
decides the texel of the few hit points that lie exactly on a texel edge (the
texture's centre lines, under the camera's mirror symmetry): a few pixels per
million.
"""
import ctypes as C
import math
import os

import numpy as np
import pytest

from blackhole_8_amd import abi, scenes
from blackhole_8_amd.render import BHError
import oracle
from tests.parity import compare, disc_texel_coords, oracle_disc_texel, straddles, summary

pytestmark = pytest.mark.gpu

ALL = ("rgb", "hit", "phi", "hitpt", "steps")
THREADS = min(16, os.cpu_count() or 1)


def _log(name, st):
    print(f"[parity] {name}: {summary(st)}", flush=True)


def test_fp64_plain_hole_matches_oracle(renderer):
    p = scenes.plain_hole()
    print("[case] fp64 plain_hole 320x180", flush=True)
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log("fp64 plain_hole", st)
    assert st["class_mismatch_frac"] <= 1e-4
    assert st["rgb_mismatch_frac"] <= 1e-4
    assert st["steps_mismatch"] == 0
    assert st["phi_max_err"] <= 1e-9
    # both renders actually exercise the two classes and the background
    assert (r["hit"] == abi.HIT_HORIZON).sum() > 1000
    assert (r["hit"] == abi.HIT_NONE).sum() > 1000


def test_fp64_shipped_scene_matches_oracle(renderer):
    p = scenes.shipped(240, 135)
    print("[case] fp64 shipped 240x135", flush=True)
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log("fp64 shipped", st)
    assert st["class_mismatch_frac"] <= 1e-4
    assert st["rgb_mismatch_frac"] <= 1e-4
    assert st["phi_max_err"] <= 1e-9
    assert (r["hit"] == abi.HIT_DISC).sum() > 1000


def test_fp64_disc_orbit_matches_oracle(renderer):
    p = scenes.disc_orbit(480, 270, frames=1, precision=abi.PREC_FP64)
    print("[case] fp64 disc_orbit 480x270", flush=True)
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log("fp64 disc_orbit", st)
    assert st["class_mismatch_frac"] <= 1e-4
    assert st["rgb_mismatch_frac"] <= 1e-4
    assert st["phi_max_err"] <= 1e-9


def _check_fp32(renderer, p, name):
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    return _check_fp32_outputs(p, g, r, name)


RGB_FRAC = 1e-4      # RGB mismatches (same class), fraction of pixels
HITPT_ERR = 1e-4     # hit point error on pixels of the same class


def _allowed(n_pixels, frac=RGB_FRAC):
    """Mismatch count allowed by a fraction of a frame: at least one pixel,
    since a texel tie on a centre line can fall either way."""
    return max(1, int(frac * n_pixels))


def _check_fp32_outputs(p, g, r, name):
    st = compare(g, r)
    _log(name, st)
    bad = st["rgb_bad_mask"]
    if bad.any() and p.scene.has_disc:
        # diagnostics only: how many are straddles of nearly equal points
        tex = p.textures[scenes.SLOT_DISC]
        vert = np.array(p.scene.disc.vertex)
        sel = bad & (g["hit"] == abi.HIT_DISC)
        ga, gb = disc_texel_coords(g["hitpt"][sel], vert, tex.shape[0], tex.shape[1])
        ra, rb = oracle_disc_texel(r["hitpt"][sel], vert, tex.shape[0], tex.shape[1])
        near = straddles(ga, ra) | straddles(gb, rb)
        print(f"[parity] {name}: rgb mismatches {int(bad.sum())}, disc texel straddles "
              f"within 0.01 texel {int(near.sum())}", flush=True)
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])
    assert st["phi_max_err"] <= 1e-5
    assert st["hit_point_max_err"] <= HITPT_ERR
    return st



def test_fp32_disc_orbit_small_matches_oracle(renderer):
    print("[case] fp32 disc_orbit 480x270", flush=True)
    _check_fp32(renderer, scenes.disc_orbit(480, 270, frames=1), "fp32 disc_orbit 480x270")


@pytest.mark.parametrize("nstep", [16, 23])
def test_fp32_other_step_counts_match_oracle(renderer, nstep):
    # the FP32 march's instance for a step count known only at run time (the
    # reference's 20 steps take an instance compiled for them): blocks of four
    # steps with remainders of 2 and 1 inbound, 3 and 2 outbound (odd sizes:
    # no pixel centre on the camera's mirror plane, so no texel ties)
    p = scenes.disc_orbit(317, 179, frames=1)
    p.options.nstep = nstep
    print(f"[case] fp32 disc_orbit 317x179 nstep {nstep}", flush=True)
    _check_fp32(renderer, p, f"fp32 disc_orbit nstep {nstep}")


def test_fp32_disc_orbit_1080p_matches_oracle(renderer):
    print("[case] fp32 disc_orbit 1920x1080 frame 0", flush=True)
    st = _check_fp32(renderer, scenes.disc_orbit(frames=1), "fp32 disc_orbit 1080p")
    assert st["rgb_mismatch_frac"] <= 1e-4


def test_fp32_path_frames_1080p_match_oracle(renderer):
    # the benchmarked workload itself: the 120-frame 1080p path as one graph
    # (one march launch over all frames, one FP64 launch over their queues);
    # frames 0, 59 and 119 of its output against the oracle
    p = scenes.disc_orbit(frames=120)
    renderer.set_textures(p.textures)
    print("[case] fp32 disc_orbit path 120 frames 1920x1080", flush=True)
    path = renderer.path(p.scene, p.options, p.cameras)
    try:
        path.launch()
        path.launch()  # a replay leaves the same frames
        for k in (0, 59, 119):
            print(f"[case] path frame {k}", flush=True)
            rgb, hit = path.read(k)
            r = oracle.render_preset(p.with_camera(p.cameras[k]), THREADS, ("rgb", "hit"))
            st = compare({"rgb": rgb, "hit": hit}, r)
            _log(f"path frame {k}", st)
            assert st["class_mismatch_frac"] <= 1e-4
            assert st["rgb_mismatch_frac"] <= 1e-4
    finally:
        path.close()


@pytest.mark.parametrize("precision", [abi.PREC_FP32, abi.PREC_FP64])
def test_camera_turned_away_matches_oracle(renderer, precision):
    # quirk Q2: phi0 = atan(b / Y0) with Y0 < 0 for pixel lines that point
    # away from the hole (blackhole_solution_test.cc:193).  The camera, 1.5
    # above the disc at 8 from the hole, looks 100 degrees away from it, so
    # most of its 60-degree frame has Y0 < 0 (phi0 then points away from the
    # camera, and those rays end mostly on the horizon, as in the reference);
    # the FP32 march starts their first chord at the camera (angle pi in its
    # frame) and forces the first step's tests.  The rest of the frame sees
    # the disc.
    import math

    cam = scenes.look_at(480, 270, math.pi / 3, (-8.0, 0.0, 1.5),
                         target=(-8.0 + 20 * math.cos(math.radians(100)),
                                 20 * math.sin(math.radians(100)), 1.5))
    p = scenes.disc_orbit(480, 270, frames=1, precision=precision).with_camera(cam)
    print(f"[case] camera turned away 480x270 precision {precision}", flush=True)
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    y0neg = sum(oracle.ray_setup(p.scene, p.options, x, y).Fp[1] < 0
                for y in range(0, 270, 15) for x in range(0, 480, 15))
    assert 100 < y0neg < 18 * 32 - 100  # both signs of Y0
    assert (r["hit"] == abi.HIT_DISC).sum() > 5000
    if precision == abi.PREC_FP64:
        st = compare(g, r)
        _log("camera turned away fp64", st)
        assert st["class_mismatch"] <= _allowed(st["pixels"])
        assert st["rgb_mismatch"] <= _allowed(st["pixels"])
        assert st["steps_mismatch"] == 0
        assert st["phi_max_err"] <= 1e-9
    else:
        _check_fp32_outputs(p, g, r, "fp32 camera turned away")


def test_fp32_disc_orbit_other_frames(renderer):
    base = scenes.disc_orbit(640, 360, frames=120)
    for k in (17, 59, 101):
        print(f"[case] fp32 disc_orbit 640x360 frame {k}", flush=True)
        _check_fp32(renderer, base.with_camera(base.cameras[k]), f"fp32 disc_orbit 640x360 frame {k}")


def _check_spheres(renderer, p, name):
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log(name, st)
    for cls in (abi.HIT_HORIZON, abi.HIT_OBJECT, abi.HIT_NONE):
        assert (r["hit"] == cls).sum() > 100, cls
    return st


def test_fp32_spheres_spheres_match_oracle(renderer):
    print("[case] fp32 spheres 640x360 (16 spheres, escape, background)", flush=True)
    st = _check_spheres(renderer, scenes.spheres(640, 360), "fp32 spheres 640x360")
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])
    assert st["phi_max_err"] <= 1e-5
    assert st["hit_point_max_err"] <= HITPT_ERR


def test_fp32_spheres_720p_matches_oracle(renderer):
    print("[case] fp32 spheres 1280x720", flush=True)
    st = _check_spheres(renderer, scenes.spheres(), "fp32 spheres 1280x720")
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])
    assert st["hit_point_max_err"] <= HITPT_ERR


@pytest.mark.parametrize("precision", [abi.PREC_FP32, abi.PREC_FP64])
def test_sphere_on_the_captured_chord(renderer, precision):
    # a sphere between the photon sphere and the horizon: captured rays end
    # with a straight chord from r ~ 3M to B (blackhole_solution_test.cc:
    # 264-272), and the reference tests every object on it, so the sphere is
    # hit before the horizon by the rays whose chord passes through it
    p = scenes.spheres(320, 180, precision=precision)
    s = abi.Scene.from_buffer_copy(p.scene)
    s.n_spheres = 1
    s.spheres[0].center[:] = (-0.9, 0.35, 0.0)
    s.spheres[0].radius = 0.3
    p.scene = s
    print(f"[case] sphere on the captured chord 320x180 precision {precision}", flush=True)
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log(f"sphere on the captured chord {precision}", st)
    # the sphere is met on the chord to the centre (steps = nstep)
    assert ((r["hit"] == abi.HIT_OBJECT) & (r["steps"] == 20)).sum() > 100
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])


def _sphere_scene(cam, spheres, precision=abi.PREC_FP32):
    """The spheres preset's hole, background and options with this camera and
    these (centre, radius) spheres (at most 16, textures of the preset)."""
    p = scenes.spheres(cam.width, cam.height, precision=precision)
    s = abi.Scene.from_buffer_copy(p.scene)
    s.camera = cam
    s.n_spheres = len(spheres)
    for i, (c, rad) in enumerate(spheres):
        s.spheres[i].center[:] = c
        s.spheres[i].radius = rad
    p.scene = s
    return p


def _check_fp32_spheres(renderer, p, name, min_objects=100):
    g = renderer.render_preset(p, ALL)
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(g, r)
    _log(name, st)
    assert (r["hit"] == abi.HIT_OBJECT).sum() > min_objects
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])
    assert st["hit_point_max_err"] <= HITPT_ERR
    return g, r


def test_fp32_spheres_ring_in_the_orbit_planes(renderer):
    # 12 spheres on a ring around the hole, in the plane of the camera: the
    # orbit planes of the middle rows cut more spheres than a lane keeps
    # (those rays are settled by the FP64 pass), and rays bent around the
    # hole meet the spheres behind it on a later revolution.
    cam = scenes.look_at(320, 180, math.pi / 3, (-16.0, 0.0, 0.3))
    ring = [((9.0 * math.cos(math.pi * (2 * k + 1) / 12),
              9.0 * math.sin(math.pi * (2 * k + 1) / 12), 0.0), 1.0) for k in range(12)]
    p = _sphere_scene(cam, ring)
    print("[case] fp32 ring of 12 spheres 320x180", flush=True)
    g, r = _check_fp32_spheres(renderer, p, "fp32 ring of 12 spheres")
    # spheres on both sides of the hole are seen
    assert len(np.unique(r["steps"][r["hit"] == abi.HIT_OBJECT])) > 5


def test_fp32_sphere_around_the_hole(renderer):
    # a sphere that contains the hole (its circle in every orbit plane
    # surrounds B: every swept angle is in its window), and two outside it
    cam = scenes.look_at(320, 180, math.pi / 3, (-15.0, 0.0, 2.0))
    p = _sphere_scene(cam, [((0.6, 0.2, 0.0), 3.5), ((4.0, -6.0, 1.0), 1.2),
                            ((-2.0, 7.0, -1.0), 0.8)])
    print("[case] fp32 sphere around the hole 320x180", flush=True)
    _check_fp32_spheres(renderer, p, "fp32 sphere around the hole", min_objects=5000)


def test_fp32_spheres_camera_turned_away(renderer):
    # rays with Y0 < 0, whose first chord runs from the camera (angle pi in
    # their frame, forced tests) past the hole, meeting spheres near the hole
    # on it; spheres in front of the camera for the other rays
    cam = scenes.look_at(320, 180, math.pi / 3, (-8.0, 0.0, 1.5),
                         target=(-8.0 + 20 * math.cos(math.radians(100)),
                                 20 * math.sin(math.radians(100)), 1.5))
    near = [((2.2 * math.cos(math.radians(a)), 2.2 * math.sin(math.radians(a)), z), 0.45)
            for a, z in ((0, 0.1), (60, 0.3), (120, 0.5), (180, 0.6), (240, 0.0), (300, -0.3))]
    ahead = [((-8.0 + 6 * math.cos(math.radians(a)), 6 * math.sin(math.radians(a)), z), 0.9)
             for a, z in ((80, 1.0), (100, 2.5), (120, 0.0), (95, -1.0))]
    p = _sphere_scene(cam, near + ahead)
    print("[case] fp32 spheres, camera turned away 320x180", flush=True)
    g, r = _check_fp32_spheres(renderer, p, "fp32 spheres camera turned away")
    assert ((r["hit"] == abi.HIT_OBJECT) & (r["steps"] == 1)).sum() > 1000


def test_fp64_spheres_matches_oracle(renderer):
    print("[case] fp64 spheres 320x180", flush=True)
    st = _check_spheres(renderer, scenes.spheres(320, 180, precision=abi.PREC_FP64),
                        "fp64 spheres 320x180")
    assert st["class_mismatch_frac"] <= 1e-4
    assert st["rgb_mismatch_frac"] <= 1e-4
    assert st["phi_max_err"] <= 1e-9


@pytest.mark.parametrize("precision", [abi.PREC_FP32, abi.PREC_FP64])
def test_path_graph_matches_eager(renderer, precision):
    # FP32: one march launch over all frames and one band launch (the frames'
    # band queues must be drained and cleared between replays); FP64: a launch
    # per frame.
    p = scenes.disc_orbit(320, 180, frames=6)
    o = scenes.options(precision=precision, band=1e-3)
    renderer.set_textures(p.textures)
    print(f"[case] path graph 6 frames 320x180 precision {precision}", flush=True)
    path = renderer.path(p.scene, o, p.cameras)
    try:
        for replay in range(3):  # every replay must equal the eager render
            path.launch()
            for k in range(len(p.cameras)):
                rgb, hit = path.read(k)
                e = renderer.render(p.with_camera(p.cameras[k]).scene, o, ("rgb", "hit"))
                assert np.array_equal(rgb, e["rgb"]), (replay, k)
                assert np.array_equal(hit, e["hit"]), (replay, k)
        if precision == abi.PREC_FP32:
            # the graph's own timing events around its two kernels
            m_ms, b_ms = path.last_ms()
            assert m_ms > 0 and b_ms > 0
            march_ms, band_ms = path.time(iters=2)
            assert march_ms > 0 and band_ms > 0
            rgb, hit = path.read(3)  # timing launches leave the same frames behind
            e = renderer.render(p.with_camera(p.cameras[3]).scene, o, ("rgb", "hit"))
            assert np.array_equal(rgb, e["rgb"]) and np.array_equal(hit, e["hit"])
        else:
            with pytest.raises(BHError) as err:
                path.time(iters=1)
            assert err.value.code == abi.BH_ERR_UNSUPPORTED
            with pytest.raises(BHError) as err:
                path.last_ms()
            assert err.value.code == abi.BH_ERR_UNSUPPORTED
    finally:
        path.close()


def test_misaligned_device_planes(renderer):
    # bh_render_device into caller planes at odd byte offsets gives the frame
    # bh_render gives, band-pass pixels included, and writes nothing outside
    # the planes
    import torch

    p = scenes.disc_orbit(640, 360, frames=1)
    o = abi.Options.from_buffer_copy(p.options)
    o.band = 1e-2  # a wide band: many pixels go through the band pass
    renderer.set_textures(p.textures)
    print("[case] misaligned device planes 640x360", flush=True)
    want = renderer.render(p.scene, o, ("rgb", "hit"))
    npix = 640 * 360
    buf_rgb = torch.zeros(3 * npix + 8, dtype=torch.uint8, device="cuda")
    buf_hit = torch.zeros(npix + 8, dtype=torch.uint8, device="cuda")
    torch.cuda.synchronize()
    outs = abi.Outputs()
    outs.rgb = buf_rgb.data_ptr() + 1
    outs.hit = buf_hit.data_ptr() + 3
    assert renderer.lib.bh_render_device(renderer.ctx, C.byref(p.scene), C.byref(o),
                                         C.byref(outs)) == abi.BH_OK
    assert renderer.lib.bh_synchronize(renderer.ctx) == abi.BH_OK
    rgb = buf_rgb.cpu().numpy()[1:1 + 3 * npix].reshape(360, 640, 3)
    hit = buf_hit.cpu().numpy()[3:3 + npix].reshape(360, 640)
    assert np.array_equal(rgb, want["rgb"]) and np.array_equal(hit, want["hit"])
    # bytes outside the planes are untouched
    assert buf_rgb[0].item() == 0 and buf_rgb[1 + 3 * npix:].sum().item() == 0
    assert buf_hit[:3].sum().item() == 0 and buf_hit[3 + npix:].sum().item() == 0


@pytest.mark.parametrize("size", [(317, 179), (322, 181)])
def test_fp32_odd_widths_match_oracle(renderer, size):
    # partial 16-pixel tiles at the right edge and bottom
    w, h = size
    print(f"[case] fp32 disc_orbit {w}x{h}", flush=True)
    _check_fp32(renderer, scenes.disc_orbit(w, h, frames=1), f"fp32 disc_orbit {w}x{h}")


@pytest.mark.parametrize("precision", [abi.PREC_FP32, abi.PREC_FP64])
def test_tiny_frames_match_oracle(renderer, precision):
    # frames of one row or column and fewer pixels than a wave: every lane
    # outside the frame idles, the rest match the oracle
    for w, h in ((1, 1), (1, 7), (7, 1), (2, 3), (3, 2), (17, 1)):
        p = scenes.disc_orbit(w, h, frames=1, precision=precision)
        print(f"[case] {w}x{h} precision {precision}", flush=True)
        g = renderer.render_preset(p, ALL)
        r = oracle.render_preset(p, 1, ALL)
        assert np.array_equal(g["hit"], r["hit"]), (w, h)
        assert np.array_equal(g["rgb"], r["rgb"]), (w, h)
        if precision == abi.PREC_FP64:
            assert np.array_equal(g["steps"], r["steps"]), (w, h)
        st = compare(g, r)
        assert st["phi_max_err"] <= (1e-9 if precision == abi.PREC_FP64 else 1e-5), (w, h)


def test_camera_inside_photon_sphere_takes_the_fp64_march(renderer):
    # r0 = 1.3 < 3M = 1.5: every ray's march runs outward (du < 0) and its
    # swept angle shrinks, which the FP32 event tests do not allow for, so an
    # FP32 request is served by the FP64 reference march: both precisions
    # give the same bytes in every plane, and the oracle's classes and angles
    print("[case] camera inside the photon sphere 160x90", flush=True)
    cam = scenes.orbit_camera(1.3, 80.0, 0.0, 160, 90, math.pi / 3)
    out = {}
    for prec in (abi.PREC_FP32, abi.PREC_FP64):
        p = scenes.disc_orbit(160, 90, frames=1, precision=prec).with_camera(cam)
        out[prec] = renderer.render_preset(p, ALL)
    for k in ALL:
        assert np.array_equal(out[abi.PREC_FP32][k], out[abi.PREC_FP64][k], equal_nan=True), k
    r = oracle.render_preset(p, THREADS, ALL)
    st = compare(out[abi.PREC_FP64], r)
    _log("camera inside the photon sphere", st)
    assert st["class_mismatch_frac"] == 0
    assert st["steps_mismatch"] == 0
    assert st["phi_max_err"] <= 1e-9


def test_fp32_8k_frame_matches_oracle(renderer):
    # 7680x4320 (33 M rays in one launch, 270 tile rows): the oracle on three
    # bands of rows, through the shadow, the photon ring and the disc edge
    p = scenes.disc_orbit(7680, 4320, frames=1)
    print("[case] fp32 disc_orbit 7680x4320", flush=True)
    g = renderer.render_preset(p, ALL)
    for y0, y1 in ((0, 24), (2100, 2220), (3000, 3048)):
        r = oracle.render_preset(p, THREADS, ALL, rows=(y0, y1))
        gs = {k: v[y0:y1] for k, v in g.items()}
        rs = {k: v[y0:y1] for k, v in r.items()}
        _check_fp32_outputs(p, gs, rs, f"fp32 8k rows {y0}..{y1}")
    assert (g["hit"] == abi.HIT_DISC).sum() > 1_000_000


def test_error_paths(renderer):
    p = scenes.disc_orbit(64, 36, frames=1)
    renderer.set_textures(p.textures)
    s = abi.Scene.from_buffer_copy(p.scene)
    assert renderer.lib.bh_texture_clear(renderer.ctx, 63) == abi.BH_OK
    s.disc.texture = 63  # empty slot -> Material::SetTexture failure
    assert renderer.check_scene(s, p.options) == abi.BH_ERR_EMPTY_TEXTURE
    s = abi.Scene.from_buffer_copy(p.scene)
    s.hole.mass = 0.0
    assert renderer.check_scene(s, p.options) == abi.BH_ERR_INVALID_ARG
    s = abi.Scene.from_buffer_copy(p.scene)
    s.camera.width = 0
    assert renderer.check_scene(s, p.options) == abi.BH_ERR_INVALID_ARG
    assert renderer.check_scene(p.scene, p.options) == abi.BH_OK


@pytest.mark.parametrize("case", ["disc_orbit", "plain_hole", "turned_away", "near_band"])
def test_turning_points_match_the_reference_bisection(renderer, case):
    # the device's per-ray turning point (closed-form start, one FP64 Newton
    # step, the 20-step bisection cell; bisection near a cell edge) lies in
    # the bisection cell of the reference's SolveG (blackhole_solution.h:
    # 38-55, the oracle's or_solve_g) for the pixel's own impact parameter:
    # the device forms the cell's lower end as l0 + k cell, the reference by
    # 20 midpoint halvings, so the two agree to a few ulp (a cell is ~3e-7 of
    # the range: a wrong cell would differ by far more than 1e-12)
    if case == "disc_orbit":
        p = scenes.disc_orbit(320, 180, frames=1)
    elif case == "plain_hole":
        p = scenes.plain_hole()
        p.options.precision = abi.PREC_FP32
    elif case == "turned_away":
        cam = scenes.look_at(320, 180, math.pi / 3, (-8.0, 0.0, 1.5),
                             target=(-8.0 + 20 * math.cos(math.radians(100)),
                                     20 * math.sin(math.radians(100)), 1.5))
        p = scenes.disc_orbit(320, 180, frames=1).with_camera(cam)
    else:
        # a narrow frame around the shadow's edge: every ray near b_c
        # (q = (b_c/b)^2 up to the band), with no band (every ray takes the
        # closed form)
        cam = scenes.look_at(320, 180, math.pi / 900, (-20.0, 0.0, 0.0),
                             target=(0.0, 2.62, 0.0))
        p = scenes.disc_orbit(320, 180, frames=1).with_camera(cam)
        p.options.band = 0.0
    s, o = p.scene, p.options
    W, H = s.camera.width, s.camera.height
    renderer.set_textures(p.textures)
    print(f"[case] turning points {case} {W}x{H}", flush=True)
    got = np.empty((H, W), dtype=np.float64)
    assert renderer.lib.bh_turning_points(renderer.ctx, C.byref(s), C.byref(o),
                                          got.ctypes.data_as(C.POINTER(C.c_double))) == abi.BH_OK
    lib = oracle.load()
    M = s.hole.mass
    want = np.zeros_like(got)
    for y in range(H):
        for x in range(W):
            ray = oracle.ray_setup(s, o, x, y)
            if not (ray.captured or ray.degenerate):
                want[y, x] = lib.or_solve_g(ray.b, M, o.root)
    live = want > 0
    assert np.array_equal(got > 0, live)  # the same rays captured
    rel = np.abs(got[live] - want[live]) / want[live]
    print(f"[turning] {case}: {int(live.sum())} of {W * H} not captured, "
          f"max relative difference {rel.max():.3g}, "
          f"{int((rel > 0).sum())} not bit-equal", flush=True)
    assert live.sum() > W * H // 10
    assert rel.max() < 1e-12


def test_widest_edge_guard_settles_every_disc_hit_in_fp64(renderer):
    # edge_guard 0.01 of r_out spans several texels, so every disc hit of the
    # FP32 march lies "near a texel edge" and goes through the FP64 pass: the
    # FP64 crossing point for every disc hit
    p = scenes.disc_orbit(320, 180, frames=1)
    o = abi.Options.from_buffer_copy(p.options)
    o.edge_guard = 0.01
    renderer.set_textures(p.textures)
    print("[case] edge guard 0.01 disc_orbit 320x180", flush=True)
    g = renderer.render(p.scene, o, ALL)
    _, _, queued = renderer.time_render(p.scene, o, iters=2)
    r = oracle.render_preset(p, THREADS, ALL)
    disc = int((r["hit"] == abi.HIT_DISC).sum())
    print(f"[edge guard] queued {queued}, disc hits {disc}", flush=True)
    assert disc > 5000 and queued >= disc
    st = compare(g, r)
    _log("edge guard 0.01", st)
    assert st["class_mismatch"] <= _allowed(st["pixels"])
    assert st["rgb_mismatch"] <= _allowed(st["pixels"])
    # the disc hits' points are the FP64 pass's (the FP32 ones are ~1e-5 off)
    both = (g["hit"] == abi.HIT_DISC) & (r["hit"] == abi.HIT_DISC)
    err = np.abs(g["hitpt"][both] - r["hitpt"][both]).max()
    print(f"[edge guard] disc hit point max error {err:.3g}", flush=True)
    assert err <= 1e-9


def test_chord_sincos_within_its_stated_bounds(renderer):
    # the FP32 chord tests' sin/cos (hardware sine and cosine) against FP64 over
    # the angles a chord can span: [-pi, pi] uniformly and, log-spaced, the
    # small spans of chords far from the hole down to 1e-8 rad
    x = np.concatenate([np.linspace(-np.pi, np.pi, 1 << 20),
                        np.logspace(-8, 0, 4096), -np.logspace(-8, 0, 4096)]).astype(np.float32)
    sc = np.empty((x.size, 2), dtype=np.float32)
    print(f"[case] chord sincos, {x.size} angles", flush=True)
    assert renderer.lib.bh_chord_sincos(renderer.ctx, x.ctypes.data_as(C.POINTER(C.c_float)),
                                        sc.ctypes.data_as(C.POINTER(C.c_float)),
                                        x.size) == abi.BH_OK
    xd = x.astype(np.float64)
    es = np.abs(sc[:, 0] - np.sin(xd))
    ec = np.abs(sc[:, 1] - np.cos(xd))
    small = np.abs(xd) <= 1.0
    tiny = (np.abs(xd) >= 1e-8) & small
    rel = es[tiny] / np.abs(np.sin(xd[tiny]))
    print(f"[chord sincos] max |sin err| {es.max():.3g} ({es[small].max():.3g} for |x| <= 1), "
          f"max |cos err| {ec.max():.3g} ({ec[small].max():.3g}), "
          f"max relative sin err {rel.max():.3g}", flush=True)
    assert np.isfinite(sc).all()
    assert es.max() <= 4e-7 and ec.max() <= 4e-7
    assert es[small].max() <= 2e-7 and ec[small].max() <= 2e-7
    assert rel.max() <= 2e-6
