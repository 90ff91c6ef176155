"""The two demo programs (blackhole_8_amd/bin/blackhole_solution and
ray_tracer), the headless counterparts of the reference's
blackhole_solution_test.cc and ray_tracer_test.cc, and the frame ring they
write through.

CPU: the programs' scene sequences (--dry-run) against an independent Python
replay of the reference's loops: the disc turned by pi/180 about its axis
after each frame with its construction-time normal kept
(blackhole_solution_test.cc:407, vector_object.h:320-326), the key bindings
(:351-406; ray_tracer_test.cc:196-235) and the flat scene's movement test
(ray_tracer_test.cc:237-261); frame 0 equals the library presets.
GPU: the files the programs write equal the library's own render of the same
scenes byte for byte, and the frame ring returns what bh_render returns.
"""
import ctypes as C
import json
import math
import subprocess
from pathlib import Path

import numpy as np
import pytest

from blackhole_8_amd import Ring, abi, images, scenes
from tests import riff

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "blackhole_8_amd" / "bin"
SOLUTION = BIN / "blackhole_solution"
TRACER = BIN / "ray_tracer"


def run(exe, *args, check=True, timeout=120):
    r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True,
                       timeout=timeout)
    if check and r.returncode != 0:
        raise AssertionError(f"{exe.name} {args} -> {r.returncode}\n{r.stdout}\n{r.stderr}")
    return r


def dry(exe, *args):
    out = run(exe, "--dry-run", *args).stdout
    return [json.loads(line) for line in out.splitlines() if line.startswith("{")]


# ---------------------------------------------------------------------------
# Python replay of the reference's Object poses (object.h:58-88)
# ---------------------------------------------------------------------------
class Pose:
    def __init__(self, verts, axes=((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        self.v = np.array(verts, dtype=np.float64).reshape(-1, 3)
        self.ax = np.array(axes, dtype=np.float64)

    def move(self, which, d):
        self.v = self.v + self.ax[which] * d

    def rotate(self, which, rad):
        axes = np.ascontiguousarray(self.ax.ravel())
        verts = np.ascontiguousarray(self.v.ravel())
        abi.load().bh_object_rotate(axes.ctypes.data_as(C.POINTER(C.c_double)),
                                    verts.ctypes.data_as(C.POINTER(C.c_double)),
                                    len(self.v), which, rad)
        self.ax = axes.reshape(3, 3)
        self.v = verts.reshape(-1, 3)


def key(cam, fov, k, move_d, rotate_d, faster=10):
    up = k.isupper()
    m = move_d * faster if up else move_d
    r = rotate_d * faster if up else rotate_d
    kl = k.lower()
    table = {"w": ("m", 0, m), "s": ("m", 0, -m), "a": ("m", 1, -m), "d": ("m", 1, m),
             "z": ("m", 2, -m), "x": ("m", 2, m), "q": ("r", 0, -r), "e": ("r", 0, r),
             "i": ("r", 1, r), "k": ("r", 1, -r), "j": ("r", 2, -r), "l": ("r", 2, r)}
    if k == "-":
        return min(max(0.0, fov - math.pi / 54), math.nextafter(math.pi, 0)), True
    if k == "+":
        return min(max(0.0, fov + math.pi / 54), math.nextafter(math.pi, 0)), True
    if kl in table:
        op, which, amt = table[kl]
        (cam.move if op == "m" else cam.rotate)(which, amt)
        return fov, True
    return fov, False


def assert_vec(got, want):
    # the programs print every double round-trip exact
    assert np.array_equal(np.array(got, dtype=np.float64), np.asarray(want, dtype=np.float64)), \
        (got, want)


def assert_camera(got, cam_pose, fov, w, h):
    assert_vec(got["position"], cam_pose.v[0])
    assert_vec(got["axis_x"], cam_pose.ax[0])
    assert_vec(got["axis_y"], cam_pose.ax[1])
    assert_vec(got["axis_z"], cam_pose.ax[2])
    assert (got["width"], got["height"]) == (w, h)
    assert got["fov"] == fov


# ---------------------------------------------------------------------------
# CPU
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("exe", [SOLUTION, TRACER])
def test_programs_are_built_and_print_usage(exe):
    assert exe.exists(), "built by __graft_entry__.build()"
    r = run(exe, "--help")
    assert r.stdout.startswith("usage:")


@pytest.mark.parametrize("exe", [SOLUTION, TRACER])
def test_bad_arguments_fail_with_status(exe):
    for args in (("--frames", "0"), ("--frames", "x"), ("--format", "gif"),
                 ("--depth", "1", "--dry-run")):
        r = run(exe, *args, check=False)
        if "--dry-run" in args:
            assert r.returncode == 0  # the ring is not built in a dry run
            continue
        assert r.returncode == 2 and r.stderr.startswith("error:"), args


def test_solution_without_device_fails_cleanly():
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    r = run(SOLUTION, "--frames", "1", check=False)
    assert r.returncode == 2 and "no HIP device" in r.stderr


def test_solution_scene_sequence_replays_the_reference_loop():
    keys = "wQ^+,lv-Ai"
    n = len(keys) + 2
    frames = dry(SOLUTION, "--frames", n, "--keys", keys)
    assert [f["frame"] for f in frames] == list(range(n))
    cam = Pose([(-2000.0, 0.0, 400.0)], ((1, 0, 0), (0, -1, 0), (0, 0, -1)))
    fov = math.pi / 2
    hole = Pose([(0.0, 0.0, 0.0)])
    disc = Pose([(0, 0, 0), (1000, 1000, 0), (-1000, 1000, 0), (-1000, -64, 0), (1000, -1000, 0)])
    normal0 = np.array(frames[0]["normal"])
    # the construction-time normal: normalize((p3 - p1) x (p2 - p1))
    n0 = np.cross(disc.v[3] - disc.v[1], disc.v[2] - disc.v[1])
    np.testing.assert_allclose(normal0, n0 / np.linalg.norm(n0), rtol=0, atol=1e-15)
    for f, rec in enumerate(frames):
        assert_camera(rec["camera"], cam, fov, 960, 540)
        assert_vec(rec["hole"], hole.v[0])
        assert_vec(rec["disc"], disc.v)
        assert_vec(rec["normal"], normal0)  # kept through the rotations
        if f < len(keys):
            fov, handled = key(cam, fov, keys[f], 10, math.pi / 1800.0)
            if not handled:
                d = {"^": (2, 5), "v": (2, -5), "<": (1, 5), ">": (1, -5), ",": (0, 5),
                     ".": (0, -5)}[keys[f]]
                hole.move(*d)
        disc.rotate(2, math.pi / 180)
    # frame 0 is the shipped preset
    p = scenes.shipped()
    c0 = frames[0]["camera"]
    assert_vec(c0["position"], p.scene.camera.position[:])
    assert_vec(c0["axis_x"], p.scene.camera.axis_x[:])
    assert_vec(frames[0]["disc"], [p.scene.disc.vertex[i][:] for i in range(5)])
    assert_vec(frames[0]["normal"], p.scene.disc.normal[:])


def test_tracer_scene_sequence_replays_the_reference_loop():
    keys = "wlE-+k"
    n = 40  # long enough for the sliding rectangle to turn back
    frames = dry(TRACER, "--frames", n, "--keys", keys)
    cam = Pose([(-400.0, 120.0, 100.0)], ((1, 0, 0), (0, -1, 0), (0, 0, -1)))
    cam.rotate(1, -math.pi / 24)
    cam.rotate(2, math.pi / 12)
    fov = math.pi / 2
    mooni = Pose([(0, 0, 0), (100, 120, 100), (100, 0, 100), (100, 0, 0), (100, 120, 0)])
    karina = Pose([(0, 0, 0), (100, 0, 150), (100, -100, 150), (100, -100, 0), (100, 0, 0)])
    winter = Pose([(0, 0, 0), (-100, 50, 100), (-100, -50, 100), (-100, -50, 0), (-100, 50, 0)])
    direction = 1
    turned = False
    for f, rec in enumerate(frames):
        assert_camera(rec["camera"], cam, fov, 800, 450)
        for got, obj in zip(rec["rects"], (mooni, karina, winter)):
            assert_vec(got, obj.v)
        if f < len(keys):
            fov, _ = key(cam, fov, keys[f], 0.9, math.pi / 180.0)
        karina.move(0, 3 * direction)
        x = karina.v[0][0]
        if x > 100:
            direction, turned = -1, True
        elif x < 0:
            direction = 1
        mooni.move(1, 60.0)
        mooni.move(2, 50.0)
        mooni.rotate(0, math.pi / 180)
        mooni.move(1, -60.0)
        mooni.move(2, -50.0)
        winter.move(0, -100.0)
        winter.move(2, 50.0)
        winter.rotate(1, math.pi / 120)
        winter.move(0, 100.0)
        winter.move(2, -50.0)
    assert turned
    s, _ = scenes.flat_test()
    c0 = frames[0]["camera"]
    for k in ("position", "axis_x", "axis_y", "axis_z"):
        assert_vec(c0[k], getattr(s.camera, k)[:])
    for i in range(3):
        assert_vec(frames[0]["rects"][i], [s.rects[i].vertex[v][:] for v in range(5)])


# ---------------------------------------------------------------------------
# GPU
# ---------------------------------------------------------------------------
def _camera(d):
    c = abi.Camera()
    c.position[:], c.axis_x[:], c.axis_y[:], c.axis_z[:] = (
        d["position"], d["axis_x"], d["axis_y"], d["axis_z"])
    c.width, c.height, c.fov = d["width"], d["height"], d["fov"]
    return c


def _solution_scene(rec, precision):
    p = scenes.shipped(rec["camera"]["width"], rec["camera"]["height"], precision=precision)
    s = p.scene
    s.camera = _camera(rec["camera"])
    s.hole.position[:] = rec["hole"]
    for i in range(5):
        s.disc.vertex[i][:] = rec["disc"][i]
    s.disc.normal[:] = rec["normal"]
    return s, p.options, p.textures


@pytest.mark.gpu
@pytest.mark.parametrize("fp32", [False, True])
def test_solution_frames_equal_library_renders(renderer, tmp_path, fp32):
    args = ["--frames", 3, "--width", 320, "--height", 180, "--keys", "dq"]
    if fp32:
        args.append("--fp32")
    print(f"[case] blackhole_solution {args}", flush=True)
    r = run(SOLUTION, *args, "--out", tmp_path, "--format", "png" if fp32 else "ppm")
    print(r.stdout.strip().splitlines()[-1], flush=True)
    recs = dry(SOLUTION, *args)
    prec = abi.PREC_FP32 if fp32 else abi.PREC_FP64
    for rec in recs:
        s, o, tex = _solution_scene(rec, prec)
        renderer.set_textures(tex)
        want = renderer.render(s, o, ("rgb", "hit"))
        got = images.load(tmp_path / f"frame_{rec['frame']:05d}.{'png' if fp32 else 'ppm'}")
        assert np.array_equal(got, want["rgb"]), rec["frame"]
        assert (want["hit"] == abi.HIT_DISC).sum() > 1000


@pytest.mark.gpu
def test_tracer_frames_equal_library_renders(renderer, tmp_path):
    args = ["--frames", 4, "--keys", "jW"]
    print(f"[case] ray_tracer {args}", flush=True)
    r = run(TRACER, *args, "--out", tmp_path, "--depth", 3)
    print(r.stdout.strip().splitlines()[-1], flush=True)
    recs = dry(TRACER, *args)
    s, tex = scenes.flat_test()
    renderer.set_textures(tex)
    for rec in recs:
        s.camera = _camera(rec["camera"])
        for i in range(3):
            for v in range(5):
                s.rects[i].vertex[v][:] = rec["rects"][i][v]
        want = renderer.render_flat(s, ("rgb", "hit"))
        got = images.load(tmp_path / f"frame_{rec['frame']:05d}.ppm")
        assert np.array_equal(got, want["rgb"]), rec["frame"]


@pytest.mark.gpu
def test_avi_output_holds_the_library_renders(renderer, tmp_path):
    # --format avi: one MJPG AVI whose frames are the JPEG encodings of the
    # library's renders, in order
    args = ["--frames", 3, "--width", 320, "--height", 180, "--keys", "Wd", "--fp32"]
    print(f"[case] blackhole_solution avi {args}", flush=True)
    run(SOLUTION, *args, "--out", tmp_path, "--format", "avi", "--quiet")
    total, frames = riff.avi_frames((tmp_path / "video.avi").read_bytes())
    recs = dry(SOLUTION, *args)
    assert total == len(frames) == len(recs) == 3
    for rec, jpg in zip(recs, frames):
        s, o, tex = _solution_scene(rec, abi.PREC_FP32)
        renderer.set_textures(tex)
        want = renderer.render(s, o, ("rgb",))["rgb"]
        assert jpg == images.encode_jpeg(want, 95), rec["frame"]


@pytest.mark.gpu
@pytest.mark.parametrize("depth", [2, 3])
def test_ring_returns_what_render_returns(renderer, depth):
    p = scenes.disc_orbit(192, 108, frames=6)
    renderer.set_textures(p.textures)
    print(f"[case] frame ring depth {depth}", flush=True)
    with Ring(renderer, 192, 108, depth) as ring:
        pending = []
        got = []
        for cam in p.cameras:
            s = p.scene
            s.camera = cam
            if len(pending) == depth:  # every slot held: hand the oldest back
                slot = pending.pop(0)
                got.append(ring.wait(slot))
                ring.release(slot)
            pending.append(ring.render((s, p.options)))
        for slot in pending:
            got.append(ring.wait(slot))
            ring.release(slot)
    assert len(got) == len(p.cameras)
    for cam, (rgb, hit, ms) in zip(p.cameras, got):
        s = p.scene
        s.camera = cam
        want = renderer.render(s, p.options, ("rgb", "hit"))
        assert np.array_equal(rgb, want["rgb"]) and np.array_equal(hit, want["hit"])
        assert ms > 0


@pytest.mark.gpu
def test_ring_errors(renderer):
    from blackhole_8_amd import BHError

    p = scenes.disc_orbit(64, 36, frames=1)
    renderer.set_textures(p.textures)
    with Ring(renderer, 64, 36, 2) as ring:
        with pytest.raises(BHError) as e:
            ring.wait(0)  # holds no frame yet
        assert e.value.code == abi.BH_ERR_INVALID_ARG
        big = scenes.disc_orbit(65, 36, frames=1)
        with pytest.raises(BHError) as e:
            ring.render((big.scene, big.options))
        assert e.value.code == abi.BH_ERR_INVALID_ARG
        slot = ring.render((p.scene, p.options))
        ring.wait(slot)
        ring.release(slot)
