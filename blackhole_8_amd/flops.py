"""Algorithmic work of the reference march, per ray.

The count follows the reduced orbit-plane form of the reference's step
(blackhole_solution_test.cc:217-239): add/mul/compare = 1, FMA = 2, and each
sqrt, rsqrt, rcp, div, sin, cos, atan, acos = 1.  It is a property of the
algorithm, not of this implementation; the kernels issue fewer instructions
where they skip work the reference does (per-step positions are only formed
on steps where an object can be reached).

  F_STEP_DISC    35  u update 1, G 5, rsqrt 1, phi update 3, rcp 1, sin+cos 2,
                     position 2, 2-D horizon chord test 14, disc sign test 5,
                     loop 1
  F_STEP_NODISC  30  the same without the disc sign test
  F_SETUP       240  pixel point, orbit-plane frame, b, Y0, r0, phi0 (60 FP32)
                     plus SolveG's 20 bisection steps (180 FP64)
  F_HIT          40  crossing interpolation, radius window, texel address

Scenes with spheres (the reference tests every sphere on every chord,
blackhole_solution_test.cc:217-239 with object_manager.h:69-85; in the orbit
plane only the spheres that cut it can be met):
  F_SPH_RAY       9  the orbit plane's normal (a cross product), per ray
  F_SPH_PLANE     7  per ray and sphere: the centre's distance from the plane
                     (dot product 5) and its comparison with the radius 2
  F_SPH_STEP     20  per step and sphere that cuts the plane: the centre in the
                     chord's frame 6 and the in-plane circle test 14 (the
                     horizon chord test's count)
  F_BG           40  per ray that escapes to a background: direction of the
                     last chord, atan2, asin, texel address

Work of a frame = F_SETUP * rays + F_STEP * sum(steps) + F_HIT * hits, plus
for spheres (F_SPH_RAY + F_SPH_PLANE * spheres) * rays +
F_SPH_STEP * sum(steps * spheres in the pixel's plane), plus F_BG * escaped
rays, with the per-pixel step counts and hit classes of the render itself.
"""
from __future__ import annotations

import numpy as np

F_STEP_DISC = 35
F_STEP_NODISC = 30
F_SETUP = 240
F_HIT = 40
F_SPH_RAY = 9
F_SPH_PLANE = 7
F_SPH_STEP = 20
F_BG = 40


def spheres_in_plane(scene) -> np.ndarray:
    """Per pixel (H, W), the number of the scene's spheres that cut the
    pixel's orbit plane: the plane through the hole B spanned by F = C - B and
    the pixel point p = f axis_x - s_x axis_y - s_y axis_z - B with
    s_x = W/2 - x, s_y = H/2 - y (the kernels' pixel point, Camera::PixelVector
    at camera.h:49-60 relative to B), normal p x F.  Degenerate pixels count 0."""
    from . import abi

    cam = scene.camera
    w, h = cam.width, cam.height
    f = abi.load().bh_focal_length(w, cam.fov)
    b = np.array(scene.hole.position[:], dtype=np.float64)
    ax = np.array(cam.axis_x[:], dtype=np.float64)
    ay = np.array(cam.axis_y[:], dtype=np.float64)
    az = np.array(cam.axis_z[:], dtype=np.float64)
    sx = w / 2.0 - np.arange(w, dtype=np.float64)
    sy = h / 2.0 - np.arange(h, dtype=np.float64)
    pv = (ax * f - b)[None, None, :] - sx[None, :, None] * ay - sy[:, None, None] * az
    F = np.array(cam.position[:], dtype=np.float64) - b
    n = np.cross(pv, F)
    with np.errstate(invalid="ignore", divide="ignore"):
        n /= np.linalg.norm(n, axis=-1, keepdims=True)
    cnt = np.zeros((h, w), dtype=np.int32)
    for i in range(scene.n_spheres):
        sp = scene.spheres[i]
        c = np.array(sp.center[:], dtype=np.float64) - b
        cnt += np.abs(n @ c) < sp.radius
    return cnt


def frame_flops(steps: np.ndarray, hit: np.ndarray, has_disc: bool,
                in_plane: np.ndarray | None = None, n_spheres: int = 0,
                background: bool = False) -> float:
    """Algorithmic work of one frame from its per-pixel steps and hit classes;
    in_plane (spheres_in_plane) and n_spheres for a scene with spheres,
    background for an escape scene with a background texture."""
    f_step = F_STEP_DISC if has_disc else F_STEP_NODISC
    n = steps.size
    st = steps.astype(np.int64)
    work = F_SETUP * n + f_step * int(st.sum()) + F_HIT * int((hit != 0).sum())
    if n_spheres > 0:
        work += (F_SPH_RAY + F_SPH_PLANE * n_spheres) * n
        if in_plane is not None:
            work += F_SPH_STEP * int((st * in_plane.astype(np.int64).reshape(st.shape)).sum())
    if background:
        work += F_BG * int((hit == 0).sum())
    return float(work)
