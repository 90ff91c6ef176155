"""Host-side code of the C ABI that needs no GPU: Object pose operations,
camera framing, the Annulus constructor, the seeded input generators."""
import ctypes as C
import math

import numpy as np
import pytest

from blackhole_8_amd import abi, scenes

LIB = abi.load()
EPS_CBRT = np.cbrt(np.finfo(np.float64).eps)  # float_equal tolerance (utility.h:17-27)


def rotate(axes, which, rad, verts=None):
    a = np.ascontiguousarray(axes, dtype=np.float64).ravel().copy()
    v = None if verts is None else np.ascontiguousarray(verts, dtype=np.float64).copy()
    LIB.bh_object_rotate(a.ctypes.data_as(C.POINTER(C.c_double)),
                         None if v is None else v.ctypes.data_as(C.POINTER(C.c_double)),
                         0 if v is None else len(v), which, rad)
    return a.reshape(3, 3), v


def test_object_rotations_known_answers():
    """include/blackhole/object/object_test.cc:38-54 on Object's default axes."""
    I = np.eye(3)
    cases = [
        (0, math.pi / 2, [(1, 0, 0), (0, 0, 1), (0, -1, 0)]),   # RotateX(pi/2)
        (1, math.pi / 2, [(0, 0, -1), (0, 1, 0), (1, 0, 0)]),   # RotateY(pi/2)
        (2, math.pi / 2, [(0, 1, 0), (-1, 0, 0), (0, 0, 1)]),   # RotateZ(pi/2)
    ]
    for which, rad, want in cases:
        ax, _ = rotate(I, which, rad)
        assert np.all(np.abs(ax - np.array(want, float)) <= EPS_CBRT), (which, ax)
        back, _ = rotate(ax, which, -rad)
        assert np.all(np.abs(back - I) <= EPS_CBRT)


def test_object_rotation_moves_vertices_about_vertex0():
    # Object::RotateZ rotates every vertex about position() = vertex[0]
    verts = np.array([[1.0, 2.0, 0.0], [2.0, 2.0, 0.0], [1.0, 3.0, 5.0]])
    _, v = rotate(np.eye(3), 2, math.pi / 2, verts)
    np.testing.assert_allclose(v, [[1, 2, 0], [1, 3, 0], [0, 2, 5]], atol=1e-15)


def test_camera_default_frame_and_look_at():
    # camera.h:32: default axes (1,0,0), (0,-1,0), (0,0,-1); a camera on -x
    # looking at the origin with +z up has exactly that frame
    c = scenes.look_at(320, 180, 1.0, (-15.0, 0.0, 0.0))
    np.testing.assert_allclose(np.array([c.axis_x, c.axis_y, c.axis_z]),
                               [[1, 0, 0], [0, -1, 0], [0, 0, -1]], atol=1e-15)
    # orbit camera: looks at the hole, right-handed like the default frame
    c = scenes.disc_orbit_camera(7)
    x, y, z = (np.array(v) for v in (c.axis_x, c.axis_y, c.axis_z))
    pos = np.array(c.position)
    np.testing.assert_allclose(x, -pos / np.linalg.norm(pos), atol=1e-15)
    np.testing.assert_allclose(np.cross(x, y), z, atol=1e-15)
    assert np.linalg.norm(pos) == pytest.approx(20.0)
    assert math.degrees(math.acos(pos[2] / 20.0)) == pytest.approx(80.0)
    assert math.degrees(math.atan2(pos[1], pos[0])) == pytest.approx(21.0)


def test_annulus_constructor():
    # vector_object.h:320-326 with the shipped scene's octal p3.y = -0100 = -64
    a = scenes.annulus((1000, 1000, 0), (-1000, 1000, 0), (-1000, -64, 0), (1000, -1000, 0),
                       1000.0, 100.0, 3)
    v = np.array(a.vertex)
    np.testing.assert_array_equal(v[0], [0, 0, 0])
    np.testing.assert_array_equal(v[3], [-1000, -64, 0])
    np.testing.assert_allclose(np.array(a.normal), [0, 0, -1], atol=0)
    assert (a.r_outer, a.r_inner, a.texture) == (1000.0, 100.0, 3)


def test_splitmix64_published_vectors():
    # SplitMix64 seeded with 0 (Vigna's reference implementation)
    s = C.c_uint64(0)
    got = [LIB.bh_splitmix64(C.byref(s)) for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    s = C.c_uint64(12345)
    u = [LIB.bh_uniform(C.byref(s), 2.0, 3.0) for _ in range(1000)]
    assert min(u) >= 2.0 and max(u) < 3.0


def test_background_structure_and_determinism():
    bg = scenes.gen_background(64, 128, (16, 8), seed=99)
    assert bg.shape == (64, 128, 3)
    ys, xs = np.meshgrid(np.arange(64), np.arange(128), indexing="ij")
    light = ((xs // 8) + (ys // 8)) % 2 == 1
    assert (bg[light] >= 192).all() and (bg[~light] < 64).all()
    assert np.array_equal(bg, scenes.gen_background(64, 128, (16, 8), seed=99))
    assert not np.array_equal(bg, scenes.gen_background(64, 128, (16, 8), seed=100))


def test_disc_texture_bands():
    t = scenes.gen_disc_texture(256, 256, 16, seed=5)
    assert t.shape == (256, 256, 4) and (t[..., 3] == 255).all()
    # warm bands (R >= 192) at even band index, cool (B >= 190) at odd
    ys, xs = np.meshgrid(np.arange(256), np.arange(256), indexing="ij")
    rho2 = np.sqrt((2 * xs + 1 - 256) ** 2 + (2 * ys + 1 - 256) ** 2).astype(int)
    band = (rho2 // 32) % 2
    assert (t[band == 0][:, 0] >= 192).all()
    assert (t[band == 1][:, 2] >= 190).all()


def test_spheres_spheres_placement():
    p = scenes.spheres(64, 36)
    s = p.scene
    assert s.n_spheres == 16
    cam = np.array(s.camera.position)
    for i in range(16):
        sp = s.spheres[i]
        r = np.linalg.norm(np.array(sp.center))
        assert 5.0 <= r <= 30.0
        assert 0.3 <= sp.radius <= 1.5
        assert np.linalg.norm(np.array(sp.center) - cam) >= sp.radius + 0.5
        assert p.textures[sp.texture].shape == (256, 256, 3)


def test_spheres_in_plane_matches_the_oracle_ray_frames():
    """flops.spheres_in_plane against the orbit planes of the oracle's own ray
    set-up (pv and F of or_ray_setup: the plane through B spanned by them),
    pixel by pixel on a small frame of the 16-sphere scene."""
    import oracle
    from blackhole_8_amd import flops

    p = scenes.spheres(48, 27)
    got = flops.spheres_in_plane(p.scene)
    B = np.array(p.scene.hole.position[:])
    cs = [np.array(p.scene.spheres[i].center[:]) - B for i in range(p.scene.n_spheres)]
    rs = [p.scene.spheres[i].radius for i in range(p.scene.n_spheres)]
    want = np.zeros_like(got)
    for y in range(27):
        for x in range(48):
            ray = oracle.ray_setup(p.scene, p.options, x, y)
            n = np.cross(np.array(ray.pv[:]), np.array(ray.F[:]))
            n /= np.linalg.norm(n)
            want[y, x] = sum(abs(n @ c) < r for c, r in zip(cs, rs))
    assert np.array_equal(got, want)
    assert got.max() >= 1 and (got == 0).any()


def test_frame_flops_sphere_and_background_terms():
    from blackhole_8_amd import flops

    steps = np.array([[3, 5], [0, 2]], dtype=np.int32)
    hit = np.array([[0, 3], [1, 0]], dtype=np.uint8)
    in_plane = np.array([[1, 0], [2, 3]], dtype=np.int32)
    base = flops.frame_flops(steps, hit, False)
    assert base == 4 * flops.F_SETUP + 10 * flops.F_STEP_NODISC + 2 * flops.F_HIT
    sph = flops.frame_flops(steps, hit, False, in_plane, n_spheres=5, background=True)
    # steps x spheres in the plane: 3*1 + 5*0 + 0*2 + 2*3 = 9; two escaped rays
    assert sph - base == (4 * (flops.F_SPH_RAY + 5 * flops.F_SPH_PLANE) +
                          9 * flops.F_SPH_STEP + 2 * flops.F_BG)
