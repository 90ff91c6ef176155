"""GPU batches: many scenes in one persistent launch (bh_render_batch_device).

Invariant: every scene of a batch equals its solo render (bh_render) byte for
byte, whatever the scheduling knobs; the solo FP32 render is itself held to
the FP64 oracle in test_gpu_parity.py, and a few batch scenes are compared
with the oracle directly here (FP32 bar: classes identical except <= 1e-4 of
pixels, colour within 1/255 except texel straddles).
"""
import os

import numpy as np
import pytest

from blackhole_8_amd import abi, scenes, tune
import oracle
from tests.parity import compare, summary

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


@pytest.fixture(scope="module")
def c3():
    return scenes.uneven_batch()


def _small(items, limit=160, n=48):
    return [it for it in items if it[0].camera.width <= limit][:n]


def _assert_same_as_solo(renderer, items, res):
    for i, ((s, o), r) in enumerate(zip(items, res)):
        solo = renderer.render(s, o, ("rgb", "hit"))
        assert np.array_equal(r["hit"], solo["hit"]), f"scene {i} hit"
        assert np.array_equal(r["rgb"], solo["rgb"]), f"scene {i} rgb"


def test_batch_equals_solo_renders(renderer, c3):
    items, tex = c3
    renderer.set_textures(tex)
    sub = _small(items)
    print(f"[case] batch of {len(sub)} small uneven_batch scenes", flush=True)
    res, ms, st = renderer.render_batch(sub, timing=True)
    assert st.rays == sum(s.camera.width * s.camera.height for s, _ in sub)
    assert st.solo_scenes == 0
    assert (ms > 0).all() and (ms <= st.total_ms + 1e-3).all()
    _assert_same_as_solo(renderer, sub, res)
    assert sum(int((r["hit"] == abi.HIT_DISC).sum()) for r in res) > 1000


@pytest.mark.parametrize("tuning", [
    dict(mode=abi.BATCH_GRID, block=64),
    dict(mode=abi.BATCH_TILES, block=256, waves_per_simd=8, chunk=1),
    dict(mode=abi.BATCH_TILES, block=64, waves_per_simd=2, chunk=16),
    dict(mode=abi.BATCH_LANES, block=64, waves_per_simd=1, refill_steps=1, refill_idle=1),
    dict(mode=abi.BATCH_LANES, block=64, waves_per_simd=2, refill_steps=16, refill_idle=48),
    dict(mode=abi.BATCH_LANES, block=128, waves_per_simd=4, refill_steps=3, refill_idle=32)])
def test_scheduling_knobs_do_not_change_pixels(renderer, c3, tuning):
    items, tex = c3
    renderer.set_textures(tex)
    sub = _small(items, 128, 24)
    print(f"[case] batch tuning {tuning}", flush=True)
    ref = renderer.render_batch(sub)
    res = renderer.render_batch(sub, tuning=abi.BatchTuning(**tuning))
    for a, b in zip(ref, res):
        assert np.array_equal(a["rgb"], b["rgb"]) and np.array_equal(a["hit"], b["hit"])


def test_full_uneven_batch_batch(renderer, c3):
    items, tex = c3
    renderer.set_textures(tex)
    print(f"[case] full uneven_batch batch: {len(items)} scenes", flush=True)
    res, ms, st = renderer.render_batch(items, timing=True)
    rays = sum(s.camera.width * s.camera.height for s, _ in items)
    print(f"[batch] {rays} rays in {st.total_ms:.3f} ms ({rays / st.total_ms / 1e3:.0f} Mrays/s), "
          f"march {st.march_ms:.3f} ms, band {st.band_rays} rays, "
          f"scene latency median {np.median(ms):.3f} ms max {ms.max():.3f} ms", flush=True)
    assert st.rays == rays
    rng = np.random.default_rng(11)
    pick = sorted(rng.choice(len(items), 12, replace=False).tolist())
    _assert_same_as_solo(renderer, [items[i] for i in pick], [res[i] for i in pick])


def test_batch_scenes_match_oracle(renderer, c3):
    items, tex = c3
    renderer.set_textures(tex)
    sub = [it for it in items if it[0].has_disc and it[0].camera.width <= 200][:4]
    print(f"[case] batch vs oracle: {len(sub)} disc scenes", flush=True)
    res = renderer.render_batch(sub)
    for (s, o), r in zip(sub, res):
        ref = oracle.render(s, o, tex, THREADS, ("rgb", "hit"))
        st = compare(r, ref)
        print(f"[batch-oracle] {s.camera.width}px: {summary(st)}", flush=True)
        assert st["class_mismatch_frac"] <= 1e-4
        assert st["rgb_mismatch_frac"] <= 2e-3


def test_batch_with_fp64_and_sphere_scenes(renderer, c3):
    items, tex = c3
    p2 = scenes.spheres(96, 54)
    renderer.set_textures(tex)
    renderer.set_textures(p2.textures)
    s64, o64 = items[0][0], scenes.options(precision=abi.PREC_FP64)
    mixed = [items[1], (p2.scene, p2.options), (s64, o64), items[2]]
    print("[case] mixed batch: FP32, spheres, FP64", flush=True)
    res, ms, st = renderer.render_batch(mixed, timing=True)
    assert st.solo_scenes == 1  # the FP64 scene; the sphere scene is in the FP32 march
    _assert_same_as_solo(renderer, mixed, res)
    # the lane-refill mode keeps no sphere table: there the sphere scene goes solo
    res, ms, st = renderer.render_batch(mixed, timing=True,
                                        tuning=abi.BatchTuning(mode=abi.BATCH_LANES))
    assert st.solo_scenes == 2
    _assert_same_as_solo(renderer, mixed, res)


def test_autotune_picks_and_caches(renderer, c3, tmp_path):
    items, tex = c3
    renderer.set_textures(tex)
    sub = _small(items, 256, 64)
    grid = [dict(mode=abi.BATCH_GRID, block=256),
            dict(mode=abi.BATCH_TILES, block=256, waves_per_simd=4, chunk=4),
            dict(mode=abi.BATCH_LANES, block=256, waves_per_simd=4, refill_steps=8, refill_idle=32)]
    cache = tmp_path / "tune.json"
    print("[case] autotune 3 candidates", flush=True)
    best, rec = tune.autotune(renderer, sub, grid=grid, reps=1, cache=cache)
    assert cache.exists()
    assert rec["best_ms"] == min(g["ms"] for g in rec["grid"])
    best2, rec2 = tune.autotune(renderer, sub, grid=grid, reps=1, cache=cache)
    assert rec2 == rec
    assert (best2.mode, best2.block) == (best.mode, best.block)
