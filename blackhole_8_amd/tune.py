"""Autotuner of the persistent batch kernel: times a grid of scheduling knobs
(workgroup size, persistent waves per SIMD, refill interval and threshold) on
the batch itself with bh_batch_autotune and caches the winner per device and
batch shape in a JSON file."""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import math
import os
from pathlib import Path

import numpy as np

from . import abi

CACHE_ENV = "BH8_TUNE_CACHE"
DEFAULT_CACHE = Path.home() / ".cache" / "blackhole_8_amd" / "batch_tuning.json"

# the three modes: grid (block size only), persistent tiles, lane refill
GRID = ([dict(mode=abi.BATCH_GRID, block=b) for b in (64, 128, 256)] +
        [dict(mode=abi.BATCH_TILES, block=b, waves_per_simd=w, chunk=c)
         for b in (64, 256) for w in (2, 4, 8) for c in (1, 4, 16)] +
        [dict(mode=abi.BATCH_LANES, block=b, waves_per_simd=w, refill_steps=k, refill_idle=i)
         for b in (64, 256) for w in (2, 4) for k in (4, 8, 16) for i in (16, 32, 48)])


def batch_signature(items) -> str:
    """Shape of a batch: scene count, log2 pixel histogram, disc fraction."""
    px = np.array([s.camera.width * s.camera.height for s, _ in items], dtype=np.float64)
    hist = np.bincount(np.clip(np.log2(px).astype(int), 0, 40), minlength=41).tolist()
    disc = round(sum(s.has_disc for s, _ in items) / max(1, len(items)), 2)
    return json.dumps({"n": len(items), "hist": hist, "disc": disc}, sort_keys=True)


def _key(renderer, items) -> str:
    info = renderer.device_info()
    raw = f"{info['name']}|{info['cu']}|{batch_signature(items)}"
    return hashlib.sha1(raw.encode()).hexdigest()


def _cache_path(path) -> Path:
    return Path(path or os.environ.get(CACHE_ENV) or DEFAULT_CACHE)


def autotune(renderer, items, want=("rgb", "hit"), grid=None, reps: int = 2, cache=None,
             refresh: bool = False):
    """Best abi.BatchTuning for this batch on this device (from the cache when
    present).  Returns (tuning, record) where record holds the timings."""
    import torch

    path = _cache_path(cache)
    key = _key(renderer, items)
    db = {}
    if path.exists():
        try:
            db = json.loads(path.read_text())
        except (OSError, ValueError):
            db = {}
    if key in db and not refresh:
        r = db[key]
        return abi.BatchTuning(**r["best"]), r
    grid = grid or GRID
    n = len(items)
    scenes = (abi.Scene * n)(*[s for s, _ in items])
    opts = (abi.Options * n)(*[o for _, o in items])
    outs = (abi.Outputs * n)()
    px = [s.camera.width * s.camera.height for s, _ in items]
    off = np.concatenate([[0], np.cumsum(px)]).astype(np.int64)
    dev = torch.device("cuda", 0)
    rgb = torch.empty(int(off[-1]) * 3, dtype=torch.uint8, device=dev) if "rgb" in want else None
    hit = torch.empty(int(off[-1]), dtype=torch.uint8, device=dev) if "hit" in want else None
    torch.cuda.synchronize(dev)
    for i in range(n):
        outs[i].rgb = rgb.data_ptr() + 3 * int(off[i]) if rgb is not None else None
        outs[i].hit = hit.data_ptr() + int(off[i]) if hit is not None else None
    cand = (abi.BatchTuning * len(grid))(*[abi.BatchTuning(**g) for g in grid])
    best = abi.BatchTuning()
    best_ms = C.c_double()
    cand_ms = (C.c_double * len(grid))()
    renderer._check(renderer.lib.bh_batch_autotune(renderer.ctx, n, scenes, opts, outs, cand,
                                                   len(grid), reps, C.byref(best),
                                                   C.byref(best_ms), cand_ms))
    rec = {"best": {f: getattr(best, f) for f, _ in abi.BatchTuning._fields_},
           "best_ms": best_ms.value,
           "grid": [dict(g, ms=cand_ms[i]) for i, g in enumerate(grid)],
           "signature": batch_signature(items)}
    db[key] = rec
    try:
        path.parent.mkdir(parents=True, exist_ok=True)
        path.write_text(json.dumps(db, indent=1))
    except OSError:
        pass
    return best, rec
