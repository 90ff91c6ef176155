"""Comparison helpers shared by the parity tests (GPU output vs oracle)."""
from __future__ import annotations

import numpy as np


def compare(gpu: dict, ref: dict) -> dict:
    """Mismatch statistics between two renders of the same scene."""
    out = {}
    n = ref["hit"].size
    cls_bad = gpu["hit"] != ref["hit"]
    out["pixels"] = n
    out["class_mismatch"] = int(cls_bad.sum())
    out["class_mismatch_frac"] = float(cls_bad.sum()) / n
    same = ~cls_bad
    d = np.abs(gpu["rgb"].astype(np.int16) - ref["rgb"].astype(np.int16)).max(axis=-1)
    rgb_bad = (d > 1) & same
    out["rgb_mismatch"] = int(rgb_bad.sum())
    out["rgb_mismatch_frac"] = float(rgb_bad.sum()) / n
    out["rgb_bad_mask"] = rgb_bad
    if "steps" in gpu and "steps" in ref:
        st = same & (gpu["steps"] == ref["steps"])
        out["steps_mismatch"] = int((same & ~st).sum())
    else:
        st = same
    if "phi" in gpu and "phi" in ref:
        ok = st & np.isfinite(ref["phi"]) & np.isfinite(gpu["phi"])
        err = np.abs(gpu["phi"] - ref["phi"])[ok]
        out["phi_max_err"] = float(err.max()) if err.size else 0.0
        out["phi_p999_err"] = float(np.quantile(err, 0.999)) if err.size else 0.0
        out["phi_median_err"] = float(np.median(err)) if err.size else 0.0
    if "hitpt" in gpu and "hitpt" in ref:
        ok = same & np.isfinite(ref["hitpt"][..., 0]) & np.isfinite(gpu["hitpt"][..., 0])
        e = np.linalg.norm(gpu["hitpt"] - ref["hitpt"], axis=-1)[ok]
        out["hit_point_max_err"] = float(e.max()) if e.size else 0.0
        out["hit_point_p999_err"] = float(np.quantile(e, 0.999)) if e.size else 0.0
    return out


def summary(stats: dict) -> str:
    return ", ".join(f"{k}={v:.3g}" if isinstance(v, float) else f"{k}={v}"
                     for k, v in stats.items() if not isinstance(v, np.ndarray))


def disc_texel_coords(hitpt: np.ndarray, vertex: np.ndarray, rows: int, cols: int):
    """Continuous texel coordinates of world points on the disc
    (Rectangle::color, vector_object.h:163-179, in closed form)."""
    v1 = vertex[1]
    s1 = vertex[2] - v1
    s2 = vertex[4] - v1
    v = hitpt - v1
    l1 = np.linalg.norm(s1)
    proj = (v @ s1) / l1
    perp = np.linalg.norm(np.cross(v, s1), axis=-1) / l1
    return proj / l1 * cols, perp / np.linalg.norm(s2) * rows


def straddles(a: np.ndarray, b: np.ndarray, tol: float = 1e-3,
              max_span: float = 0.01) -> np.ndarray:
    """True where two continuous texel coordinates at most max_span texels
    apart lie on either side of a texel edge (the span widened by tol): a
    nearest-texel difference between two nearly equal points.  Points further
    apart are never a straddle."""
    lo = np.minimum(a, b) - tol
    hi = np.maximum(a, b) + tol
    return (np.floor(hi) > np.floor(lo)) & (np.abs(a - b) <= max_span)


def oracle_disc_texel(hitpt: np.ndarray, vertex, rows: int, cols: int):
    """Continuous texel coordinates by the oracle's own restatement of
    Rectangle::color (acos round trip included)."""
    import ctypes as C

    import oracle

    lib = oracle.load()
    v = np.ascontiguousarray(np.asarray(vertex, dtype=np.float64).reshape(5, 3))
    a = np.empty(len(hitpt))
    b = np.empty(len(hitpt))
    fa, fb = C.c_double(), C.c_double()
    ia, ib = C.c_int(), C.c_int()
    for i, q in enumerate(np.ascontiguousarray(hitpt, dtype=np.float64)):
        q = np.ascontiguousarray(q)
        lib.or_rect_texel(v.ctypes.data_as(C.POINTER(C.c_double)), rows, cols,
                          q.ctypes.data_as(C.POINTER(C.c_double)), C.byref(fa), C.byref(fb),
                          C.byref(ia), C.byref(ib))
        a[i], b[i] = fa.value, fb.value
    return a, b
