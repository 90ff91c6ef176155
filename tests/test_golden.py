"""The committed golden fixtures (tests/golden/make_golden.py) against the
oracle (CPU) and against the HIP march (GPU)."""
from pathlib import Path

import numpy as np
import pytest

from blackhole_8_amd import abi
import oracle
from tests.golden.make_golden import FRAMES

GOLD = Path(__file__).resolve().parent / "golden"


def fixtures():
    return dict(np.load(GOLD / "renders.npz"))


@pytest.mark.parametrize("name", sorted(FRAMES))
def test_oracle_reproduces_golden_frames(name):
    g = fixtures()
    r = oracle.render_preset(FRAMES[name](), 2, ("rgb", "hit", "steps", "phi"))
    for k in ("rgb", "hit", "steps"):
        assert np.array_equal(r[k], g[f"{name}_{k}"]), k
    np.testing.assert_array_equal(np.isnan(r["phi"]), np.isnan(g[f"{name}_phi"]))
    ok = ~np.isnan(r["phi"])
    assert np.array_equal(r["phi"][ok], g[f"{name}_phi"][ok])


def test_oracle_reproduces_golden_turning_points():
    t = np.load(GOLD / "solveg.npz")
    lib = oracle.load()
    for i, M in enumerate(t["M"]):
        for j, b in enumerate(t["b_over_rs"]):
            assert lib.or_solve_g(b * 2 * M, M, abi.ROOT_BISECT20) == t["l"][i, j]


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(FRAMES))
def test_gpu_matches_golden_frames(renderer, name):
    g = fixtures()
    p = FRAMES[name]()
    for prec in (abi.PREC_FP64, abi.PREC_FP32):
        p.options.precision = prec
        print(f"[case] golden {name} precision {prec}", flush=True)
        r = renderer.render_preset(p, ("rgb", "hit", "steps", "phi"))
        n = r["hit"].size
        assert (r["hit"] != g[f"{name}_hit"]).sum() <= max(1, 1e-4 * n)
        same = (r["hit"] == g[f"{name}_hit"]) & (r["steps"] == g[f"{name}_steps"])
        ok = same & ~np.isnan(g[f"{name}_phi"])
        tol = 1e-9 if prec == abi.PREC_FP64 else 1e-5
        assert np.abs(r["phi"][ok] - g[f"{name}_phi"][ok]).max() <= tol
