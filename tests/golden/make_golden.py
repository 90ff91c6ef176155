"""Writes the golden fixtures under tests/golden/ from the FP64 oracle.

  python tests/golden/make_golden.py

renders.npz  small frames of the workloads (FP64 reference arithmetic):
             plain_hole 64x36, disc_orbit 96x54 frame 0, the shipped scene 96x54;
             per frame rgb, hit, steps, phi.
solveg.npz   SolveG's 20-bisection turning point for a grid of (b, M).
The oracle is itself pinned by closed forms (tests/test_oracle.py); these
fixtures freeze its output so that any later change to it, or a GPU path
compared against them, is caught without rerunning it.
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

from blackhole_8_amd import abi, scenes  # noqa: E402
import oracle  # noqa: E402

FRAMES = {
    "plain_hole": lambda: scenes.plain_hole(64, 36),
    "disc_orbit": lambda: scenes.disc_orbit(96, 54, frames=1, precision=abi.PREC_FP64),
    "shipped": lambda: scenes.shipped(96, 54),
}


def main():
    out = {}
    for name, make in FRAMES.items():
        r = oracle.render_preset(make(), 4, ("rgb", "hit", "steps", "phi"))
        for k, v in r.items():
            out[f"{name}_{k}"] = v
    np.savez_compressed(HERE / "renders.npz", **out)
    lib = oracle.load()
    Ms = np.array([0.25, 0.5, 2.0])
    bs = np.concatenate([np.array([2.5981, 2.6, 2.7, 3.0, 5.0, 10.0, 100.0]), np.logspace(0.5, 3, 20)])
    table = np.array([[lib.or_solve_g(b * 2 * M, M, abi.ROOT_BISECT20) for b in bs] for M in Ms])
    np.savez_compressed(HERE / "solveg.npz", M=Ms, b_over_rs=bs, l=table)
    print("wrote", HERE / "renders.npz", HERE / "solveg.npz")


if __name__ == "__main__":
    main()
