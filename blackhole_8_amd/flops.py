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

Work of a frame = F_SETUP * rays + F_STEP * sum(steps) + F_HIT * hits, with
the per-pixel step counts and hit classes of the render itself.
"""
from __future__ import annotations

import numpy as np

F_STEP_DISC = 35
F_STEP_NODISC = 30
F_SETUP = 240
F_HIT = 40


def frame_flops(steps: np.ndarray, hit: np.ndarray, has_disc: bool) -> float:
    f_step = F_STEP_DISC if has_disc else F_STEP_NODISC
    n = steps.size
    return float(F_SETUP * n + f_step * int(steps.astype(np.int64).sum()) +
                 F_HIT * int((hit != 0).sum()))
