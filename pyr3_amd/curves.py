"""Post-tonemap channel curves: validation, evaluation, 5 x 256 LUT bake.

Restates src/channel-curves.ts. A curve is 2..8 control points (x, y) in
[0, 1] with strictly increasing x. Evaluation clamps to the end values
outside [x_first, x_last], interpolates linearly between two points, and
otherwise uses the Catmull-Rom cardinal spline with tension 0.5 on the
segment holding x, duplicating the end points as phantom neighbours
(:71-113). `bake_curves` samples each channel at i / 255, clamps to [0, 1]
and stacks the channels in the order composite, r, g, b, luma (:115-133);
it returns None when every channel is the identity (0,0)-(1,1), which the
tonemap treats as "no curves" (`active_mask` == 0, :59-67).
"""
from __future__ import annotations

from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np

CHANNELS = ("composite", "r", "g", "b", "luma")
IDENTITY_POINTS: List[Tuple[float, float]] = [(0.0, 0.0), (1.0, 1.0)]
MIN_POINTS, MAX_POINTS = 2, 8

Curve = Sequence[Tuple[float, float]]


def validate(points: Curve) -> None:
    """src/channel-curves.ts:30-48."""
    if len(points) < MIN_POINTS:
        raise ValueError(f"channel curve must have at least 2 points, got {len(points)}")
    if len(points) > MAX_POINTS:
        raise ValueError(f"channel curve must have at most 8 points, got {len(points)}")
    for i, (x, y) in enumerate(points):
        if x < 0 or x > 1 or y < 0 or y > 1:
            raise ValueError(f"channel curve point {i} out of range [0,1]: x={x:g} y={y:g}")
        if i > 0 and x <= points[i - 1][0]:
            raise ValueError(f"channel curve x not strictly monotonic at index {i}: "
                             f"{points[i - 1][0]:g} -> {x:g}")


def is_identity(points: Curve) -> bool:
    return (len(points) == 2 and tuple(points[0]) == (0.0, 0.0)
            and tuple(points[1]) == (1.0, 1.0))


def identity_curves() -> Dict[str, List[Tuple[float, float]]]:
    return {ch: list(IDENTITY_POINTS) for ch in CHANNELS}


def active_mask(c: Optional[Dict[str, Curve]]) -> int:
    """Bit i set when channel CHANNELS[i] is not the identity."""
    if not c:
        return 0
    mask = 0
    for i, ch in enumerate(CHANNELS):
        if not is_identity(c[ch]):
            mask |= 1 << i
    return mask


def _spline(u: float, ya: float, yb: float, yc: float, yd: float) -> float:
    # cardinal spline, tension B = 0.5, between yb (u = 0) and yc (u = 1)
    b = 0.5
    v = u * u * u * (-b * ya + (2 - b) * yb + (b - 2) * yc + b * yd)
    v += u * u * (2 * b * ya + (b - 3) * yb + (3 - 2 * b) * yc - b * yd)
    v += u * (-b * ya + b * yc)
    return v + yb


def eval_curve(points: Curve, x: float) -> float:
    """src/channel-curves.ts:80-113."""
    x0, y0 = points[0]
    xl, yl = points[-1]
    if x <= x0:
        return y0
    if x >= xl:
        return yl
    if len(points) < 3:
        x1, y1 = points[1]
        dx = x1 - x0
        if dx < 1e-9:
            return y0
        return y0 + (x - x0) / dx * (y1 - y0)
    i = 0
    while i < len(points) - 1 and points[i + 1][0] < x:
        i += 1
    ya = points[max(0, i - 1)][1]
    yb = points[i][1]
    yc = points[i + 1][1]
    yd = points[min(len(points) - 1, i + 2)][1]
    seg = points[i + 1][0] - points[i][0]
    if seg < 1e-9:
        return yb
    return _spline((x - points[i][0]) / seg, ya, yb, yc, yd)


def bake_one(points: Curve) -> np.ndarray:
    """256 f32 samples at i / 255, clamped to [0, 1]."""
    return np.array([min(1.0, max(0.0, eval_curve(points, i / 255))) for i in range(256)],
                    dtype=np.float32)


def bake_curves(c: Optional[Dict[str, Curve]]) -> Optional[np.ndarray]:
    """(5, 256) f32 LUT (composite, r, g, b, luma), or None when inactive."""
    if active_mask(c) == 0:
        return None
    return np.stack([bake_one(c[ch]) for ch in CHANNELS])
