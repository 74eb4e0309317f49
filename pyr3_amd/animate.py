"""Keyframe animation: interpolation, motion overlays, easing, temporal filter,
and the motion-blurred frame render.

The step before the chaos path for animations: a genome per sub-frame time,
then the sub-frames' dispatches into one histogram and one present.

* `interpolate(animation, t)` restates src/interpolate.ts:58-257 (the
  reference's port of flam3's flam3_interpolate): keyframe pick with
  endpoint extrapolation (:264-300), symmetry baked into xforms first
  (src/symmetry.ts:64-83), per-segment easing of the blend weight (:78-84),
  Catmull-Rom smooth interpolation on interior segments (:90-98, :870-1005),
  xform-count alignment with zero-weight identity pads, 180-degree flipped
  for the spherical family under log interpolation (:311-356, :1008-1033),
  per-segment xform permutation (:115-118, :1252-1279), asymmetric rotation
  wind (:358-387), motion overlays on the raw segment clock (:126-136),
  per-xform blend with log-polar or linear affines (:389-576), variation
  union (:1109-1139), xaos (:1141-1152), palettes in HSV / RGB /
  hsv_circular / sweep (:611-733, :1154-1211), tonemap (:735-756), HSL
  adjust and channel curves fading from identity (:758-868), and the
  continuous render fields (:201-241).
* `apply_motion_parameters` restates src/motion.ts:41-168 (flam3's
  motion_funcs and apply_motion_parameters).
* `eval_easing` restates src/easing.ts:15-93.
* `temporal_filter` restates src/temporal-filter.ts:25-77 (flam3's
  flam3_create_temporal_filter).
* `render_animation_frame` follows src/animate-render.ts:73-189: the
  centre-time genome's quality sizes the dispatch, walkers are apportioned
  over the filter weights (largest remainder), sub-frame i runs at
  t + delta[i] with seed (seed + i * 0x9e3779b1) mod 2^32, and one present
  of the centre genome calibrates on the total samples. Consecutive
  sub-frames with the same kernel shape go out as one batched launch (up
  to 64), whose histogram equals one dispatch per sub-frame.

All arithmetic is float64, as in the reference (JavaScript numbers).
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

from . import curves as cv
from . import pack
from .catalog import VARIATION_INDEX
from .genome import (MAX_VARIATION_PARAMS, MAX_XFORMS, Affine, ColorStop, Density, Genome,
                     Palette, Tonemap, Variation, Xform)

EPS = 1e-9                  # src/interpolate.ts:45
LOG_MAG_THRESHOLD = -10.0   # src/interpolate.ts:50
TWO_PI = 2.0 * math.pi
MAX_VARIATIONS_PER_XFORM = 8
PALETTE_SIZE = 256
LINEAR = VARIATION_INDEX["linear"]
# spherical, polar, julian, juliascope, ngon, wedge_sph, wedge_julia
# (src/interpolate.ts:330): their padded identity is rotated 180 degrees under
# log interpolation
FLIPPED_IDENTITY_VARS = frozenset(VARIATION_INDEX[n] for n in (
    "spherical", "polar", "julian", "juliascope", "ngon", "wedge_sph", "wedge_julia"))
# the missing-keyframe side of a tonemap blend (src/interpolate.ts:735-738)
INTERP_TONEMAP_FALLBACK = Tonemap(gamma=4.0, vibrancy=1.0, highlight_power=-1.0,
                                  brightness=4.0, gamma_threshold=0.01)
IDENTITY_HSL = {"hue": 0.0, "sat": 100.0, "light": 0.0}
SUBFRAME_SEED_STEP = 0x9E3779B1


@dataclass
class Animation:
    """src/animation.ts:35-90; defaults are flam3's (FLAM3_ANIMATION_DEFAULTS)."""
    keyframes: List[Genome]
    interpolation: str = "linear"             # 'linear' | 'smooth'
    interpolation_type: str = "log"           # 'linear' | 'log' | 'compat' | 'older'
    palette_interpolation: str = "hsv_circular"   # 'hsv' | 'sweep' | 'rgb' | 'hsv_circular'
    hsv_rgb_palette_blend: float = 0.0
    ntemporal_samples: float = 1000.0
    temporal_filter_type: str = "box"         # 'box' | 'gaussian' | 'exp'
    temporal_filter_width: float = 1.0
    temporal_filter_exp: float = 0.0
    # per keyframe gap: an easing curve ({'kind': 'preset', 'name': ...} or
    # {'kind': 'cubicBezier', 'x1', 'y1', 'x2', 'y2'}) and an xform permutation
    segment_easing: Optional[List[Optional[Dict]]] = None
    segment_permutation: Optional[List[Optional[List[int]]]] = None

    @classmethod
    def from_flame(cls, imp) -> "Animation":
        """The animation of a multi-flame `.flam3` import (pyr3_amd.flam3):
        keyframes sorted by time, settings from the first flame
        (src/flame-import.ts:1091-1144)."""
        if len(imp.keyframes) < 2:
            raise ValueError("pyr3: an animation needs at least 2 keyframes")
        a = cls(keyframes=list(imp.keyframes))
        s = imp.animation
        for key in ("interpolation", "interpolation_type", "palette_interpolation",
                    "temporal_filter_type", "hsv_rgb_palette_blend", "temporal_filter_width",
                    "temporal_filter_exp"):
            if key in s:
                setattr(a, key, s[key])
        if "temporal_samples" in s:
            a.ntemporal_samples = s["temporal_samples"]
        return a


# ------------------------------------------------------------ numerics --

def _blend(v0: float, v1: float, c0: float, c1: float) -> float:
    return c0 * v0 + c1 * v1


def _blend_n(values: Sequence[float], weights: Sequence[float]) -> float:
    s = 0.0
    for v, w in zip(values, weights):
        s += w * v
    return s


def _clamp_ge0(x: float) -> float:
    return 0.0 if x < 0 else x


def _clamp01(x: float) -> float:
    return 0.0 if x < 0 else (1.0 if x > 1 else x)


def _js_round(x: float) -> int:
    return int(math.floor(x + 0.5))


# ------------------------------------------------------- temporal filter --

@dataclass
class TemporalFilter:
    deltas: List[float]
    filter: List[float]
    sumfilt: float


def temporal_filter(numsteps: int, kind: str = "box", width: float = 1.0,
                    exp: float = 0.0) -> TemporalFilter:
    """Sub-frame time offsets over [-width/2, width/2] and weights normalised
    to max 1; sumfilt = mean weight (src/temporal-filter.ts:25-77)."""
    if numsteps <= 1:
        return TemporalFilter([0.0], [1.0], 1.0)
    deltas = [(i / (numsteps - 1) - 0.5) * width for i in range(numsteps)]
    if kind == "exp":
        filt = [((i + 1) / numsteps if exp >= 0 else (numsteps - i) / numsteps) ** abs(exp)
                for i in range(numsteps)]
    elif kind == "gaussian":
        half = numsteps / 2.0
        filt = []
        for i in range(numsteps):
            x = 1.5 * abs(i - half) / half      # flam3's gaussian support 1.5
            filt.append(math.exp(-2.0 * x * x) * math.sqrt(2.0 / math.pi))
    else:
        filt = [1.0] * numsteps
    mx = 1.0 if kind not in ("exp", "gaussian") else max(0.0, *filt)
    filt = [f / mx for f in filt]
    total = 0.0
    for f in filt:
        total += f
    return TemporalFilter(deltas, filt, total / numsteps)


# --------------------------------------------------------------- easing --

def _preset(name: str, t: float) -> float:
    if name == "easeIn":
        return t * t
    if name == "easeOut":
        return 1 - (1 - t) * (1 - t)
    if name == "easeInOut":
        return t * t * (3 - 2 * t)
    if name == "hold":
        return 0.0 if t < 1 else 1.0
    return t


def _bezier(c1: float, c2: float, s: float) -> float:
    mt = 1 - s
    return 3 * mt * mt * s * c1 + 3 * mt * s * s * c2 + s * s * s


def _bezier_d(c1: float, c2: float, s: float) -> float:
    mt = 1 - s
    return 3 * mt * mt * c1 + 6 * mt * s * (c2 - c1) + 3 * s * s * (1 - c2)


def _solve_bezier(x1: float, x2: float, x: float) -> float:
    """Newton steps (clamped to [0, 1]), then bisection (src/easing.ts:49-73)."""
    eps = 1e-6
    s = x
    for _ in range(8):
        err = _bezier(x1, x2, s) - x
        if abs(err) < eps:
            return s
        d = _bezier_d(x1, x2, s)
        if abs(d) < eps:
            break
        s = _clamp01(s - err / d)
    lo, hi = 0.0, 1.0
    s = x
    while hi - lo > eps:
        xs = _bezier(x1, x2, s)
        if abs(xs - x) < eps:
            break
        if x > xs:
            lo = s
        else:
            hi = s
        s = (lo + hi) / 2
    return s


def eval_easing(curve: Optional[Dict], t: float) -> float:
    """Remap t in [0, 1] through a preset or cubic-bezier curve; output
    clamped to [0, 1]; unknown kinds are the identity."""
    x = _clamp01(t)
    if not curve:
        return x
    if curve.get("kind") == "preset":
        return _clamp01(_preset(curve.get("name"), x))
    if curve.get("kind") == "cubicBezier":
        s = _solve_bezier(curve["x1"], curve["x2"], x)
        return _clamp01(_bezier(curve["y1"], curve["y2"], s))
    return x


# --------------------------------------------------------------- motion --

def motion_funcs(func: int, t: float) -> float:
    """1 sin(2 pi t); 2 triangle (+1 at 1/4, -1 at 3/4); 3 hill
    (1 - cos(2 pi t)) / 2 (src/motion.ts:41-59)."""
    if func == 1:
        return math.sin(TWO_PI * t)
    if func == 2:
        fr = math.fmod(t, 1.0)
        if fr < 0:
            fr += 1
        if fr <= 0.25:
            return 4.0 * fr
        if fr <= 0.75:
            return -4.0 * fr + 2.0
        return 4.0 * fr - 4.0
    return (1 - math.cos(TWO_PI * t)) * 0.5


def apply_motion_parameters(xf: Xform, blend: float) -> Xform:
    """The xform with each motion element's fields added, scaled by
    motion_funcs(func, freq * blend); colour clamped to [0, 1] and weight to
    >= 0 afterwards. A motion variation absent from the xform is added at
    weight 0 unless the xform already has 8 (then that contribution is
    dropped). Returns `xf` itself when it has no motion."""
    if not xf.motion:
        return xf
    out = copy.deepcopy(xf)
    for m in xf.motion:
        func = m.motion_func
        if not func:
            continue
        w = motion_funcs(func, (m.motion_freq or 0.0) * blend)
        a, ma = out.affine, m.affine
        a.a += ma.a * w; a.b += ma.b * w; a.c += ma.c * w
        a.d += ma.d * w; a.e += ma.e * w; a.f += ma.f * w
        out.weight += m.weight * w
        out.color += m.color * w
        out.color_speed += m.color_speed * w
        if m.opacity_given:
            out.opacity = out.opacity + m.opacity * w
        if m.animate is not None:
            out.animate = (out.animate if out.animate is not None else 1.0) + m.animate * w
        if m.post is not None:
            if out.post is None:
                out.post = Affine()
            p, mp = out.post, m.post
            p.a += mp.a * w; p.b += mp.b * w; p.c += mp.c * w
            p.d += mp.d * w; p.e += mp.e * w; p.f += mp.f * w
        if m.variations:
            by_idx = {v.index: v for v in out.variations}
            for mv in m.variations:
                tgt = by_idx.get(mv.index)
                if tgt is None:
                    if len(out.variations) >= MAX_VARIATIONS_PER_XFORM:
                        continue
                    tgt = Variation(mv.index, 0.0)
                    out.variations.append(tgt)
                    by_idx[mv.index] = tgt
                tgt.weight += mv.weight * w
                for k in range(len(tgt.params)):
                    tgt.params[k] += mv.params[k] * w
    out.color = min(1.0, max(0.0, out.color))
    if out.weight < 0:
        out.weight = 0.0
    return out


# ----------------------------------------------------------- keyframes --

def pick_keyframes(keyframes: Sequence[Genome], time: float) -> Tuple[int, int, float, float]:
    """(i1, i2, c0, c1) for `time`; outside the keyframes' span the end
    segment extrapolates (src/interpolate.ts:264-300)."""
    n = len(keyframes)
    if n < 2:
        raise ValueError("pyr3: pickKeyframes requires N >= 2")
    tk = [k.time or 0.0 for k in keyframes]
    if tk[0] >= time:
        return _compute_blend(tk, 0, 1, time)
    if tk[n - 1] <= time:
        return _compute_blend(tk, n - 2, n - 1, time)
    i1 = 0
    while tk[i1] < time:
        i1 += 1
    i1 -= 1
    return _compute_blend(tk, i1, i1 + 1, time)


def _compute_blend(tk, i1, i2, time):
    t1, t2 = tk[i1], tk[i2]
    if abs(t2 - t1) < EPS:
        return i1, i2, 0.0, 1.0
    c0 = (t2 - time) / (t2 - t1)
    return i1, i2, c0, 1 - c0


def catmull_rom_weights(t: float) -> Tuple[float, float, float, float]:
    """Cubic basis of keyframes [i1-1, i1, i1+1, i1+2] (src/interpolate.ts:870-880)."""
    t2 = t * t
    t3 = t2 * t
    return ((2 * t2 - t - t3) / 2, (3 * t3 - 5 * t2 + 2) / 2,
            (4 * t2 - 3 * t3 + t) / 2, (t3 - t2) / 2)


def bake_symmetry(g: Genome) -> Genome:
    """Symmetry declaration -> appended rotation/reflection xforms, field
    cleared (src/symmetry.ts:64-83)."""
    if not g.symmetry:
        return g
    extra = pack.symmetry_xforms(g.symmetry)
    if not extra:
        return g
    total = len(g.xforms) + len(extra)
    if total > MAX_XFORMS:
        raise ValueError(f"pyr3: symmetry expansion exceeds MAX_XFORMS ({MAX_XFORMS}): "
                         f"{len(g.xforms)} source + {len(extra)} generated = {total}. "
                         "Reduce source xforms or symmetry n.")
    out = copy.copy(g)
    out.xforms = list(g.xforms) + extra
    out.symmetry = None
    return out


def _identity_xform(hint: Optional[int] = None, use_log: bool = False) -> Xform:
    flip = use_log and hint is not None and hint in FLIPPED_IDENTITY_VARS
    s = -1.0 if flip else 1.0
    return Xform(weight=0.0, color=0.0, color_speed=0.5, affine=Affine(s, 0.0, 0.0, 0.0, s, 0.0),
                 variations=[Variation(LINEAR, s)])


def align_xform_counts(kfs: Sequence[Genome], interpolation_type: str) -> List[Genome]:
    """Pad every keyframe to the largest xform count, and to a final xform if
    any has one (src/interpolate.ts:1008-1033)."""
    n = max(len(k.xforms) for k in kfs)
    any_final = any(k.finalxform is not None for k in kfs)
    use_log = interpolation_type == "log"

    def hint_for(j):
        for k in kfs:
            if j < len(k.xforms):
                for v in k.xforms[j].variations:
                    if v.index in FLIPPED_IDENTITY_VARS and v.active:
                        return v.index
        return None

    out = []
    for g in kfs:
        a = copy.copy(g)
        if len(g.xforms) < n:
            a.xforms = list(g.xforms) + [_identity_xform(hint_for(len(g.xforms) + k), use_log)
                                         for k in range(n - len(g.xforms))]
        if g.finalxform is None and any_final:
            a.finalxform = _identity_xform()
        out.append(a)
    return out


def _is_permutation(perm, n) -> bool:
    if perm is None or len(perm) != n:
        return False
    seen = [False] * n
    for v in perm:
        if not isinstance(v, int) or isinstance(v, bool) or v < 0 or v >= n or seen[v]:
            return False
        seen[v] = True
    return True


def resolve_segment_permutation(perm, n: int) -> Optional[List[int]]:
    """A bijection over [0, n), a shorter valid one extended with an
    identity tail, else None (src/interpolate.ts:1252-1279)."""
    if not perm:
        return None
    if len(perm) == n:
        return list(perm) if _is_permutation(perm, n) else None
    if len(perm) < n and _is_permutation(perm, len(perm)):
        return list(perm) + list(range(len(perm), n))
    return None


def establish_wind(g0: Genome, g1: Genome) -> List[Optional[Tuple[float, float]]]:
    """Per xform, a reference angle per affine column when exactly one side
    is stationary (animate == 0), pinning the log-polar winding
    (src/interpolate.ts:358-387)."""
    out = []
    for x0, x1 in zip(g0.xforms, g1.xforms):
        s0, s1 = x0.animate == 0, x1.animate == 0
        if s0 == s1:
            out.append(None)
            continue
        w = []
        for col in (0, 1):
            p, q = x0.affine, x1.affine
            a0 = math.atan2(p.d, p.a) if col == 0 else math.atan2(p.e, p.b)
            a1 = math.atan2(q.d, q.a) if col == 0 else math.atan2(q.e, q.b)
            d = a1 - a0
            if d > math.pi + EPS:
                a1 -= TWO_PI
            elif d < -(math.pi - EPS):
                a1 += TWO_PI
            w.append((a0 if s1 else a1) + TWO_PI)
        out.append((w[0], w[1]))
    return out


# -------------------------------------------------------------- affines --

def _is_identity_affine(p: Affine) -> bool:
    return p.a == 1 and p.b == 0 and p.c == 0 and p.d == 0 and p.e == 1 and p.f == 0


def _affine_linear_n(affs: Sequence[Affine], cmc) -> Affine:
    return Affine(*(_blend_n([getattr(x, k) for x in affs], cmc) for k in "abcdef"))


def _polar_columns(p: Affine):
    """(ang0, mag0, ang1, mag1) of the columns (a, d) and (b, e); a zero
    column takes the other column's angle (src/interpolate.ts:521-531)."""
    mag0 = math.hypot(p.a, p.d)
    mag1 = math.hypot(p.b, p.e)
    ang0 = math.atan2(p.d, p.a)
    ang1 = math.atan2(p.e, p.b)
    z0, z1 = mag0 < EPS, mag1 < EPS
    if z0 and not z1:
        ang0 = ang1
    elif z1 and not z0:
        ang1 = ang0
    return ang0, mag0, ang1, mag1


def _blend_polar_column(ang0, mag0, ang1, mag1, c0, c1, wind=None):
    """src/interpolate.ts:533-576."""
    if wind is not None and wind > 0:
        ref = wind - TWO_PI
        while ang0 < ref:
            ang0 += TWO_PI
        while ang0 > ref + TWO_PI:
            ang0 -= TWO_PI
        while ang1 < ref:
            ang1 += TWO_PI
        while ang1 > ref + TWO_PI:
            ang1 -= TWO_PI
    else:
        d = ang1 - ang0
        if d > math.pi + EPS:
            ang1 -= TWO_PI
        elif d < -(math.pi - EPS):
            ang1 += TWO_PI
    log_mag = (math.log(mag0 + EPS) >= LOG_MAG_THRESHOLD
               and math.log(mag1 + EPS) >= LOG_MAG_THRESHOLD)
    ang = c0 * ang0 + c1 * ang1
    mag = (math.exp(c0 * math.log(mag0 + EPS) + c1 * math.log(mag1 + EPS)) if log_mag
           else c0 * mag0 + c1 * mag1)
    return mag * math.cos(ang), mag * math.sin(ang)


def _affine_log_polar(x0: Affine, x1: Affine, c0, c1, wind=None) -> Affine:
    p0, p1 = _polar_columns(x0), _polar_columns(x1)
    ax, ay = _blend_polar_column(p0[0], p0[1], p1[0], p1[1], c0, c1,
                                 None if wind is None else wind[0])
    bx, by = _blend_polar_column(p0[2], p0[3], p1[2], p1[3], c0, c1,
                                 None if wind is None else wind[1])
    return Affine(ax, bx, _blend(x0.c, x1.c, c0, c1), ay, by, _blend(x0.f, x1.f, c0, c1))


def _blend_polar_column_n(angs, mags, cmc):
    """Sequential +-pi unwrap, then the weighted blend (src/interpolate.ts:1093-1107)."""
    angs = list(angs)
    for k in range(1, len(angs)):
        d = angs[k] - angs[k - 1]
        if d > math.pi + EPS:
            angs[k] -= TWO_PI
        elif d < -(math.pi - EPS):
            angs[k] += TWO_PI
    log_mag = all(math.log(m + EPS) >= LOG_MAG_THRESHOLD for m in mags)
    ang = _blend_n(angs, cmc)
    mag = (math.exp(_blend_n([math.log(m + EPS) for m in mags], cmc)) if log_mag
           else _blend_n(mags, cmc))
    return mag * math.cos(ang), mag * math.sin(ang)


def _affine_log_polar_n(affs: Sequence[Affine], cmc) -> Affine:
    cols = [_polar_columns(a) for a in affs]
    ax, ay = _blend_polar_column_n([c[0] for c in cols], [c[1] for c in cols], cmc)
    bx, by = _blend_polar_column_n([c[2] for c in cols], [c[3] for c in cols], cmc)
    return Affine(ax, bx, _blend_n([a.c for a in affs], cmc), ay, by,
                  _blend_n([a.f for a in affs], cmc))


# --------------------------------------------------------------- xforms --

def interpolate_variations(vss: Sequence[Sequence[Variation]], cmc) -> List[Variation]:
    """Union of variation indices in first-seen order; weights (0 where
    absent or inactive) and parameters (0 where absent) blended; at most 8;
    an empty union gives linear 1 (src/interpolate.ts:1109-1139)."""
    maps = [{v.index: v for v in vs} for vs in vss]
    order: List[int] = []
    for vs in vss:
        for v in vs:
            if v.index not in order:
                order.append(v.index)
    out: List[Variation] = []
    for idx in order:
        got = [m.get(idx) for m in maps]
        w = _blend_n([0.0 if v is None or not v.active else v.weight for v in got], cmc)
        params = [_blend_n([0.0 if v is None else v.params[k] for v in got], cmc)
                  for k in range(MAX_VARIATION_PARAMS)]
        nv = Variation(idx, w, params)
        if all(v is not None and not v.active for v in got):
            nv.active = False
        out.append(nv)
    del out[MAX_VARIATIONS_PER_XFORM:]
    if not out:
        out.append(Variation(LINEAR, 1.0))
    return out


def interpolate_xaos(xss: Sequence[Optional[List[float]]], cmc) -> Optional[List[float]]:
    """Per-cell blend, missing cells 1, clamped >= 0; all-ones -> None."""
    n = max(len(x) if x is not None else 0 for x in xss)
    if n == 0:
        return None
    out = [_clamp_ge0(_blend_n([x[i] if x is not None and i < len(x) else 1.0 for x in xss], cmc))
           for i in range(n)]
    return None if all(v == 1 for v in out) else out


def interpolate_xform(xfs: Sequence[Xform], cmc, use_log: bool,
                      wind: Optional[Tuple[float, float]] = None) -> Xform:
    """src/interpolate.ts:389-460 (two keyframes, with wind) and :1035-1068
    (N keyframes)."""
    density = _clamp_ge0(_blend_n([x.weight if x.active else 0.0 for x in xfs], cmc))
    color = _clamp01(_blend_n([x.color for x in xfs], cmc))
    speed = _clamp01(_blend_n([x.color_speed for x in xfs], cmc))
    opacity = _blend_n([x.opacity for x in xfs], cmc)
    affs = [x.affine for x in xfs]
    if use_log:
        aff = (_affine_log_polar(affs[0], affs[1], cmc[0], cmc[1], wind) if len(xfs) == 2
               else _affine_log_polar_n(affs, cmc))
    else:
        aff = _affine_linear_n(affs, cmc)
    post = None
    if any(x.post is not None for x in xfs):
        if not all(x.post is None or _is_identity_affine(x.post) for x in xfs):
            posts = [x.post if x.post is not None else Affine() for x in xfs]
            if use_log:
                ip = (_affine_log_polar(posts[0], posts[1], cmc[0], cmc[1]) if len(xfs) == 2
                      else _affine_log_polar_n(posts, cmc))
            else:
                ip = _affine_linear_n(posts, cmc)
            if not _is_identity_affine(ip):
                post = ip
    out = Xform(weight=density, color=color, color_speed=speed, affine=aff,
                variations=interpolate_variations([x.variations for x in xfs], cmc),
                opacity=opacity, post=post)
    if all(not x.active for x in xfs):
        out.active = False
    if any(x.xaos is not None for x in xfs):
        out.xaos = interpolate_xaos([x.xaos for x in xfs], cmc)
    return out


# -------------------------------------------------------------- palette --

def _rgb2hsv(r, g, b):
    """flam3's rgb2hsv, hue in [0, 6) (src/interpolate.ts:687-702)."""
    mx, mn = max(r, g, b), min(r, g, b)
    delta = mx - mn
    s = delta / mx if mx > 0 else 0.0
    h = 0.0
    if delta > 0:
        if mx == r:
            h = (g - b) / delta
        elif mx == g:
            h = 2 + (b - r) / delta
        else:
            h = 4 + (r - g) / delta
        if h < 0:
            h += 6
    return [h, s, mx]


def _hsv2rgb(h, s, v):
    hh = math.fmod(h, 6.0)
    if hh < 0:
        hh += 6
    i = math.floor(hh)
    f = hh - i
    p, q, t = v * (1 - s), v * (1 - s * f), v * (1 - s * (1 - f))
    return [(v, t, p), (q, v, p), (p, v, t), (p, q, v), (t, p, v)][i] if i < 5 else (v, p, q)


def _lut(g: Genome):
    lut = pack.bake_lut(g.palette.stops, g.palette.hue, g.palette.mode)
    return [[float(lut[i, c]) for c in range(3)] for i in range(PALETTE_SIZE)]


def interpolate_palette(anim: Animation, kfs: Sequence[Genome], cmc, name: str) -> Palette:
    """Both (all) palettes baked to 256 entries and blended per entry
    (src/interpolate.ts:611-685 for two keyframes, :1154-1211 for N)."""
    luts = [_lut(k) for k in kfs]
    mode = anim.palette_interpolation
    two = len(kfs) == 2
    dom = max(range(len(cmc)), key=lambda k: (cmc[k], -k))   # first argmax
    stops = []
    if mode == "sweep":
        for i in range(PALETTE_SIZE):
            src = (luts[0] if i < PALETTE_SIZE * cmc[0] else luts[1]) if two else luts[dom]
            stops.append(ColorStop(i / (PALETTE_SIZE - 1), *src[i]))
        return Palette(name, stops)
    rgb_frac = 1.0 if mode == "rgb" else (anim.hsv_rgb_palette_blend if mode == "hsv_circular"
                                         else 0.0)
    for i in range(PALETTE_SIZE):
        cols = [lt[i] for lt in luts]
        rgb = [_blend_n([c[ch] for c in cols], cmc) for ch in range(3)]
        hsvs = [_rgb2hsv(*c) for c in cols]
        if mode == "hsv_circular":
            if two:   # the first keyframe's hue moves by 6 towards the second's
                dh = hsvs[1][0] - hsvs[0][0]
                if dh > 3.0:
                    hsvs[0][0] += 6
                elif dh < -3.0:
                    hsvs[0][0] -= 6
            else:     # every hue moves towards the dominant keyframe's
                dom_h = hsvs[dom][0]
                for h in hsvs:
                    dh = h[0] - dom_h
                    if dh > 3.0:
                        h[0] -= 6
                    elif dh < -3.0:
                        h[0] += 6
        hsv_rgb = _hsv2rgb(*(_blend_n([h[k] for h in hsvs], cmc) for k in range(3)))
        stops.append(ColorStop(i / (PALETTE_SIZE - 1),
                               *(_clamp01(rgb_frac * rgb[ch] + (1 - rgb_frac) * hsv_rgb[ch])
                                 for ch in range(3))))
    return Palette(name, stops)


# -------------------------------------------------------------- grading --

def interpolate_tonemap(ts: Sequence[Optional[Tonemap]], cmc) -> Tonemap:
    a = [t if t is not None else INTERP_TONEMAP_FALLBACK for t in ts]
    return Tonemap(**{f: _blend_n([getattr(t, f) for t in a], cmc)
                      for f in ("gamma", "vibrancy", "highlight_power", "brightness",
                                "gamma_threshold")})


def _blend_hue_short_arc(hues, weights) -> float:
    h = list(hues)
    for k in range(1, len(h)):
        d = h[k] - h[k - 1]
        if d > 180:
            h[k] -= 360
        elif d < -180:
            h[k] += 360
    b = _blend_n(h, weights)
    return math.fmod(math.fmod(b + 180, 360) + 360, 360) - 180


def interpolate_hsl_adjust(adjs, cmc) -> Optional[Dict[str, float]]:
    """Fades from the identity (0, 100, 0) where a keyframe has none; hue on
    the shorter arc, normalised to [-180, 180) (src/interpolate.ts:768-794)."""
    if all(a is None for a in adjs):
        return None
    f = [a if a is not None else IDENTITY_HSL for a in adjs]
    return {"hue": _blend_hue_short_arc([a["hue"] for a in f], cmc),
            "sat": _blend_n([a["sat"] for a in f], cmc),
            "light": _blend_n([a["light"] for a in f], cmc)}


def _downsample_keeping_ends(xs, k):
    n = len(xs)
    out = []
    for i in range(k):
        x = xs[_js_round(i * (n - 1) / (k - 1))]
        if not out or x != out[-1]:
            out.append(x)
    return out


def _blend_curve(points, weights):
    """The weighted sum of the curves, sampled at the union of their control
    x's (deduplicated within 1e-6, at most 8 kept, ends included)
    (src/interpolate.ts:829-868)."""
    xs = []
    for x in sorted(p[0] for pts in points for p in pts):
        if not xs or x - xs[-1] > 1e-6:
            xs.append(x)
    if len(xs) > cv.MAX_POINTS:
        xs = _downsample_keeping_ends(xs, cv.MAX_POINTS)
    return [(x, _clamp01(_blend_n([cv.eval_curve(p, x) for p in points], weights))) for x in xs]


def interpolate_channel_curves(cs, cmc):
    if all(c is None for c in cs):
        return None
    f = [c if c is not None else cv.identity_curves() for c in cs]
    return {ch: _blend_curve([c[ch] for c in f], cmc) for ch in cv.CHANNELS}


# ---------------------------------------------------------- interpolate --

def _render_fields(out: Genome, kfs: Sequence[Genome], cmc, shape_from: Genome) -> None:
    """Continuous render fields: blended when every keyframe has them, else
    the first one present (src/interpolate.ts:201-241, :957-995)."""
    def first(get):
        for k in kfs:
            v = get(k)
            if v is not None:
                return v
        return None

    def have(get):
        return all(get(k) is not None for k in kfs)
    if have(lambda k: k.quality):
        out.quality = _blend_n([k.quality for k in kfs], cmc)
    else:
        out.quality = first(lambda k: k.quality)
    if have(lambda k: k.density):
        out.density = Density(*(_blend_n([getattr(k.density, f) for k in kfs], cmc)
                                for f in ("max_rad", "min_rad", "curve")))
    else:
        out.density = first(lambda k: k.density)
    if have(lambda k: k.filter_radius):
        out.filter_radius = _blend_n([k.filter_radius for k in kfs], cmc)
    else:
        out.filter_radius = first(lambda k: k.filter_radius)
    if have(lambda k: k.background):
        out.background = [_blend_n([k.background[c] for k in kfs], cmc) for c in range(3)]
    else:
        out.background = first(lambda k: k.background)
    if have(lambda k: k.size):
        out.size = {d: _js_round(_blend_n([k.size[d] for k in kfs], cmc))
                    for d in ("width", "height")}
    else:
        out.size = first(lambda k: k.size)
    if have(lambda k: k.oversample):
        out.oversample = _js_round(_blend_n([k.oversample for k in kfs], cmc))
    else:
        out.oversample = first(lambda k: k.oversample)
    out.palette_mode = shape_from.palette_mode


def _finish(out: Genome, kfs: Sequence[Genome], cmc, anchor: Genome) -> Genome:
    out.nick = anchor.nick
    if any(k.tonemap is not None for k in kfs):
        out.tonemap = interpolate_tonemap([k.tonemap for k in kfs], cmc)
    if any(k.rotate != 0 for k in kfs):
        out.rotate = _blend_n([k.rotate for k in kfs], cmc)
    if any(k.xform_blend != 0 for k in kfs):
        out.xform_blend = _blend_n([k.xform_blend for k in kfs], cmc)
    _render_fields(out, kfs, cmc, anchor)
    out.hsl_adjust = interpolate_hsl_adjust([k.hsl_adjust for k in kfs], cmc)
    out.channel_curves = interpolate_channel_curves([k.channel_curves for k in kfs], cmc)
    return out


def interpolate(anim: Animation, time: float) -> Genome:
    """The genome of `anim` at `time` (src/interpolate.ts:58-257)."""
    kf = anim.keyframes
    if len(kf) < 2:
        raise ValueError("pyr3: interpolate requires Animation.keyframes.length >= 2")
    i1, i2, raw_c0, raw_c1 = pick_keyframes(kf, time)
    k0, k1 = bake_symmetry(kf[i1]), bake_symmetry(kf[i2])
    c0, c1 = raw_c0, raw_c1
    ease = anim.segment_easing[i1] if anim.segment_easing and i1 < len(anim.segment_easing) \
        else None
    if ease and 0 <= raw_c1 <= 1:
        c1 = eval_easing(ease, raw_c1)
        c0 = 1 - c1
    use_log = anim.interpolation_type == "log"
    if anim.interpolation == "smooth" and i1 != 0 and i2 != len(kf) - 1:
        return _interpolate_catmull_rom(
            [bake_symmetry(kf[j]) for j in (i1 - 1, i1, i1 + 1, i1 + 2)],
            catmull_rom_weights(c1), anim, time, raw_c1)
    a0, a1 = align_xform_counts([k0, k1], anim.interpolation_type)
    perms = anim.segment_permutation
    perm = resolve_segment_permutation(perms[i1] if perms and i1 < len(perms) else None,
                                       len(a1.xforms))
    b1 = a1
    if perm is not None:
        b1 = copy.copy(a1)
        b1.xforms = [a1.xforms[j] for j in perm]
    wind = establish_wind(a0, b1) if use_log else None
    xforms = []
    for i in range(len(a0.xforms)):
        x0 = apply_motion_parameters(a0.xforms[i], raw_c1)
        x1 = apply_motion_parameters(b1.xforms[i], raw_c1)
        xforms.append(interpolate_xform([x0, x1], (c0, c1), use_log,
                                        wind[i] if wind is not None else None))
    final = None
    if a0.finalxform is not None and a1.finalxform is not None:
        final = interpolate_xform([a0.finalxform, a1.finalxform], (c0, c1), use_log)
    out = Genome(name=k0.name, xforms=xforms, scale=_blend(k0.scale, k1.scale, c0, c1),
                 cx=_blend(k0.cx, k1.cx, c0, c1), cy=_blend(k0.cy, k1.cy, c0, c1),
                 palette=interpolate_palette(anim, [k0, k1], (c0, c1), k0.palette.name),
                 finalxform=final, time=time)
    return _finish(out, [k0, k1], (c0, c1), k0)


def _interpolate_catmull_rom(kfs: List[Genome], cmc, anim: Animation, time: float,
                             raw_c1: float) -> Genome:
    """src/interpolate.ts:888-1005: the four keyframes blended with the cubic
    basis; no wind, no permutation."""
    use_log = anim.interpolation_type == "log"
    al = align_xform_counts(kfs, anim.interpolation_type)
    xforms = [interpolate_xform([apply_motion_parameters(g.xforms[i], raw_c1) for g in al],
                                cmc, use_log)
              for i in range(len(al[0].xforms))]
    final = None
    if all(g.finalxform is not None for g in al):
        final = interpolate_xform([g.finalxform for g in al], cmc, use_log)
    out = Genome(name=kfs[1].name, xforms=xforms,
                 scale=_blend_n([k.scale for k in kfs], cmc),
                 cx=_blend_n([k.cx for k in kfs], cmc), cy=_blend_n([k.cy for k in kfs], cmc),
                 palette=interpolate_palette(anim, kfs, cmc, kfs[1].palette.name),
                 finalxform=final, time=time)
    return _finish(out, kfs, cmc, kfs[1])


# ------------------------------------------------------------ rendering --

def packed_shape(pg) -> tuple:
    """The chaos-kernel shape of a packed genome: per record its slot count,
    post-affine presence and variation indices, plus the final slot."""
    x = pg.xforms
    rows = []
    for r in range(x.shape[0]):
        n = min(int(x[r, 7]), 8)
        rows.append((n, bool(x[r, 15] != 0), tuple(int(x[r, 20 + 4 * k]) for k in range(n))))
    return (pg.distrib.shape[0], pack.final_xform_slot(pg.expanded), tuple(rows))


@dataclass
class FrameResult:
    center: Genome
    subframes: int
    total_walkers: int
    total_samples: int
    walkers: List[int] = field(default_factory=list)
    seeds: List[int] = field(default_factory=list)
    genomes: List[Genome] = field(default_factory=list)


def render_animation_frame(renderer, anim: Animation, time: float, seed: int,
                           walkers: Optional[int] = None, force_de_off: bool = False,
                           transparent: bool = False, walker_jitter: Optional[float] = None,
                           fuse: int = pack.FUSE) -> FrameResult:
    """One output frame at `time` with the animation's motion blur
    (src/animate-render.ts:73-189): reset with the centre genome, one dispatch
    per sub-frame (batched per kernel shape), present with the centre genome
    calibrated on the total samples. `walkers` pins the total walker count
    (the reference's --walkers override); by default the centre genome's
    quality is sized as the reference's computeDispatch does."""
    n = max(1, int(math.floor(anim.ntemporal_samples)))
    center = interpolate(anim, time)
    spp = center.quality if center.quality is not None else pack.DEFAULT_SPP
    total_w, iters, _ = pack.compute_dispatch(spp, renderer.width, renderer.height, walkers)
    renderer.reset(center)
    base = seed & 0xFFFFFFFF
    if n == 1:
        renderer.iterate(center, base, total_w, iters, walker_jitter, fuse)
        renderer.present(center, total_w * iters, force_de_off, transparent)
        return FrameResult(center, 1, total_w, total_w * iters, [total_w], [base], [center])
    tf = temporal_filter(n, anim.temporal_filter_type, anim.temporal_filter_width,
                         anim.temporal_filter_exp)
    from .batch import apportion
    counts = apportion(total_w, tf.filter)
    subs, seeds, ws = [], [], []
    for i in range(n):
        if counts[i] == 0:
            continue
        subs.append(interpolate(anim, time + tf.deltas[i]))
        seeds.append((base + i * SUBFRAME_SEED_STEP) & 0xFFFFFFFF)
        ws.append(counts[i])
    from .renderer import pack_genome
    pgs = [pack_genome(g) for g in subs]
    i = 0
    while i < len(subs):   # runs of one kernel shape, at most 64 per launch
        key = packed_shape(pgs[i])
        j = i + 1
        while j < len(subs) and j - i < 64 and packed_shape(pgs[j]) == key:
            j += 1
        if j - i == 1:
            renderer.iterate(subs[i], seeds[i], ws[i], iters, walker_jitter, fuse, packed=pgs[i])
        else:
            renderer.iterate_multi(subs[i:j], seeds[i:j], ws[i:j], iters, walker_jitter, fuse,
                                   packed=pgs[i:j])
        i = j
    tot = sum(ws)
    renderer.present(center, tot * iters, force_de_off, transparent)
    return FrameResult(center, n, tot, tot * iters, ws, seeds, subs)
