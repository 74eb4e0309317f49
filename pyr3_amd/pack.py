"""Host-side packers: genome -> the byte blobs the render path consumes.

Every blob is byte-compatible with the reference's GPU uploads so the same
bytes could feed either backend:

* xform records, 116 f32 = 464 B each   (src/genome.ts:328-421 packXformInto/packXforms)
* final-xform slot index                 (src/genome.ts:432-436 finalXformSlot)
* xform-pick table, u32 x 16384 per row  (src/genome.ts:495-534 packXformDistrib)
* palette LUT, 256 x RGBA f32            (src/palette.ts:71-125 bakeLUT, :211 packPalette)
* chaos uniforms, 32 x 4 B               (src/chaos.ts:381-456)
* DE uniforms, 8 x 4 B                   (src/density.ts:147-175)
* tonemap uniforms, 24 x 4 B             (src/visualize.ts:159-205)
* dispatch sizing and calibration        (src/renderer.ts:57-78, src/calibration.ts:37-43)
* spatial filter taps                    (src/spatial-filter.ts:217-256)
* symmetry expansion                     (src/symmetry.ts:21-124)

The pick table is built with a vectorised exact form of the reference's f64
scan: the running slot position r and the running weight total t are both
sequential f64 sums (numpy's add.accumulate is sequential), and the while-loop
``while r >= t: j++`` picks the first j with t_j > r, i.e. searchsorted(right).
"""
from __future__ import annotations

import bisect
import copy
import dataclasses
import math
import struct
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np

from .catalog import DC_VARIATIONS
from .genome import Affine, Genome, Tonemap, Variation, Xform

XFORM_FLOATS = 116
XFORM_BYTES = XFORM_FLOATS * 4
MAX_VARS = 8
GRAIN = 16384
MAX_XFORMS = 128
PALETTE_SIZE = 256
CHAOS_UNIFORM_WORDS = 32
DE_UNIFORM_WORDS = 8
VIZ_UNIFORM_WORDS = 24
DEFAULT_WALKER_JITTER = 1e-7        # src/chaos.ts:96
FUSE = 200                          # src/renderer.ts:40
CHAOS_BLOCK = 640                   # walkers per chaos block (kChaosBlock in p3_args.h)
CHIP_WALKERS = 256 * CHAOS_BLOCK    # one chaos block on each of the 256 CUs
TARGET_WALKERS = 1024               # src/renderer.ts:26
MIN_ITERS_PER_WALKER = 4096
MAX_ITERS_PER_WALKER = 1048576
MAX_WALKERS = 65535 * 64
DEFAULT_SPP = 16
DEFAULT_FILTER_RADIUS = 0.5
PRE_BLUR = 97


# ------------------------------------------------------------ symmetry --

def _round6(x: float) -> float:
    # JS Math.round(x * 1e6) / 1e6 (round half up toward +inf)
    return math.floor(x * 1e6 + 0.5) / 1e6


def _identity_op(a, b, d, e, color) -> Xform:
    return Xform(weight=1.0, color=color, color_speed=0.0, affine=Affine(a, b, 0.0, d, e, 0.0),
                 variations=[Variation(index=0, weight=1.0)])


def symmetry_xforms(sym) -> List[Xform]:
    out: List[Xform] = []
    n = sym["n"]
    if sym["kind"] == "rotational" and n == 1:
        return out
    if sym["kind"] == "dihedral":
        out.append(_identity_op(-1.0, 0.0, 0.0, 1.0, 1.0))
    a = 2.0 * math.pi / n
    for k in range(1, n):
        c = _round6(math.cos(k * a))
        s = _round6(math.sin(k * a))
        color = 0.0 if n < 3 else (k - 1) / (n - 2)
        out.append(_identity_op(c, -s, s, c, color))
    return out


def expand_genome_for_gpu(g: Genome) -> Genome:
    """Zero the weights of inactive xforms/variations, then append symmetry
    xforms. The result shares every object it does not change with `g`
    (palette, records of active xforms): it is read, never written."""
    w = copy.copy(g)
    xs = []
    for x in g.xforms:
        if not x.active or not all(v.active for v in x.variations):
            x = copy.copy(x)
            if not x.active:
                x.weight = 0.0
            x.variations = [v if v.active else dataclasses.replace(v, weight=0.0)
                            for v in x.variations]
        xs.append(x)
    w.xforms = xs
    if w.symmetry:
        extra = symmetry_xforms(w.symmetry)
        if extra:
            total = len(w.xforms) + len(extra)
            if total > MAX_XFORMS:
                raise ValueError(
                    f"pyr3: symmetry expansion exceeds MAX_XFORMS ({MAX_XFORMS}): "
                    f"{len(w.xforms)} source + {len(extra)} generated = {total}. "
                    "Reduce source xforms or symmetry n.")
            w.xforms = w.xforms + extra
        w.symmetry = None
    return w


# -------------------------------------------------------------- xforms --

def _pack_xform(buf: np.ndarray, x: Xform) -> None:
    nv = min(len(x.variations), MAX_VARS)
    a = x.affine
    buf[0:4] = (a.a, a.b, a.c, x.weight)
    buf[4:8] = (a.d, a.e, a.f, nv)
    has_dc = any(v.active and v.index in DC_VARIATIONS for v in x.variations)
    buf[8:12] = (x.color, x.color_speed, max(0.0, min(1.0, x.opacity)), 1.0 if has_dc else 0.0)
    if x.post is not None:
        p = x.post
        buf[12:16] = (p.a, p.b, p.c, 1.0)
        buf[16:20] = (p.d, p.e, p.f, 0.0)
    for k in range(nv):
        v = x.variations[k]
        buf[20 + 4 * k: 24 + 4 * k] = (v.index, v.weight, v.params[0], v.params[1])
        buf[52 + 4 * k: 56 + 4 * k] = v.params[2:6]
        buf[84 + 4 * k: 88 + 4 * k] = v.params[6:10]


def pack_xforms(g: Genome) -> np.ndarray:
    """(n_xforms [+1 final]) x 116 float32, final xform last."""
    n = len(g.xforms) + (1 if g.finalxform is not None else 0)
    out = np.zeros((n, XFORM_FLOATS), dtype=np.float32)
    for i, x in enumerate(g.xforms):
        _pack_xform(out[i], x)
    if g.finalxform is not None:
        _pack_xform(out[len(g.xforms)], g.finalxform)
    return out


def final_xform_slot(g: Genome) -> int:
    return len(g.xforms) if (g.finalxform is not None and g.finalxform.active) else -1


def _distrib_row(weights: np.ndarray) -> np.ndarray:
    n = len(weights)
    if not np.all(np.isfinite(weights)) or np.any(weights < 0):
        bad = weights[~(np.isfinite(weights) & (weights >= 0))][0]
        raise ValueError(f"xform weight must be non-negative finite: got {bad}")
    total = 0.0
    for d in weights:          # sequential f64 sum, as the reference scan
        total += float(d)
    if total == 0.0:
        return np.zeros(GRAIN, dtype=np.uint32)
    dr = total / GRAIN
    steps = np.full(GRAIN, dr, dtype=np.float64)
    steps[0] = 0.0
    r = np.add.accumulate(steps)                       # r_i = r_{i-1} + dr
    t = np.add.accumulate(weights.astype(np.float64))  # t_j = t_{j-1} + w_j
    j = np.searchsorted(t, r, side="right")            # first j with t_j > r_i
    return np.minimum(j, n - 1).astype(np.uint32)


def pack_xform_distrib(g: Genome) -> np.ndarray:
    """(n + 1) x 16384 uint32: rows "previous xform = i" then the no-previous row."""
    n = len(g.xforms)
    w = np.array([x.weight for x in g.xforms], dtype=np.float64)
    rows = []
    for i, xi in enumerate(g.xforms):
        m = np.ones(n)
        if xi.xaos is not None:
            for j in range(min(n, len(xi.xaos))):
                m[j] = xi.xaos[j]
        rows.append(_distrib_row(w * m))
    rows.append(_distrib_row(w))
    return np.stack(rows)


# ------------------------------------------------------------- palette --

def _rgb_to_hsv(r, g, b):
    mx, mn = max(r, g, b), min(r, g, b)
    d = mx - mn
    v, s = mx, (0.0 if mx == 0 else d / mx)
    h = 0.0
    if d != 0:
        if mx == r:
            h = math.fmod((g - b) / d, 6.0)
        elif mx == g:
            h = (b - r) / d + 2
        else:
            h = (r - g) / d + 4
        h *= 60
        if h < 0:
            h += 360
    return h, s, v


def _hsv_to_rgb(h, s, v):
    h = math.fmod(math.fmod(h, 360.0) + 360.0, 360.0)
    c = v * s
    x = c * (1 - abs(math.fmod(h / 60, 2) - 1))
    m = v - c
    rp = gp = bp = 0.0
    if h < 60:
        rp, gp = c, x
    elif h < 120:
        rp, gp = x, c
    elif h < 180:
        gp, bp = c, x
    elif h < 240:
        gp, bp = x, c
    elif h < 300:
        rp, bp = x, c
    else:
        rp, bp = c, x
    return rp + m, gp + m, bp + m


def _catmull_rom(u, xa, xb, xc, xd):
    B = 0.5
    c = u * u * u * (-B * xa + (2 - B) * xb + (B - 2) * xc + B * xd)
    c += u * u * (2 * B * xa + (B - 3) * xb + (3 - 2 * B) * xc - B * xd)
    c += u * (-B * xa + B * xc)
    return c + xb


def bake_lut(stops, hue: float = 0.0, mode: str = "linear") -> np.ndarray:
    """256 x RGBA float32 (alpha 0) from gradient stops."""
    if hue != 0:
        adj = []
        for s in stops:
            h, sat, v = _rgb_to_hsv(s.r, s.g, s.b)
            r, g, b = _hsv_to_rgb(h + hue, sat, v)
            adj.append((s.t, r, g, b))
    else:
        adj = [(s.t, s.r, s.g, s.b) for s in stops]
    srt = sorted(adj, key=lambda s: s[0])
    ts = [s[0] for s in srt]
    rows = []  # f64 per entry, rounded to f32 once at the end
    for i in range(PALETTE_SIZE):
        t = i / (PALETTE_SIZE - 1)
        # first k with ts[k] <= t <= ts[k+1] (ts sorted), else the whole span
        lo_i, hi_i = 0, len(srt) - 1
        k = max(bisect.bisect_left(ts, t) - 1, 0)
        if k <= len(srt) - 2 and k < bisect.bisect_right(ts, t):
            lo_i, hi_i = k, k + 1
        lo, hi = srt[lo_i], srt[hi_i]
        if mode == "step":
            rows.append((lo[1], lo[2], lo[3], 0.0))
        else:
            span = (hi[0] - lo[0]) or 1
            u = max(0.0, min(1.0, (t - lo[0]) / span))
            if mode == "smooth" and len(srt) >= 3:
                a = srt[max(0, lo_i - 1)]
                d = srt[min(len(srt) - 1, hi_i + 1)]
                rows.append(tuple(max(0.0, min(1.0, _catmull_rom(u, a[1 + c], lo[1 + c],
                                                                 hi[1 + c], d[1 + c])))
                                  for c in range(3)) + (0.0,))
            else:
                rows.append(tuple(lo[1 + c] + (hi[1 + c] - lo[1 + c]) * u for c in range(3))
                            + (0.0,))
    return np.array(rows, dtype=np.float32)


def pack_palette(g: Genome) -> np.ndarray:
    return bake_lut(g.palette.stops, g.palette.hue, g.palette.mode)


# ------------------------------------------------------------ uniforms --

def _f32_bits(v: float) -> int:
    return struct.unpack("<I", struct.pack("<f", v))[0]


COLOR_MODES = {"palette": 0, "flow": 1, "trap-distance": 2, "phase": 3}
DEFAULT_FLOW_SCALE = 2.0             # src/chaos.ts:139


@dataclass
class ColorOpts:
    """Splat-colour options of a dispatch (src/chaos.ts:60-80, 421-450): the
    colour mode, its parameters, and capture of the colour index."""
    mode: str = "palette"
    flow_strength: float = 1.0
    flow_scale: float = DEFAULT_FLOW_SCALE
    trap_kind: str = "point"         # point | circle | line
    trap_mode: str = "glow"          # glow | rings
    trap_cx: float = 0.0
    trap_cy: float = 0.0
    trap_radius: float = 0.5
    trap_angle: float = 0.0          # degrees (line normal (-sin, cos))
    trap_falloff: float = 2.0
    trap_freq: float = 4.0
    trap_strength: float = 1.0
    phase_strength: float = 1.0
    phase_freq: float = 1.0
    capture_index: bool = False


def chaos_uniforms(g: Genome, width_s: int, height_s: int, oversample: int, iters: int,
                   fuse: int, seed: int, walkers: int, jitter: float = DEFAULT_WALKER_JITTER,
                   trace: bool = False, color: "ColorOpts" = None) -> np.ndarray:
    """The 128-byte chaos uniform block (src/chaos.ts:381-456)."""
    c = color or ColorOpts()
    if c.mode not in COLOR_MODES:
        raise ValueError(f"color mode must be one of {sorted(COLOR_MODES)}, got {c.mode!r}")
    if c.trap_kind not in ("point", "circle", "line") or c.trap_mode not in ("glow", "rings"):
        raise ValueError(f"bad trap kind/mode: {c.trap_kind!r}/{c.trap_mode!r}")
    u = np.zeros(CHAOS_UNIFORM_WORDS, dtype=np.uint32)
    u[0], u[1], u[2], u[3] = width_s, height_s, iters, fuse
    u[4] = _f32_bits(g.scale * oversample)
    u[5] = _f32_bits(g.cx)
    u[6] = _f32_bits(g.cy)
    u[7] = seed & 0xFFFFFFFF
    u[8] = final_xform_slot(g) & 0xFFFFFFFF
    u[9] = _f32_bits((g.rotate or 0.0) * math.pi / 180.0)
    u[10] = 1 if g.palette_mode == "linear" else 0
    u[11] = 1 if trace else 0
    u[12] = walkers
    u[13] = _f32_bits(jitter)
    u[14] = 1 if c.capture_index else 0
    u[15] = COLOR_MODES[c.mode]
    u[16] = _f32_bits(c.flow_strength)
    u[17] = _f32_bits(c.flow_scale)
    u[18] = {"point": 0, "circle": 1, "line": 2}[c.trap_kind]
    u[19] = 1 if c.trap_mode == "rings" else 0
    ang = c.trap_angle * math.pi / 180.0
    for i, v in ((20, c.trap_cx), (21, c.trap_cy), (22, c.trap_radius), (23, -math.sin(ang)),
                 (24, math.cos(ang)), (25, c.trap_falloff), (26, c.trap_freq),
                 (27, c.trap_strength), (28, c.phase_strength), (29, c.phase_freq)):
        u[i] = _f32_bits(v)
    u[30] = _f32_bits(min(1.0, max(0.0, g.xform_blend or 0.0)))
    return u


def de_uniforms(width_s: int, height_s: int, max_rad: float, min_rad: float, curve: float,
                k1: float, k2: float, oversample: int) -> np.ndarray:
    """The 32-byte DEUniforms block; radii scaled to super-pixels (r * os + 1)."""
    u = np.zeros(DE_UNIFORM_WORDS, dtype=np.uint32)
    u[0], u[1] = width_s, height_s
    u[2] = _f32_bits(max_rad * oversample + 1)
    u[3] = _f32_bits(min_rad * oversample + 1)
    u[4] = _f32_bits(curve)
    u[5] = _f32_bits(k1)
    u[6] = _f32_bits(k2)
    return u


def viz_uniforms(width: int, height: int, k1: float, k2: float, tm: Tonemap, oversample: int,
                 fwidth: int, background=(0.0, 0.0, 0.0), transparent: bool = False,
                 curves_mask: int = 0, hsl_adjust=None) -> np.ndarray:
    """The 96-byte VizUniforms block (src/visualize.ts:177-201): word 16 the
    channel-curves bit mask, 17..20 HSL on, hue (degrees), sat / 100 and
    light / 100 when `hsl_adjust` ({'hue', 'sat', 'light'}) is given."""
    u = np.zeros(VIZ_UNIFORM_WORDS, dtype=np.uint32)
    u[16] = curves_mask
    if hsl_adjust is not None:
        u[17] = 1
        u[18] = _f32_bits(hsl_adjust["hue"])
        u[19] = _f32_bits(hsl_adjust["sat"] / 100.0)
        u[20] = _f32_bits(hsl_adjust["light"] / 100.0)
    u[0], u[1] = width, height
    for i, v in enumerate((k1, k2, tm.gamma, tm.vibrancy, tm.highlight_power, tm.gamma_threshold)):
        u[2 + i] = _f32_bits(v)
    u[8], u[9] = oversample, fwidth
    for i in range(3):
        u[12 + i] = _f32_bits(background[i])
    u[21] = 1 if transparent else 0
    return u


# ------------------------------------------------- sizing / calibration --

def compute_dispatch(target_spp: float, width: int, height: int,
                     walkers_override: int = None) -> Tuple[int, int, int]:
    """(walkers, iters_per_walker, samples) as src/renderer.ts:57-78."""
    target = _js_round(target_spp * width * height)
    walkers = walkers_override if walkers_override is not None else TARGET_WALKERS
    iters = -(-target // walkers)
    if walkers_override is None:
        if iters < MIN_ITERS_PER_WALKER:
            iters = MIN_ITERS_PER_WALKER
            walkers = max(1, -(-target // iters))
        elif iters > MAX_ITERS_PER_WALKER:
            iters = MAX_ITERS_PER_WALKER
            walkers = min(MAX_WALKERS, -(-target // iters))
    else:
        iters = max(1, iters)
    return walkers, iters, walkers * iters


def _js_round(x: float) -> int:
    return int(math.floor(x + 0.5))


def calibration(scale: float, samples: float, brightness: float, oversample: int):
    k1 = brightness * 255.0 * 268.0 / 256.0
    k2 = (oversample * oversample) * scale * scale / (255.0 * samples)
    return k1, k2


def spatial_kernel(radius: float, oversample: int) -> np.ndarray:
    """Gaussian collapse taps (support 1.5), parity-matched to the oversample."""
    if not radius > 0:
        raise ValueError(f"pyr3: spatial-filter radius must be positive, got {radius}")
    if oversample < 1 or int(oversample) != oversample:
        raise ValueError(f"pyr3: oversample must be a positive integer, got {oversample}")
    supp = 1.5
    fw = 2.0 * supp * oversample * radius
    fwidth = int(math.floor(fw)) + 1
    if (fwidth ^ oversample) & 1:
        fwidth += 1
    fwidth = max(3, min(17, fwidth))
    adjust = supp * fwidth / fw if fw > 0 else 1.0
    vals = []
    out = np.zeros(fwidth, dtype=np.float32)
    for i in range(fwidth):
        x = ((2.0 * i + 1.0) / fwidth - 1.0) * adjust
        v = math.exp(-2.0 * x * x)
        out[i] = v
        vals.append(v)
    s = 0.0
    for v in vals:
        s += v
    if s != 0:
        out = (out.astype(np.float64) / s).astype(np.float32)
    return out
