"""Output side of the render path: 16-bit PNG, OpenEXR, scene-linear export,
render-time auto-exposure.

* `display_to_half` rounds the tonemap's f32 RGBA to IEEE binary16 (round to
  nearest even), the values the reference's rgba16float target holds
  (bin/pyr3-render.ts:416-424); `half_to_float` decodes binary16
  (src/half-float.ts:1-19); `srgb_to_linear` is the inverse sRGB transfer
  (src/srgb.ts:14-17).
* `display_half_to_png16` / `display_half_to_linear_exr` follow
  src/gpu-readback.ts:83-115: clamp to [0, 1] (NaN -> 0), then 16-bit samples
  (x 65535, rounded half up) or linear light (alpha not sRGB-curved).
* `encode_exr` writes the reference's uncompressed single-part scanline
  OpenEXR (src/exr-encode.ts:17-94): channels A, B, G, R as FLOAT, data and
  display window (0, 0)-(W-1, H-1), increasing-y line order, one offset
  table entry and one scanline block per row.
* Scene-linear RGBA (histogramToLinearRgba, src/export-linear.ts:24-47) is
  computed on the device: `Renderer.read_linear(k1, k2)`.
* Auto-exposure (src/auto-exposure.ts): the mean luminance of a preview-tier
  probe render (1024 long edge, 16 spp, fixed seed) is the target; the
  brightness is refitted by re-presenting the same histogram until the mean
  luminance is within 10 % of it (at most 4 passes, brightness clamped to
  [1e-3, 1e4]).
"""
from __future__ import annotations

import copy
import math
import struct
from dataclasses import dataclass
from typing import Callable

import numpy as np

from .genome import Genome, Tonemap

AE_PROBE_LONG_EDGE = 1024
AE_PROBE_QUALITY = 16
AE_PROBE_SEED = 0x9E3779B9
AE_DEADBAND = 0.10
AE_MAX_ITERS = 4
AE_BRIGHTNESS_MIN = 1e-3
AE_BRIGHTNESS_MAX = 1e4


# ----------------------------------------------------------- conversions --

def display_to_half(rgba_f32: np.ndarray) -> np.ndarray:
    """f32 RGBA -> binary16 bit patterns (u16), round to nearest even."""
    return np.asarray(rgba_f32, dtype=np.float32).astype(np.float16).view(np.uint16)


def half_to_float(h: int) -> float:
    sign = (h & 0x8000) >> 15
    exp = (h & 0x7C00) >> 10
    frac = h & 0x03FF
    if exp == 0:
        val = frac * 2.0 ** -24
    elif exp == 0x1F:
        val = math.inf if frac == 0 else math.nan
    else:
        val = (1 + frac / 1024) * 2.0 ** (exp - 15)
    return -val if sign else val


def _halves(bits: np.ndarray) -> np.ndarray:
    return np.asarray(bits, dtype=np.uint16).view(np.float16).astype(np.float64)


def _clamp01_nan0(f: np.ndarray) -> np.ndarray:
    return np.where(np.isfinite(f), np.clip(f, 0.0, 1.0), 0.0)


def srgb_to_linear(c):
    c = np.asarray(c, dtype=np.float64)
    return np.where(c <= 0.04045, c / 12.92, ((c + 0.055) / 1.055) ** 2.4)


def display_half_to_png16(bits: np.ndarray) -> np.ndarray:
    f = _halves(bits)
    # NaN samples become 0 (a NaN sample coerces to 0 on the u16 store)
    v = np.where(np.isnan(f), 0.0, np.clip(f, 0.0, 1.0)) * 65535.0
    return np.floor(v + 0.5).astype(np.uint16)


def display_half_to_linear_exr(bits: np.ndarray) -> np.ndarray:
    f = _clamp01_nan0(_halves(bits))
    out = srgb_to_linear(f)
    out[..., 3] = f[..., 3]
    return out.astype(np.float32)


# ------------------------------------------------------------------ EXR --

def _cstr(s: str) -> bytes:
    return s.encode("utf-8") + b"\x00"


def _attr(name: str, kind: str, payload: bytes) -> bytes:
    return _cstr(name) + _cstr(kind) + struct.pack("<i", len(payload)) + payload


def encode_exr(rgba: np.ndarray) -> bytes:
    """(H, W, 4) f32 linear RGBA -> uncompressed scanline OpenEXR bytes."""
    img = np.ascontiguousarray(rgba, dtype=np.float32)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError(f"expected (H, W, 4) RGBA, got {img.shape}")
    h, w, _ = img.shape
    head = struct.pack("<ii", 0x01312F76, 2)
    chl = b"".join(_cstr(c) + struct.pack("<i", 2) + b"\x00\x00\x00\x00" + struct.pack("<ii", 1, 1)
                   for c in "ABGR") + b"\x00"
    box = struct.pack("<iiii", 0, 0, w - 1, h - 1)
    head += (_attr("channels", "chlist", chl) + _attr("compression", "compression", b"\x00")
             + _attr("dataWindow", "box2i", box) + _attr("displayWindow", "box2i", box)
             + _attr("lineOrder", "lineOrder", b"\x00")
             + _attr("pixelAspectRatio", "float", struct.pack("<f", 1.0))
             + _attr("screenWindowCenter", "v2f", struct.pack("<ff", 0.0, 0.0))
             + _attr("screenWindowWidth", "float", struct.pack("<f", 1.0)) + b"\x00")
    row_bytes = 4 * w * 4
    block = 8 + row_bytes
    first = len(head) + 8 * h
    offsets = struct.pack(f"<{h}Q", *(first + y * block for y in range(h)))
    planes = img[..., [3, 2, 1, 0]].transpose(0, 2, 1)       # (H, channel A B G R, W)
    rows = np.empty((h, block), dtype=np.uint8)
    rows[:, 0:4] = np.arange(h, dtype="<i4").view(np.uint8).reshape(h, 4)
    rows[:, 4:8] = np.full(h, row_bytes, dtype="<i4").view(np.uint8).reshape(h, 4)
    rows[:, 8:] = np.ascontiguousarray(planes, dtype="<f4").view(np.uint8).reshape(h, row_bytes)
    return head + offsets + rows.tobytes()


# -------------------------------------------------------- auto-exposure --

def compute_mean_luminance(rgba8: np.ndarray) -> float:
    """Mean of (R + G + B) / 3 over RGBA8 pixels in [0, 255], summed in
    pixel order in f64 (src/auto-exposure.ts:74-82)."""
    px = np.asarray(rgba8, dtype=np.uint8).reshape(-1, 4)
    if px.shape[0] == 0:
        return 0.0
    v = (px[:, 0].astype(np.float64) + px[:, 1] + px[:, 2]) / 3.0
    return float(np.cumsum(v)[-1] / px.shape[0])


def compute_mean_luminance_half(bits: np.ndarray) -> float:
    """The same measure on binary16 RGBA, each channel clamped to [0, 1]
    (non-finite -> 0) then x 255 (src/auto-exposure.ts:87-101)."""
    f = _clamp01_nan0(_halves(np.asarray(bits).reshape(-1, 4)))
    if f.shape[0] == 0:
        return 0.0
    v = ((f[:, 0] + f[:, 1] + f[:, 2]) / 3.0) * 255.0
    return float(np.cumsum(v)[-1] / f.shape[0])


def _clamp(x: float, lo: float, hi: float) -> float:
    return min(hi, max(lo, x)) if math.isfinite(x) else lo


@dataclass
class BrightnessFit:
    brightness: float
    corrected: bool
    iters: int
    final_mean: float
    factor: float


def fit_brightness_to_target(base: float, target: float, initial: float,
                             rerender: Callable[[float], float]) -> BrightnessFit:
    """src/auto-exposure.ts:147-175: multiply the brightness by
    target / mean and re-present until within the deadband."""
    finite = math.isfinite(target) and math.isfinite(initial)
    factor0 = target / initial if finite and initial > 0 else 1.0
    if not finite or target <= 0 or initial <= 0 or abs(factor0 - 1) <= AE_DEADBAND:
        return BrightnessFit(base, False, 0, initial, factor0)
    b, mean, iters = base, initial, 0
    for _ in range(AE_MAX_ITERS):
        factor = target / mean
        if abs(factor - 1) <= AE_DEADBAND:
            break
        b = _clamp(b * factor, AE_BRIGHTNESS_MIN, AE_BRIGHTNESS_MAX)
        mean = rerender(b)
        iters += 1
        if mean <= 0:
            break
    return BrightnessFit(b, iters > 0, iters, mean, factor0)


def probe_genome(g: Genome, long_edge: int = AE_PROBE_LONG_EDGE,
                 quality: float = AE_PROBE_QUALITY) -> Genome:
    """applyPreset(genome, customSpec(long_edge, quality)) (src/presets.ts:46-79,
    :128-136): quality capped, oversample 1; a sized genome is rescaled so its
    long edge is `long_edge` (short edge floored), scale and the DE / filter
    radii scaled with it."""
    p = copy.deepcopy(g)
    p.quality = min(g.quality if g.quality is not None else quality, quality)
    p.oversample = 1
    if not g.size:
        return p
    w, h = g.size["width"], g.size["height"]
    mx = max(w, h)
    s = long_edge / mx
    p.size = {"width": long_edge if w == mx else max(1, math.floor(long_edge * w / h)),
              "height": long_edge if h == mx else max(1, math.floor(long_edge * h / w))}
    p.scale = g.scale * s
    if s != 1:
        if g.density is not None:
            p.density.max_rad = g.density.max_rad * s
            p.density.min_rad = g.density.min_rad * s
        if g.filter_radius is not None:
            p.filter_radius = g.filter_radius * s
    return p


def measure_probe_luminance(g: Genome, seed: int = AE_PROBE_SEED) -> float:
    """Mean luminance of the preview-tier probe render (its own renderer)."""
    from . import pack
    from .renderer import Renderer
    p = probe_genome(g)
    size = p.size or {"width": AE_PROBE_LONG_EDGE, "height": AE_PROBE_LONG_EDGE}
    radius = p.filter_radius if p.filter_radius is not None else pack.DEFAULT_FILTER_RADIUS
    r = Renderer(size["width"], size["height"], 1, filter_radius=radius)
    try:
        r.render(p, seed=seed)
        return compute_mean_luminance(r.read_image())
    finally:
        r.destroy()


def with_brightness(g: Genome, brightness: float) -> Genome:
    """The genome with its tonemap brightness replaced (a viewing transform;
    the source genome is not changed)."""
    out = copy.copy(g)
    tm = copy.copy(g.tonemap) if g.tonemap is not None else Tonemap()
    tm.brightness = brightness
    out.tonemap = tm
    return out
