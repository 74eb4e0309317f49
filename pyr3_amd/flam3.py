"""flam3 / Apophysis `.flam3` / `.flame` XML -> Genome.

Follows the reference importer (src/flame-import.ts:720-1085, `parseFlame`)
for the fields the render path consumes; what it cannot honour is listed in
the returned report instead of being dropped silently:

* `<xform>` / `<finalxform>`: `coefs` and `post` (flam3 column-major 3x2,
  :144-151; an identity post is dropped), `weight` (final: 0), `color`
  (missing -> position parity, :485), `color_speed` or the deprecated
  `symmetry` (colour speed (1 - s) / 2, :489-499), `opacity`, `chaos` (xaos
  row), every attribute naming a catalog variation (weight, in attribute
  order) with its `<name>_<param>` attributes (missing ones take the
  variation's defaults, :420-443); legacy aliases `Re_A`.. -> mobius,
  `oscope_*` -> oscilloscope (:43-67); unknown names are reported as
  dropped; an xform left with no variation gets linear 1 (:593-596);
  `motion_frequency`, `motion_function` (none/sin/triangle/hill), `animate`
  and `<motion>` children (partial xforms of deltas: missing coefs are zero,
  no linear fallback; :642-697), consumed by `pyr3_amd.animate`.
* `<flame>`: size, center, scale (default 50), rotate, quality, supersample
  (-> oversample), filter (+ shape), estimator_radius/minimum/curve,
  gamma/vibrancy/highlight_power/brightness/gamma_threshold (partial sets
  filled with flam3's defaults, :975-989), background, palette_mode,
  `<symmetry kind>`, palettes as `<color index rgb [a]>` (premultiplied,
  sparse ones interpolated around the wheel), `<colors data>` or
  `<palette count format>` hex (:240-372).
* Several `<flame>` elements (animation keyframes) are all parsed; the
  first one is the genome, the others come back as keyframes sorted by time
  (:741-758).

Not honoured (reported): flam3's numbered palette library (a `palette="N"`
reference falls back to the default palette), spatial filter shapes other
than gaussian, and the genome fields the reference lists as ignored
(:835-846).
"""
from __future__ import annotations

import math
import re
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from .catalog import VARIATION_DEFAULTS, VARIATION_INDEX, VARIATION_PARAMS
from .genome import (MAX_RAD_CAP, MAX_VARIATION_PARAMS, MAX_XFORMS, MIN_CURVE, MAX_CURVE, Affine,
                     ColorStop, Density, Genome, GenomeError, Palette, Tonemap, Variation, Xform)

# Alias attribute names accepted by flam3's parser (parser.c:1136-1152,
# 1228-1243), restated from src/flame-import.ts:43-67.
_EXACT_ALIASES = {f"{p}_{c}": f"mobius_{p.lower()}_{c.lower()}"
                  for p in ("Re", "Im") for c in "ABCD"}
_PREFIX_ALIASES = {"oscope": "oscilloscope"}

# <xform> attributes that are not variations (src/flame-import.ts:160-164).
_XFORM_FIELDS = frozenset({"weight", "color", "color_speed", "symmetry", "coefs", "post", "opacity",
                           "animate", "chaos", "plotmode", "var", "var1", "motion_frequency",
                           "motion_function"})

_PARAM_ATTRS = frozenset(f"{v}_{p}" for v, ps in VARIATION_PARAMS.items() for p in ps)

# flam3's defaults for a partially specified tone map (clear_cp; restated
# from src/flame-import.ts:980-986): gamma 4, brightness 4, vibrancy 1,
# highlight power -1, gamma threshold as the project default.
_FLAM3_TONEMAP = dict(gamma=4.0, brightness=4.0, vibrancy=1.0, highlight_power=-1.0,
                      gamma_threshold=Tonemap().gamma_threshold)

# Fallback palette when a flame carries no colours (src/palette.ts:150-157).
_DEFAULT_STOPS = [(0.0, 0.18, 0.0, 0.02), (0.18, 0.55, 0.04, 0.02), (0.4, 0.9, 0.18, 0.02),
                  (0.6, 1.0, 0.5, 0.06), (0.8, 1.0, 0.85, 0.18), (1.0, 1.0, 0.98, 0.7)]

# Genome attributes the render path does not model (src/flame-import.ts:835-846).
_IGNORED_SINGLE = ("zoom", "oversample", "temporal_filter_type", "temporal_filter_width",
                   "temporal_filter_exp", "hue", "rot_center")
_IGNORED_MULTI = ("zoom", "oversample", "hue", "rot_center")

_SPATIAL_SHAPES = ("gaussian", "hermite", "box", "triangle", "bell", "bspline", "mitchell",
                   "blackman", "catrom", "hanning", "hamming", "lanczos3", "lanczos2", "quadratic")


@dataclass
class ImportReport:
    flame_count: int = 1
    flame_name: str = ""
    dropped_variations: List[Dict] = field(default_factory=list)
    ignored_fields: List[Dict] = field(default_factory=list)
    defaulted_fields: List[Dict] = field(default_factory=list)
    palette_fallback: Optional[Dict] = None
    clamped_xforms: Optional[Dict] = None


@dataclass
class FlameImport:
    genome: Genome
    report: ImportReport
    keyframes: List[Genome] = field(default_factory=list)   # all flames, by time
    times: List[float] = field(default_factory=list)
    animation: Dict = field(default_factory=dict)            # interpolation settings


# ------------------------------------------------------------ scalars --

def _finite(raw: str, what: str) -> float:
    try:
        v = float(raw)
    except ValueError:
        v = math.nan
    if not math.isfinite(v):
        raise GenomeError(f"pyr3: {what} must be a finite number, got: {raw!r}")
    return v


def _number_or(raw: Optional[str], fallback: float) -> Tuple[float, bool]:
    """(value, ok): the parsed number, or the fallback and ok=False when the
    attribute is present but not a finite number."""
    if raw is None:
        return fallback, True
    try:
        v = float(raw)
    except ValueError:
        return fallback, False
    return (v, True) if math.isfinite(v) else (fallback, False)


def _numbers(raw: str) -> List[float]:
    out = []
    for tok in raw.split():
        try:
            out.append(float(tok))
        except ValueError:
            out.append(math.nan)
    return out


def coefs_to_affine(raw: str) -> Affine:
    """flam3 stores x' = c00 x + c10 y + c20, y' = c01 x + c11 y + c21 as
    "c00 c01 c10 c11 c20 c21"; the genome's rows are (a b c | d e f)."""
    v = _numbers(raw)
    if len(v) != 6 or not all(math.isfinite(x) for x in v):
        raise GenomeError(f"pyr3: coefs must be 6 finite numbers, got: {raw!r}")
    return Affine(a=v[0], b=v[2], c=v[4], d=v[1], e=v[3], f=v[5])


def _canonical(name: str) -> str:
    if name in _EXACT_ALIASES:
        return _EXACT_ALIASES[name]
    head, sep, tail = name.partition("_")
    if sep and head in _PREFIX_ALIASES:
        return _PREFIX_ALIASES[head] + "_" + tail
    return name


# ------------------------------------------------------------- xforms --

def _variation(name: str, weight: float, attrs: Dict[str, str]) -> Variation:
    v = Variation(VARIATION_INDEX[name], weight)
    names = VARIATION_PARAMS.get(name, [])
    defaults = VARIATION_DEFAULTS.get(name, [])
    for i, pname in enumerate(names[:MAX_VARIATION_PARAMS]):
        raw = attrs.get(f"{name}_{pname}")
        if raw is not None:
            v.params[i] = _finite(raw, f"{name}_{pname}")
        elif i < len(defaults):
            v.params[i] = float(defaults[i])
    return v


def _motion_func(text: str) -> Optional[int]:
    """flam3 motion_function names (src/flame-import.ts:700-707)."""
    t = text.strip().lower()
    if t in ("", "none", "0"):
        return 0
    if t in ("sin", "sine", "1"):
        return 1
    if t in ("triangle", "2"):
        return 2
    if t in ("hill", "3"):
        return 3
    return None


def _xform(el: ET.Element, index: int, final: bool, parity: int, rep: ImportReport,
           in_motion: bool = False) -> Xform:
    """One <xform>/<finalxform> (src/flame-import.ts:452-697). A <motion>
    child (in_motion) is a partial xform of additive deltas: missing coefs
    are zero deltas, and no linear fallback is added."""
    where = "finalxform" if final else f"xform[{index}]"

    def number(attr: str, fallback: float) -> float:
        val, ok = _number_or(el.get(attr), fallback)
        if not ok:
            rep.ignored_fields.append({"field": f"{attr}@{where}", "value": el.get(attr)})
        return val

    weight = 0.0 if final else number("weight", 1.0)
    color = number("color", float(parity & 1))
    if el.get("color_speed") is not None:
        speed = number("color_speed", 0.0)
    elif el.get("symmetry") is not None:
        speed = (1.0 - number("symmetry", 1.0)) / 2.0
    else:
        speed = 0.5
    if el.get("coefs") is None:
        if not in_motion:
            raise GenomeError(f"pyr3: {where} missing coefs")
        affine = Affine(0.0, 0.0, 0.0, 0.0, 0.0, 0.0)
    else:
        affine = coefs_to_affine(el.get("coefs"))

    attrs: Dict[str, str] = {}
    for k, val in el.attrib.items():
        attrs[_canonical(k)] = val
    variations: List[Variation] = []
    for name, val in attrs.items():
        if name in _XFORM_FIELDS:
            continue
        if name in VARIATION_INDEX:
            variations.append(_variation(name, _finite(val, name), attrs))
            continue
        if "_" in name:
            if name in _PARAM_ATTRS:
                continue
            w, _ = _number_or(val, 0.0)
        else:
            w = _finite(val, name)
        drop = {"name": name, "weight": w, "xform": -1 if final else index}
        if final:
            drop["final"] = True
        rep.dropped_variations.append(drop)
    if not variations and not in_motion:
        variations.append(Variation(VARIATION_INDEX["linear"], 1.0))

    x = Xform(weight=weight, color=color, color_speed=speed, affine=affine,
              variations=variations)
    if el.get("post") is not None:
        post = coefs_to_affine(el.get("post"))
        if post != Affine():
            x.post = post
    if el.get("opacity") is not None:
        x.opacity = _finite(el.get("opacity"), f"opacity@{where}")
        x.opacity_given = True
    if el.get("chaos") is not None:
        row = [_finite(t, f"chaos@{where}[{i}]") for i, t in enumerate(el.get("chaos").split())]
        x.xaos = row or None
    # keyframe-animation fields (src/flame-import.ts:642-690)
    if el.get("motion_frequency") is not None:
        v = number("motion_frequency", 0.0)
        if v != 0:
            x.motion_freq = v
    if el.get("motion_function") is not None:
        code = _motion_func(el.get("motion_function"))
        if code is None:
            rep.ignored_fields.append({"field": f"motion_function@{where}",
                                       "value": el.get("motion_function")})
        elif code != 0:
            x.motion_func = code
    if el.get("animate") is not None:
        v = number("animate", 1.0)
        if v != 1:
            x.animate = v
    if not in_motion:
        motion = []
        for mi, me in enumerate(c for c in el if c.tag == "motion"):
            sub_rep = ImportReport()
            motion.append(_xform(me, index, False, 0, sub_rep, in_motion=True))
            for d in sub_rep.dropped_variations:
                rep.dropped_variations.append(dict(d, name=f"motion[{mi}].{d['name']}"))
            for ig in sub_rep.ignored_fields:
                rep.ignored_fields.append({"field": f"motion[{mi}].{ig['field']}",
                                           "value": ig["value"]})
        if motion:
            x.motion = motion
    return x


# ------------------------------------------------------------ palette --

def _hex_bytes(text: str) -> List[int]:
    s = re.sub(r"\s+", "", text or "")
    if len(s) % 2 or re.fullmatch(r"[0-9a-fA-F]*", s) is None:
        raise GenomeError(f"pyr3: palette hex data malformed: {s[:32]!r}...")
    return [int(s[i:i + 2], 16) for i in range(0, len(s), 2)]


def _palette(flame: ET.Element, name: str, rep: ImportReport) -> Palette:
    rgb = [(0.0, 0.0, 0.0)] * 256
    found = False
    colors = flame.findall("color")
    if colors:
        found = True
        given = set()
        for el in colors:
            if el.get("index") is None:
                raise GenomeError("pyr3: <color> missing index")
            idx = _finite(el.get("index"), "color index")
            if idx != int(idx) or not 0 <= idx <= 255:
                raise GenomeError(f'pyr3: <color index="{el.get("index")}"> must be 0-255')
            idx = int(idx)
            triple = el.get("rgb") if el.get("rgb") is not None else el.get("rgba")
            if triple is None:
                raise GenomeError(f'pyr3: <color index="{idx}"> missing rgb/rgba')
            parts = _numbers(triple)
            if len(parts) < 3 or not all(math.isfinite(p) for p in parts):
                raise GenomeError(f'pyr3: <color index="{idx}"> bad rgb: {triple!r}')
            alpha = 1.0
            if el.get("a") is not None:
                alpha = _finite(el.get("a"), "a")
            elif len(parts) >= 4:
                alpha = parts[3]
            if alpha > 1.0:
                alpha /= 255.0
            rgb[idx] = tuple(parts[c] / 255.0 * alpha for c in range(3))
            given.add(idx)
        if 1 < len(given) < 256:   # a single given entry leaves the rest black
            # fill each gap linearly between the given entries around it,
            # wrapping from the last given index back to the first
            order = sorted(given)
            for a, b in zip(order, order[1:] + [order[0] + 256]):
                span = b - a
                ca, cb = rgb[a], rgb[b % 256]
                for j in range(1, span):
                    t = j / span
                    rgb[(a + j) % 256] = tuple(ca[c] * (1 - t) + cb[c] * t for c in range(3))
    block = flame.find("colors")
    if block is not None:
        found = True
        data = _hex_bytes(block.get("data", ""))
        if len(data) < 1024:
            raise GenomeError(f"pyr3: <colors data> needs >= 1024 hex bytes, got {len(data)}")
        rgb = [(data[4 * i + 1] / 255.0, data[4 * i + 2] / 255.0, data[4 * i + 3] / 255.0)
               for i in range(256)]
    pal = flame.find("palette")
    if pal is not None:
        if pal.get("index0") is not None or pal.get("blend") is not None:
            raise GenomeError("pyr3: old palette format (<palette index0=... blend=...>) "
                              "not supported")
        found = True
        step = 4 if pal.get("format", "RGB") == "RGBA" else 3
        data = _hex_bytes(pal.text or "")
        for i in range(min(256, len(data) // step)):
            rgb[i] = tuple(data[i * step + c] / 255.0 for c in range(3))
    if found:
        return Palette(name=name, stops=[ColorStop(i / 255.0, *rgb[i]) for i in range(256)])
    ref = flame.get("palette")
    reason = ("flam3's numbered palette library is not bundled" if ref and ref.strip()
              else "no <color>/<colors>/<palette> block and no palette index")
    rep.palette_fallback = {"kind": "default", "reason": reason,
                            **({"index": ref} if ref and ref.strip() else {})}
    return Palette(name="pyre", stops=[ColorStop(*s) for s in _DEFAULT_STOPS])


# ------------------------------------------------------------- flames --

def _flame(flame: ET.Element, n_flames: int, rep: ImportReport) -> Tuple[Genome, float]:
    xforms: List[Xform] = []
    final = None
    for el in flame:
        if el.tag == "xform":
            xforms.append(_xform(el, len(xforms), False, len(xforms), rep))
        elif el.tag == "finalxform":
            final = _xform(el, -1, True, len(xforms), rep)
    if not xforms:
        raise GenomeError("pyr3: <flame> has no <xform> children; cannot render")
    if len(xforms) > MAX_XFORMS:
        rep.clamped_xforms = {"had": len(xforms), "cap": MAX_XFORMS}
        xforms = xforms[:MAX_XFORMS]
    for attr in (_IGNORED_MULTI if n_flames > 1 else _IGNORED_SINGLE):
        if flame.get(attr) is not None:
            rep.ignored_fields.append({"field": attr, "value": flame.get(attr)})

    name = flame.get("name", "")
    scale, ok = _number_or(flame.get("scale"), 50.0)
    if not ok:
        rep.defaulted_fields.append({"field": "scale", "value": flame.get("scale"),
                                     "used": "50"})
    cx = cy = 0.0
    if flame.get("center") is not None:
        c = _numbers(flame.get("center"))
        if len(c) == 2 and all(math.isfinite(v) for v in c):
            cx, cy = c
        else:
            rep.defaulted_fields.append({"field": "center", "value": flame.get("center"),
                                         "used": "0 0"})
    g = Genome(name=name, xforms=xforms, scale=scale, cx=cx, cy=cy,
               palette=_palette(flame, name, rep), finalxform=final)

    sym = flame.find("symmetry")
    if sym is not None:
        if sym.get("kind") is None:
            raise GenomeError("pyr3: <symmetry> missing kind")
        k = _finite(sym.get("kind"), "symmetry kind")
        if k != int(k) or k == 0:
            raise GenomeError(f'pyr3: <symmetry kind="{sym.get("kind")}"> must be a non-zero '
                              f"integer")
        g.symmetry = {"kind": "rotational" if k > 0 else "dihedral", "n": abs(int(k))}

    est = [flame.get(a) for a in ("estimator_radius", "estimator_minimum", "estimator_curve")]
    if any(v is not None for v in est):
        mx = _finite(est[0], "estimator_radius") if est[0] is not None else 9.0
        mn = _finite(est[1], "estimator_minimum") if est[1] is not None else 0.0
        cv = _finite(est[2], "estimator_curve") if est[2] is not None else 0.4
        if not 0 <= mx <= MAX_RAD_CAP:
            raise GenomeError(f"pyr3: estimator_radius out of range [0, {MAX_RAD_CAP}]: {mx}")
        if not 0 <= mn <= mx:
            raise GenomeError(f"pyr3: estimator_minimum must be in [0, estimator_radius]: {mn}")
        if not MIN_CURVE <= cv <= MAX_CURVE:
            raise GenomeError(f"pyr3: estimator_curve out of range [{MIN_CURVE}, {MAX_CURVE}]: "
                              f"{cv}")
        g.density = Density(mx, mn, cv)

    tone = {}
    for attr, key in (("gamma", "gamma"), ("vibrancy", "vibrancy"),
                      ("highlight_power", "highlight_power"), ("brightness", "brightness"),
                      ("gamma_threshold", "gamma_threshold")):
        if flame.get(attr) is not None:
            tone[key] = _finite(flame.get(attr), attr)
    if tone:
        g.tonemap = Tonemap(**{**_FLAM3_TONEMAP, **tone})

    if flame.get("rotate") is not None:
        g.rotate = _finite(flame.get("rotate"), "rotate")
    if flame.get("quality") is not None:
        q = _finite(flame.get("quality"), "quality")
        if q > 0:
            g.quality = q
    if flame.get("supersample") is not None:
        s = math.floor(_finite(flame.get("supersample"), "supersample"))
        if s > 1:
            g.oversample = int(s)
    if flame.get("size") is not None:
        v = _numbers(flame.get("size"))
        if len(v) == 2 and all(math.isfinite(x) and x == int(x) and x > 0 for x in v):
            g.size = {"width": int(v[0]), "height": int(v[1])}
        else:
            rep.defaulted_fields.append({"field": "size", "value": flame.get("size"),
                                         "used": "viewer canvas default"})
    if flame.get("filter") is not None:
        radius = _finite(flame.get("filter"), "filter")
        if radius > 0:
            shape = flame.get("filter_shape", "gaussian")
            if shape != "gaussian":
                known = "not supported by this renderer" if shape in _SPATIAL_SHAPES else "unknown"
                rep.ignored_fields.append({"field": "filter_shape",
                                           "value": f"{shape} ({known}; gaussian used)"})
            g.filter_radius = radius
            if g.density is None:
                rep.ignored_fields.append({"field": "filter",
                                           "value": f"{radius} (requires <flame estimator_*>)"})
    if flame.get("background") is not None:
        bg = _numbers(flame.get("background"))
        if len(bg) != 3 or not all(math.isfinite(v) for v in bg):
            raise GenomeError(f"pyr3: background must be 3 finite numbers, got: "
                              f"{flame.get('background')!r}")
        if any(bg):
            g.background = bg
    if flame.get("palette_mode") is not None:
        pm = flame.get("palette_mode")
        if pm in ("step", "linear"):
            g.palette_mode = pm
        else:
            rep.ignored_fields.append({"field": "palette_mode", "value": pm})
    t, ok = _number_or(flame.get("time"), 0.0)
    if not ok:
        rep.defaulted_fields.append({"field": "time", "value": flame.get("time"), "used": "0"})
    if t != 0:
        g.time = t
    return g, t


def _animation(flame: ET.Element) -> Dict:
    a = {}
    choices = {"interpolation": ("linear", "smooth"),
               "interpolation_type": ("linear", "log", "compat", "older"),
               "palette_interpolation": ("hsv", "sweep", "rgb", "hsv_circular"),
               "temporal_filter_type": ("box", "gaussian", "exp")}
    for attr, ok in choices.items():
        if flame.get(attr) in ok:
            a[attr] = flame.get(attr)
    for attr, positive in (("hsv_rgb_palette_blend", False), ("temporal_samples", True),
                           ("temporal_filter_width", True), ("temporal_filter_exp", False)):
        v, good = _number_or(flame.get(attr), math.nan)
        if flame.get(attr) is not None and good and (v > 0 or not positive):
            a[attr] = v
    return a


def parse_flame(text: str) -> FlameImport:
    """Parse `.flam3` / `.flame` XML text (one or more <flame> elements)."""
    body = text.strip()
    if body.startswith("<?xml"):
        end = body.find("?>")
        if end > 0:
            body = body[end + 2:].strip()
    try:
        root = ET.fromstring(f"<flames>{body}</flames>")
    except ET.ParseError as e:
        raise GenomeError(f"pyr3: malformed .flame XML: {e}") from None
    flames = root.findall(".//flame")
    if not flames:
        raise GenomeError("pyr3: no <flame> element found")
    rep = ImportReport(flame_count=len(flames))
    parsed = []
    for i, f in enumerate(flames):
        r = rep if i == 0 else ImportReport(flame_count=len(flames))
        parsed.append(_flame(f, len(flames), r))
    genome = parsed[0][0]
    rep.flame_name = genome.name
    out = FlameImport(genome=genome, report=rep)
    if len(parsed) > 1:
        order = sorted(range(len(parsed)), key=lambda k: parsed[k][1])
        out.keyframes = [parsed[k][0] for k in order]
        out.times = [parsed[k][1] for k in order]
        out.animation = _animation(flames[0])
    return out


def load_genome(path: str, text: Optional[str] = None) -> Tuple[Genome, str, Optional[ImportReport]]:
    """Genome from a `.flam3` / `.flame` / `.pyr3.json` file or a `.png` saved
    with its genome (src/loader.ts:32-66 sniffing: suffix, then a leading '<'
    means XML). Returns (genome, kind, report)."""
    from .genome import genome_from_text
    if path.lower().endswith(".png"):
        from .png import genome_from_png
        with open(path, "rb") as fh:
            return genome_from_png(fh.read()), "pyr3-png", None
    if text is None:
        with open(path, encoding="utf-8") as fh:
            text = fh.read()
    low = path.lower()
    if low.endswith((".flame", ".flam3")) or (not low.endswith(".json")
                                               and text.lstrip().startswith("<")):
        imp = parse_flame(text)
        return imp.genome, "flame", imp.report
    return genome_from_text(text), "pyr3-json", None
