"""Genome model and the ``.pyr3.json`` (schema v1) codec.

Mirrors the reference's genome types (src/genome.ts:80-168) and its JSON
reader/writer (src/serialize.ts:22-120 schema, :1040 genomeFromJson,
:1304-1404 parseXformBody/variationFromJson, :878 genomeToJson). Errors are
raised as ``GenomeError`` with the same "pyr3: ..." messages the reference
throws, so a host that caught the reference's exceptions reads the same text.
"""
from __future__ import annotations

import copy
import json
import math
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

from .catalog import VARIATION_INDEX, VARIATION_NAMES, VARIATION_PARAMS

PYR3_JSON_VERSION = 1
MAX_XFORMS = 128            # src/genome.ts:326
MAX_VARIATION_PARAMS = 10   # src/serialize.ts:861-876
MAX_RAD_CAP = 30            # src/density.ts:20
MIN_CURVE, MAX_CURVE = 0.1, 2.0


class GenomeError(ValueError):
    """Invalid genome input (the reference throws ``Error('pyr3: ...')``)."""


@dataclass
class Variation:
    index: int
    weight: float
    params: List[float] = field(default_factory=lambda: [0.0] * MAX_VARIATION_PARAMS)
    active: bool = True

    @property
    def name(self) -> str:
        return VARIATION_NAMES.get(self.index, f"var{self.index}")


@dataclass
class Affine:
    a: float = 1.0
    b: float = 0.0
    c: float = 0.0
    d: float = 0.0
    e: float = 1.0
    f: float = 0.0


@dataclass
class Xform:
    weight: float
    color: float
    color_speed: float
    affine: Affine
    variations: List[Variation]
    opacity: float = 1.0
    xaos: Optional[List[float]] = None
    post: Optional[Affine] = None
    active: bool = True
    # keyframe-animation fields (src/genome.ts:61-77): flam3 motion_frequency /
    # motion_function (0 none, 1 sin, 2 triangle, 3 hill), the per-xform
    # motion overlays (partial xforms of additive deltas) and `animate`
    # (1 rotating, 0 stationary; None = 1)
    motion_freq: Optional[float] = None
    motion_func: Optional[int] = None
    motion: Optional[List["Xform"]] = None
    animate: Optional[float] = None
    # the source gave an opacity (a motion overlay adds its opacity only then)
    opacity_given: bool = field(default=False, compare=False, repr=False)


@dataclass
class ColorStop:
    t: float
    r: float
    g: float
    b: float


@dataclass
class Palette:
    name: str
    stops: List[ColorStop]
    hue: float = 0.0
    mode: str = "linear"


@dataclass
class Density:
    max_rad: float = 9.0
    min_rad: float = 0.0
    curve: float = 0.4


@dataclass
class Tonemap:
    """src/tonemap.ts:10-37 (DEFAULT_TONEMAP)."""
    gamma: float = 2.4
    vibrancy: float = 0.0
    highlight_power: float = 1.0
    brightness: float = 1.0
    gamma_threshold: float = 0.01


@dataclass
class Genome:
    name: str
    xforms: List[Xform]
    scale: float
    cx: float
    cy: float
    palette: Palette
    finalxform: Optional[Xform] = None
    symmetry: Optional[Dict] = None
    density: Optional[Density] = None
    tonemap: Optional[Tonemap] = None
    rotate: float = 0.0
    xform_blend: float = 0.0
    quality: Optional[float] = None
    oversample: Optional[int] = None
    size: Optional[Dict[str, int]] = None
    filter_radius: Optional[float] = None
    background: Optional[List[float]] = None
    palette_mode: str = "step"
    nick: Optional[str] = None
    # post-tonemap grading (src/genome.ts:160-163): channel curves per channel
    # ('composite', 'r', 'g', 'b', 'luma' -> [(x, y), ...]) and HSV adjust
    # {'hue': deg, 'sat': %, 'light': %}
    channel_curves: Optional[Dict[str, List[Tuple[float, float]]]] = None
    hsl_adjust: Optional[Dict[str, float]] = None
    time: Optional[float] = None    # keyframe time (src/genome.ts:167)

    def clone(self) -> "Genome":
        return copy.deepcopy(self)


# ---------------------------------------------------------------- reading --

def _obj(v, path):
    if not isinstance(v, dict):
        raise GenomeError(f"pyr3: expected object at {path}, got: {_typeof(v)}")
    return v


def _arr(v, path):
    if not isinstance(v, list):
        raise GenomeError(f"pyr3: expected array at {path}, got: {_typeof(v)}")
    return v


def _str(v, path):
    if not isinstance(v, str):
        raise GenomeError(f"pyr3: expected string at {path}, got: {_typeof(v)}")
    return v


def _num(v, path):
    if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v):
        raise GenomeError(f"pyr3: expected finite number at {path}, got: {_typeof(v)}")
    return float(v)


def _typeof(v):
    if v is None:
        return "null"
    if isinstance(v, list):
        return "array"
    if isinstance(v, bool):
        return "boolean"
    if isinstance(v, (int, float)):
        return "number"
    if isinstance(v, str):
        return "string"
    return "object"


def _affine(o, path):
    o = _obj(o, path)
    return Affine(*(_num(o.get(k), f"{path}.{k}") for k in "abcdef"))


def _variation(o, path) -> Variation:
    o = _obj(o, path)
    name = _str(o.get("name"), f"{path}.name")
    weight = _num(o.get("weight"), f"{path}.weight")
    if name not in VARIATION_INDEX:
        raise GenomeError(f"pyr3: unknown variation name '{name}' at {path}")
    v = Variation(index=VARIATION_INDEX[name], weight=weight)
    names = VARIATION_PARAMS.get(name)
    raw = o.get("params")
    if raw is not None and names:
        raw = _obj(raw, f"{path}.params")
        for i, pn in enumerate(names[:MAX_VARIATION_PARAMS]):
            if raw.get(pn) is not None:
                v.params[i] = _num(raw[pn], f"{path}.params.{pn}")
    if o.get("active") is False:
        v.active = False
    return v


def _xform(o, path, is_final) -> Xform:
    o = _obj(o, path)
    weight = 0.0 if is_final else _num(o.get("weight"), f"{path}.weight")
    x = Xform(
        weight=weight,
        color=_num(o.get("color"), f"{path}.color"),
        color_speed=_num(o.get("colorSpeed"), f"{path}.colorSpeed"),
        affine=_affine(o.get("affine"), f"{path}.affine"),
        variations=[_variation(v, f"{path}.variations[{i}]")
                    for i, v in enumerate(_arr(o.get("variations"), f"{path}.variations"))],
    )
    if o.get("opacity") is not None:
        x.opacity = _num(o["opacity"], f"{path}.opacity")
        x.opacity_given = True
    if not is_final and o.get("xaos") is not None:
        x.xaos = [_num(v, f"{path}.xaos[{i}]") for i, v in enumerate(_arr(o["xaos"], f"{path}.xaos"))]
    if o.get("post") is not None:
        x.post = _affine(o["post"], f"{path}.post")
    if o.get("active") is False:
        x.active = False
    # src/serialize.ts:1341-1367
    if o.get("motion_freq") is not None:
        x.motion_freq = _num(o["motion_freq"], f"{path}.motion_freq")
    if o.get("motion_func") is not None:
        mf = _num(o["motion_func"], f"{path}.motion_func")
        if mf not in (0, 1, 2, 3):
            raise GenomeError(f"pyr3: {path}.motion_func must be 0..3, got: {o['motion_func']}")
        x.motion_func = int(mf)
    if o.get("animate") is not None:
        x.animate = _num(o["animate"], f"{path}.animate")
    if o.get("motion") is not None:
        x.motion = [_xform(m, f"{path}.motion[{i}]", False)
                    for i, m in enumerate(_arr(o["motion"], f"{path}.motion"))]
    return x


CURVE_CHANNELS = ("composite", "r", "g", "b", "luma")


def _channel_curves(j, path):
    """src/serialize.ts:1271-1297 parseChannelCurves."""
    o = _obj(j, path)
    out = {}
    for ch in CURVE_CHANNELS:
        arr = _arr(o.get(ch), f"{path}.{ch}")
        if len(arr) < 2 or len(arr) > 8:
            raise GenomeError(f"pyr3: {path}.{ch} must have 2..8 points, got {len(arr)}")
        pts = []
        for i, pt in enumerate(arr):
            po = _obj(pt, f"{path}.{ch}[{i}]")
            x = _num(po.get("x"), f"{path}.{ch}[{i}].x")
            y = _num(po.get("y"), f"{path}.{ch}[{i}].y")
            if x < 0 or x > 1 or y < 0 or y > 1:
                raise GenomeError(f"pyr3: {path}.{ch}[{i}] out of [0,1]: x={x:g} y={y:g}")
            pts.append((x, y))
        for i in range(1, len(pts)):
            if pts[i][0] <= pts[i - 1][0]:
                raise GenomeError(f"pyr3: {path}.{ch}[{i}].x not strictly monotonic")
        out[ch] = pts
    return out


def _hsl_adjust(j, path):
    """src/serialize.ts:1239-1257: hue wrapped into [-180, 180], sat clamped
    to [0, 200], light to [-100, 100]; missing or non-numeric fields take
    the identity values."""
    o = _obj(j, path)

    def num(v, dflt):
        if isinstance(v, bool) or not isinstance(v, (int, float)) or not math.isfinite(v):
            return dflt
        return float(v)
    raw_hue = num(o.get("hue"), 0.0)
    hue_mod = math.fmod(math.fmod(raw_hue, 360.0) + 360.0, 360.0)
    hue = hue_mod - 360.0 if hue_mod > 180.0 else hue_mod
    sat = min(200.0, max(0.0, num(o.get("sat"), 100.0)))
    light = min(100.0, max(-100.0, num(o.get("light"), 0.0)))
    return {"hue": hue, "sat": sat, "light": light}


def genome_from_json(j) -> Genome:
    """src/serialize.ts:1040 genomeFromJson (render-relevant fields)."""
    root = _obj(j, "root")
    if root.get("version") != PYR3_JSON_VERSION:
        raise GenomeError(f"pyr3: unsupported .pyr3.json version: {root.get('version')} "
                          f"(expected {PYR3_JSON_VERSION})")
    name = _str(root.get("name"), "name")
    vp = _obj(root.get("viewport"), "viewport")
    scale = _num(vp.get("scale"), "viewport.scale")
    if scale == 0:
        raise GenomeError("pyr3: viewport.scale must be non-zero (0 collapses the image to one pixel)")
    cx = _num(vp.get("cx"), "viewport.cx")
    cy = _num(vp.get("cy"), "viewport.cy")
    po = _obj(root.get("palette"), "palette")
    stops = []
    for i, s in enumerate(_arr(po.get("stops"), "palette.stops")):
        so = _obj(s, f"palette.stops[{i}]")
        stops.append(ColorStop(*(_num(so.get(k), f"palette.stops[{i}].{k}") for k in "trgb")))
    pal = Palette(name=_str(po.get("name"), "palette.name"), stops=stops)
    if po.get("hue") is not None:
        pal.hue = _num(po["hue"], "palette.hue")
    if po.get("mode") is not None:
        if po["mode"] not in ("linear", "step", "smooth"):
            raise GenomeError(f"pyr3: palette.mode must be 'linear', 'step', or 'smooth', got: {po['mode']}")
        pal.mode = po["mode"]
    xr = _arr(root.get("xforms"), "xforms")
    if not xr:
        raise GenomeError("pyr3: xforms must contain at least one xform; cannot render")
    if len(xr) > MAX_XFORMS:
        raise GenomeError(f"pyr3: xforms count {len(xr)} exceeds the {MAX_XFORMS}-xform cap")
    g = Genome(name=name, xforms=[_xform(x, f"xforms[{i}]", False) for i, x in enumerate(xr)],
               scale=scale, cx=cx, cy=cy, palette=pal)
    if root.get("finalxform") is not None:
        g.finalxform = _xform(root["finalxform"], "finalxform", True)
    if root.get("symmetry") is not None:
        s = _obj(root["symmetry"], "symmetry")
        kind = _str(s.get("kind"), "symmetry.kind")
        if kind not in ("rotational", "dihedral"):
            raise GenomeError(f"pyr3: symmetry.kind must be 'rotational' or 'dihedral', got: {kind}")
        n = _num(s.get("n"), "symmetry.n")
        if n != int(n) or n < 1:
            raise GenomeError(f"pyr3: symmetry.n must be a positive integer, got: {n}")
        g.symmetry = {"kind": kind, "n": int(n)}
    if root.get("density") is not None:
        d = _obj(root["density"], "density")
        mx = _num(d.get("maxRad"), "density.maxRad")
        mn = _num(d.get("minRad"), "density.minRad")
        cv = _num(d.get("curve"), "density.curve")
        if mx < 0 or mx > MAX_RAD_CAP:
            raise GenomeError(f"pyr3: density.maxRad must be in [0, {MAX_RAD_CAP}], got: {mx}")
        if mn < 0 or mn > mx:
            raise GenomeError(f"pyr3: density.minRad must be in [0, density.maxRad], got: {mn} (max={mx})")
        if cv < MIN_CURVE or cv > MAX_CURVE:
            raise GenomeError(f"pyr3: density.curve must be in [{MIN_CURVE}, {MAX_CURVE}], got: {cv}")
        g.density = Density(mx, mn, cv)
    if root.get("tonemap") is not None:
        t = _obj(root["tonemap"], "tonemap")
        tm = Tonemap()
        for js, attr in (("gamma", "gamma"), ("vibrancy", "vibrancy"),
                         ("highlightPower", "highlight_power"), ("brightness", "brightness"),
                         ("gammaThreshold", "gamma_threshold")):
            if t.get(js) is not None:
                setattr(tm, attr, _num(t[js], f"tonemap.{js}"))
        g.tonemap = tm
    if root.get("rotate") is not None:
        g.rotate = _num(root["rotate"], "rotate")
    if root.get("xformBlend") is not None:
        g.xform_blend = _num(root["xformBlend"], "xformBlend")
    if root.get("quality") is not None:
        g.quality = _num(root["quality"], "quality")
    if root.get("oversample") is not None:
        g.oversample = int(_num(root["oversample"], "oversample"))
    if root.get("size") is not None:
        so = _obj(root["size"], "size")
        g.size = {"width": int(_num(so.get("width"), "size.width")),
                  "height": int(_num(so.get("height"), "size.height"))}
    if root.get("spatialFilter") is not None:
        sf = _obj(root["spatialFilter"], "spatialFilter")
        g.filter_radius = _num(sf.get("radius"), "spatialFilter.radius")
    if root.get("background") is not None:
        bg = _arr(root["background"], "background")
        g.background = [_num(v, f"background[{i}]") for i, v in enumerate(bg)]
    if root.get("paletteMode") is not None:
        g.palette_mode = _str(root["paletteMode"], "paletteMode")
    if root.get("channelCurves") is not None:
        g.channel_curves = _channel_curves(root["channelCurves"], "channelCurves")
    if root.get("hslAdjust") is not None:
        g.hsl_adjust = _hsl_adjust(root["hslAdjust"], "hslAdjust")
    if root.get("nick") is not None:
        g.nick = _str(root["nick"], "nick")
    if root.get("time") is not None:
        t = _num(root["time"], "time")
        if t != 0:
            g.time = t
    return g


def genome_from_text(text: str) -> Genome:
    return genome_from_json(json.loads(text))


# ---------------------------------------------------------------- writing --

def _affine_json(a: Affine):
    return {"a": a.a, "b": a.b, "c": a.c, "d": a.d, "e": a.e, "f": a.f}


def _xform_json(x: Xform, is_final: bool):
    o = {}
    if not is_final:
        o["weight"] = x.weight
    o.update({"color": x.color, "colorSpeed": x.color_speed, "affine": _affine_json(x.affine)})
    vs = []
    for v in x.variations:
        vo = {"name": v.name, "weight": v.weight}
        names = VARIATION_PARAMS.get(v.name, [])
        if names:
            vo["params"] = {pn: v.params[i] for i, pn in enumerate(names[:MAX_VARIATION_PARAMS])}
        if not v.active:
            vo["active"] = False
        vs.append(vo)
    o["variations"] = vs
    if x.opacity != 1.0:
        o["opacity"] = x.opacity
    if x.xaos is not None and not is_final:
        o["xaos"] = list(x.xaos)
    if x.post is not None:
        o["post"] = _affine_json(x.post)
    if not x.active:
        o["active"] = False
    if x.motion_freq:
        o["motion_freq"] = x.motion_freq
    if x.motion_func:
        o["motion_func"] = x.motion_func
    if x.animate is not None and x.animate != 1:
        o["animate"] = x.animate
    if x.motion:
        o["motion"] = [_xform_json(m, False) for m in x.motion]
    return o


def genome_to_json(g: Genome) -> dict:
    """src/serialize.ts:878 genomeToJson (render-relevant fields)."""
    o = {
        "version": PYR3_JSON_VERSION,
        "name": g.name,
        "viewport": {"scale": g.scale, "cx": g.cx, "cy": g.cy},
        "palette": {"name": g.palette.name,
                    "stops": [{"t": s.t, "r": s.r, "g": s.g, "b": s.b} for s in g.palette.stops]},
        "xforms": [_xform_json(x, False) for x in g.xforms],
    }
    if g.palette.hue:
        o["palette"]["hue"] = g.palette.hue
    if g.palette.mode != "linear":
        o["palette"]["mode"] = g.palette.mode
    if g.finalxform is not None:
        o["finalxform"] = _xform_json(g.finalxform, True)
    if g.symmetry:
        o["symmetry"] = dict(g.symmetry)
    if g.density is not None:
        o["density"] = {"maxRad": g.density.max_rad, "minRad": g.density.min_rad,
                        "curve": g.density.curve}
    if g.tonemap is not None:
        t = g.tonemap
        o["tonemap"] = {"gamma": t.gamma, "vibrancy": t.vibrancy,
                        "highlightPower": t.highlight_power, "brightness": t.brightness,
                        "gammaThreshold": t.gamma_threshold}
    if g.rotate:
        o["rotate"] = g.rotate
    if g.xform_blend:
        o["xformBlend"] = g.xform_blend
    if g.quality is not None:
        o["quality"] = g.quality
    if g.oversample is not None and g.oversample != 1:
        o["oversample"] = g.oversample
    if g.size is not None:
        o["size"] = dict(g.size)
    if g.filter_radius is not None:
        o["spatialFilter"] = {"shape": "gaussian", "radius": g.filter_radius}
    if g.background is not None:
        o["background"] = list(g.background)
    if g.palette_mode != "step":
        o["paletteMode"] = g.palette_mode
    if g.channel_curves is not None and any(
            [tuple(p) for p in g.channel_curves[ch]] != [(0.0, 0.0), (1.0, 1.0)]
            for ch in CURVE_CHANNELS):
        o["channelCurves"] = {ch: [{"x": x, "y": y} for x, y in g.channel_curves[ch]]
                              for ch in CURVE_CHANNELS}
    h = g.hsl_adjust
    if h is not None and (h["hue"] != 0 or h["sat"] != 100 or h["light"] != 0):
        o["hslAdjust"] = {"hue": h["hue"], "sat": h["sat"], "light": h["light"]}
    if g.nick:
        o["nick"] = g.nick
    if g.time:
        o["time"] = g.time
    return o


def genome_to_text(g: Genome) -> str:
    return json.dumps(genome_to_json(g), indent=2)
