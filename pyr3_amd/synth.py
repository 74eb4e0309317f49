"""Seeded synthetic genomes for the benchmark workloads.

Each workload draws its genomes from numpy's PCG64 seeded with
(workload index * 1000 + k), so every figure is reproducible:

* preview-small (256x256, os 1, q50, fuse 20): 3 xforms, weights U(0.2, 1),
  affine coefficients U(-1, 1), variations {linear, sinusoidal, spherical}
  with Dirichlet(1) weights, a random 256-stop palette (t = i/255), colour
  speed 0.5, DE (9, 0, 0.4), gamma 4, vibrancy 1.
* preview (1024x768, os 2, q200): 6 xforms with 2-4 variations drawn from
  the flam3 core set, parametric variations' params U(-2, 2), post-affine on
  half of the xforms, same DE/tonemap settings.
* 4k (3840x2160, os 2, q200): 12 xforms plus a final xform, variations from
  the whole catalog (V0..V325, direct-colour entries included), log-normal
  (sigma 0.7) affine scales; entries past the flam3 core set take the
  reference's default parameters (src/serialize.ts:487-860).
* plain-1080p (1920x1080, os 1): the 12-xform + final layout of 4k with
  variations from the flam3 core set, the denoiser/benchmark workload.

The viewport is centred on the origin with 4 world units across the long
edge (scale = long_edge / 4).
"""
from __future__ import annotations

from typing import Iterable, List, Sequence

import numpy as np

from .catalog import VARIATION_DEFAULTS, VARIATION_INDEX, VARIATION_PARAMS
from .genome import Affine, ColorStop, Density, Genome, Palette, Tonemap, Variation, Xform

CONFIG0_VARIATIONS = ("linear", "sinusoidal", "spherical")


def random_palette(rng: np.random.Generator) -> Palette:
    rgb = rng.random((256, 3))
    stops = [ColorStop(i / 255.0, *map(float, rgb[i])) for i in range(256)]
    return Palette(name="synthetic", stops=stops)


def _affine(rng, lo=-1.0, hi=1.0) -> Affine:
    a, b, c, d, e, f = (float(v) for v in rng.uniform(lo, hi, 6))
    return Affine(a, b, c, d, e, f)


def _variation(rng, name: str, weight: float, param_range: float = 2.0) -> Variation:
    v = Variation(index=VARIATION_INDEX[name], weight=weight)
    n = len(VARIATION_PARAMS.get(name, []))
    if VARIATION_INDEX[name] > 98 and name in VARIATION_DEFAULTS:
        v.params[:n] = VARIATION_DEFAULTS[name]
        return v
    for i in range(n):
        v.params[i] = float(rng.uniform(-param_range, param_range))
    return v


def make_genome(rng: np.random.Generator, n_xforms: int, var_pool: Sequence[str],
                vars_per_xform: Iterable[int], width: int, height: int, quality: float,
                oversample: int, post_fraction: float = 0.0, final: bool = False,
                lognormal_sigma: float = 0.0, name: str = "synthetic") -> Genome:
    counts = list(vars_per_xform)
    xforms: List[Xform] = []
    for i in range(n_xforms):
        k = counts[i % len(counts)] if len(counts) > 1 else counts[0]
        names = [var_pool[j] for j in rng.integers(0, len(var_pool), k)]
        ws = rng.dirichlet(np.ones(k))
        aff = _affine(rng)
        if lognormal_sigma > 0:
            s = float(rng.lognormal(0.0, lognormal_sigma))
            aff.a *= s; aff.b *= s; aff.d *= s; aff.e *= s
        x = Xform(weight=float(rng.uniform(0.2, 1.0)), color=float(rng.random()),
                  color_speed=0.5, affine=aff,
                  variations=[_variation(rng, n, float(w)) for n, w in zip(names, ws)])
        if post_fraction > 0 and i < int(round(n_xforms * post_fraction)):
            x.post = _affine(rng)
        xforms.append(x)
    fx = None
    if final:
        fnames = [var_pool[j] for j in rng.integers(0, len(var_pool), 1)]
        fx = Xform(weight=0.0, color=float(rng.random()), color_speed=0.0, affine=_affine(rng),
                   variations=[_variation(rng, n, 1.0) for n in fnames])
    long_edge = max(width, height)
    return Genome(name=name, xforms=xforms, scale=long_edge / 4.0, cx=0.0, cy=0.0,
                  palette=random_palette(rng), finalxform=fx, density=Density(9.0, 0.0, 0.4),
                  tonemap=Tonemap(gamma=4.0, vibrancy=1.0), quality=quality,
                  oversample=oversample, size={"width": width, "height": height})


def config0_genome(k: int = 0) -> Genome:
    rng = np.random.default_rng(0 * 1000 + k)
    return make_genome(rng, 3, CONFIG0_VARIATIONS, [3], 256, 256, 50, 1,
                       name=f"preview-small-{k}")


def config1_genome(k: int = 0, pool: Sequence[str] = None) -> Genome:
    rng = np.random.default_rng(1 * 1000 + k)
    pool = pool or flam3_core_pool()
    counts = [int(c) for c in rng.integers(2, 5, 6)]
    return make_genome(rng, 6, pool, counts, 1024, 768, 200, 2, post_fraction=0.5,
                       name=f"preview-{k}")


def flam3_core_pool() -> List[str]:
    """V0..V98 names (the flam3 core set), restricted to implemented entries."""
    from .variations_supported import SUPPORTED
    return [n for n, i in VARIATION_INDEX.items() if i <= 98 and i in SUPPORTED]


def implemented_pool(max_index: int = 400) -> List[str]:
    from .variations_supported import SUPPORTED
    return [n for n, i in VARIATION_INDEX.items() if i in SUPPORTED and i <= max_index]


def bench_genome(k: int, width: int = 1920, height: int = 1080) -> Genome:
    """1080p denoiser-workload genome: 12 xforms + final from the flam3 core
    set, log-normal(0.7) affine scales, quality 4000 reference."""
    rng = np.random.default_rng(4 * 1000 + k)
    counts = [int(c) for c in rng.integers(1, 4, 12)]
    return make_genome(rng, 12, flam3_core_pool(), counts, width, height, 4000, 1,
                       post_fraction=0.5, final=True, lognormal_sigma=0.7,
                       name=f"plain-1080p-{k}")


def config2_genome(k: int = 0, width: int = 3840, height: int = 2160) -> Genome:
    """4k workload genome: 12 xforms + final from the whole catalog."""
    rng = np.random.default_rng(2 * 1000 + k)
    counts = [int(c) for c in rng.integers(1, 4, 12)]
    return make_genome(rng, 12, implemented_pool(), counts, width, height, 200, 2,
                       post_fraction=0.5, final=True, lognormal_sigma=0.7,
                       name=f"4k-{k}")


def redraw(g: Genome, seed: int, lognormal_sigma: float = 0.7) -> Genome:
    """A genome of the same kernel shape as `g` (the same xforms, variation
    indices per slot, post-affine presence and final xform) with its other
    parameters drawn anew as make_genome draws them: affine coefficients
    (regular xforms' linear parts scaled by a log-normal factor, as the 4k
    workload's), xform weights and colours, variation weights and parameters
    (entries past the flam3 core set keep their defaults), the palette."""
    import copy
    rng = np.random.default_rng(seed)
    h = copy.deepcopy(g)
    xs = list(h.xforms) + ([h.finalxform] if h.finalxform is not None else [])
    for i, x in enumerate(xs):
        final = h.finalxform is not None and i == len(xs) - 1
        x.affine = _affine(rng)
        if not final:
            x.weight = float(rng.uniform(0.2, 1.0))
            if lognormal_sigma > 0:
                sc = float(rng.lognormal(0.0, lognormal_sigma))
                x.affine.a *= sc; x.affine.b *= sc; x.affine.d *= sc; x.affine.e *= sc
        if x.post is not None:
            x.post = _affine(rng)
        x.color = float(rng.random())
        ws = rng.dirichlet(np.ones(len(x.variations))) if not final else [1.0] * len(x.variations)
        for v, w in zip(x.variations, ws):
            v.weight = float(w)
            if v.index <= 98:
                for j in range(len(v.params)):
                    if v.params[j] != 0.0:
                        v.params[j] = float(rng.uniform(-2.0, 2.0))
    h.palette = random_palette(rng)
    h.name = f"{g.name}-redraw{seed}"
    return h
