"""Command-line render of one genome file to a PNG on the GPU.

The counterpart of `npm run render <input.flam3 | input.pyr3.json> [output.png]`
(bin/pyr3-render.ts) for the render path: load (`.flam3`/`.flame`,
`.pyr3.json`, or a `.png` saved with its genome), reset, one chaos dispatch
at the genome's quality (or `--quality`), present (DE + tonemap), write an
RGBA8 PNG carrying the genome in its `pyr3` tEXt chunk.

    python -m pyr3_amd.render in.flam3 out.png [--quality Q] [--seed S]
        [--walkers N] [--oversample N] [--no-de] [--transparent]

Walkers default to 131,072 (the reference's `--walkers` override; its own
default of ~1,024 long walkers fills only a few CUs of this GPU).
"""
from __future__ import annotations

import argparse
import os
import random
import sys
import time

from . import pack
from .flam3 import load_genome
from .png import encode_render_png


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pyr3_amd.render")
    ap.add_argument("input")
    ap.add_argument("output", nargs="?")
    ap.add_argument("--quality", type=float, default=None, help="samples per output pixel")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--walkers", type=int, default=131072)
    ap.add_argument("--oversample", type=int, default=None)
    ap.add_argument("--no-de", action="store_true", help="skip density estimation")
    ap.add_argument("--transparent", action="store_true")
    a = ap.parse_args(argv)
    if a.walkers < 1:
        ap.error("--walkers must be positive")
    if a.quality is not None and not a.quality > 0:
        ap.error("--quality must be a positive number")
    out = a.output or os.path.splitext(a.input)[0] + ".png"

    g, kind, report = load_genome(a.input)
    if a.quality is not None:
        g.quality = a.quality
    size = g.size or {"width": 1024, "height": 1024}
    os_ = max(1, int(a.oversample or g.oversample or 1))
    radius = g.filter_radius if g.filter_radius is not None else pack.DEFAULT_FILTER_RADIUS
    dropped = len(report.dropped_variations) if report else 0
    ignored = len(report.ignored_fields) if report else 0
    print(f"[pyr3-render] genome={g.name!r} {size['width']}x{size['height']} oversample={os_} "
          f"({kind}{f', {dropped} dropped, {ignored} ignored' if dropped + ignored else ''})",
          flush=True)

    from .renderer import Renderer   # loads the HIP library; fails loudly without it
    r = Renderer(size["width"], size["height"], os_, filter_radius=radius)
    if not r.kernel_ready(g):
        print("[pyr3-render] compiling the chaos kernel for this genome's shape "
              "(cached for later renders)", flush=True)
    seed = a.seed if a.seed is not None else random.getrandbits(32)
    t0 = time.perf_counter()
    samples = r.render(g, seed=seed, force_de_off=a.no_de, transparent=a.transparent,
                       walkers=a.walkers)
    rgba = r.read_image()
    dt = time.perf_counter() - t0
    r.destroy()
    with open(out, "wb") as fh:
        fh.write(encode_render_png(rgba, g))
    print(f"[pyr3-render] {samples} samples in {dt:.3f}s (seed {seed}) -> {out}", flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
