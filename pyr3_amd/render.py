"""Command-line render of one genome file on the GPU.

The counterpart of `npm run render <input.flam3 | input.pyr3.json> [output]`
(bin/pyr3-render.ts) for the render path: load (`.flam3`/`.flame`,
`.pyr3.json`, or a `.png` saved with its genome), reset, one chaos dispatch
at the genome's quality (or `--quality`), present (DE + tonemap), then:

* auto-exposure (default on, `--no-auto-exposure` off; never for
  exr-linear): the brightness is refitted so the mean luminance matches a
  preview-tier probe render, re-presenting the same histogram
  (bin/pyr3-render.ts:480-514); the embedded genome keeps its own brightness;
* output per `--format` (bin/pyr3-render.ts:516-566): png8 (RGBA8), png16
  (the tonemap output rounded to binary16, then 16-bit samples), both with
  the genome in a `pyr3` tEXt chunk; exr (the linear light of that binary16
  display image); exr-linear (the histogram collapsed to scene-linear RGBA
  on the device, calibrated on the dispatched sample count).

    python -m pyr3_amd.render in.flam3 [out] [--quality Q] [--seed S]
        [--walkers N] [--oversample N] [--no-de] [--transparent]
        [--format png8|png16|exr|exr-linear] [--no-auto-exposure]

Walkers default to 163,840, one chaos block of 640 on each CU (the
reference's `--walkers` override; its own default of ~1,024 long walkers
fills only a few CUs of this GPU).
"""
from __future__ import annotations

import argparse
import os
import random
import sys
import time

from . import export, pack
from .flam3 import load_genome
from .genome import Tonemap
from .png import encode_render_png


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m pyr3_amd.render")
    ap.add_argument("input")
    ap.add_argument("output", nargs="?")
    ap.add_argument("--quality", type=float, default=None, help="samples per output pixel")
    ap.add_argument("--seed", type=int, default=None)
    ap.add_argument("--walkers", type=int, default=pack.CHIP_WALKERS)
    ap.add_argument("--oversample", type=int, default=None)
    ap.add_argument("--no-de", action="store_true", help="skip density estimation")
    ap.add_argument("--transparent", action="store_true")
    ap.add_argument("--format", choices=("png8", "png16", "exr", "exr-linear"), default="png8")
    ap.add_argument("--no-auto-exposure", action="store_true")
    a = ap.parse_args(argv)
    if a.walkers < 1:
        ap.error("--walkers must be positive")
    if a.quality is not None and not a.quality > 0:
        ap.error("--quality must be a positive number")
    is_exr = a.format in ("exr", "exr-linear")
    out = a.output or os.path.splitext(a.input)[0] + (".exr" if is_exr else ".png")
    if is_exr and os.path.splitext(out)[1].lower() != ".exr":
        out = os.path.splitext(out)[0] + ".exr"

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
    half = a.format in ("png16", "exr")
    r.set_image_f32(half)
    t0 = time.perf_counter()
    samples = r.render(g, seed=seed, force_de_off=a.no_de, transparent=a.transparent,
                       walkers=a.walkers)

    def measure() -> float:
        if half:
            return export.compute_mean_luminance_half(export.display_to_half(r.read_image_f32()))
        return export.compute_mean_luminance(r.read_image())

    if not a.no_auto_exposure and a.format != "exr-linear":
        base = (g.tonemap or Tonemap()).brightness
        target = export.measure_probe_luminance(g)
        initial = measure()

        def rerender(b: float) -> float:
            r.present(export.with_brightness(g, b), samples, a.no_de, a.transparent)
            return measure()
        fit = export.fit_brightness_to_target(base, target, initial, rerender)
        if fit.corrected:
            print(f"[pyr3-render] auto-exposure: brightness {base:.3f} -> {fit.brightness:.3f} "
                  f"(target {target:.2f}, was {initial:.2f}, now {fit.final_mean:.2f}, "
                  f"{fit.iters} pass{'' if fit.iters == 1 else 'es'})", flush=True)
    if a.format == "exr-linear":
        k1, k2 = pack.calibration(g.scale, samples, (g.tonemap or Tonemap()).brightness, os_)
        data = export.encode_exr(r.read_linear(k1, k2))
    elif a.format == "exr":
        data = export.encode_exr(export.display_half_to_linear_exr(
            export.display_to_half(r.read_image_f32())))
    elif a.format == "png16":
        data = encode_render_png(export.display_half_to_png16(
            export.display_to_half(r.read_image_f32())), g)
    else:
        data = encode_render_png(r.read_image(), g)
    dt = time.perf_counter() - t0
    r.destroy()
    with open(out, "wb") as fh:
        fh.write(data)
    print(f"[pyr3-render] {samples} samples in {dt:.3f}s (seed {seed}, {a.format}) -> {out}",
          flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
