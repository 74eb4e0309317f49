"""Score the learned path on held-out genomes against the plain path.

Held-out genomes are drawn from the training distribution (4k workload
genomes framed at 1920x1080) with seeds disjoint from the training set's:
`--shapes` kernel shapes (synth.config2_genome), each with `--per-shape`
parameter draws (synth.redraw: new affines, weights, variation parameters,
palette), so the set needs few kernel compiles. For each genome: the plain
render at `--ref-quality` is the judge, the learned frame (learned-path
input at each of `--qualities`, then the denoiser through the C ABI) is
scored by PSNR over RGB, next to the plain path at the same quality for
comparison; with several `--fuses` the learned input is rendered (same
seed) at each fuse. Genomes whose attractor covers less than 5% of the
frame are skipped (as in training). Writes a JSON report (per quality: mean, median,
10th percentile PSNR and the share of genomes at or above the 38 dB target)
and prints a progress line every 25 genomes.

    python -m pyr3_amd.denoise.evaluate --weights weights.npz --report eval.json
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from .. import learned, pack, synth
from .train import H, W, compile_ahead, log, shape_key

HELDOUT_BASE = 40000


def _lkey(q: float, fuse: int, fuses) -> str:
    """Row key of a learned score: the quality, with the fuse when several
    fuses are scored."""
    return str(q) if len(fuses) == 1 else f"{q}/f{fuse}"


def heldout_genomes(n_shapes: int, per_shape: int):
    out = []
    for k in range(HELDOUT_BASE, HELDOUT_BASE + n_shapes):
        g = synth.config2_genome(k, W, H)
        out.append(g)
        out += [synth.redraw(g, 7919 * k + j) for j in range(1, per_shape)]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", default=None)
    ap.add_argument("--report", required=True)
    ap.add_argument("--shapes", type=int, default=125)
    ap.add_argument("--per-shape", type=int, default=8)
    ap.add_argument("--qualities", default="3,4,6")
    ap.add_argument("--fuses", default=str(learned.FUSE),
                    help="input fuses of the learned path to score (comma list)")
    ap.add_argument("--ref-quality", type=float, default=4000)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--max-genomes", type=int, default=0)
    ap.add_argument("--time-budget", type=float, default=0.0,
                    help="stop starting new genomes after this many seconds (0: none)")
    a = ap.parse_args(argv)
    from ..renderer import Denoiser, Renderer
    t0 = time.time()
    qs = [float(q) for q in a.qualities.split(",")]
    fuses = [int(f) for f in a.fuses.split(",")]
    genomes = heldout_genomes(a.shapes, a.per_shape)
    if a.max_genomes:
        genomes = genomes[:a.max_genomes]
    ex, futs, _ = compile_ahead(genomes, a.workers)
    for f in futs:
        key, rc, _ = f.result()
        if rc != 0:
            raise RuntimeError(f"kernel compile failed for {key}")
    ex.shutdown()
    log(f"eval: {len(futs)} shapes compiled for {len(genomes)} genomes ({time.time() - t0:.0f}s)")
    r = Renderer(W, H, 1)
    den = Denoiser.from_npz(a.weights)
    rows = []
    t1 = time.time()
    for i, g in enumerate(genomes):
        if a.time_budget and time.time() - t0 > a.time_budget:
            log(f"eval: time budget reached after {i} genomes")
            break
        w, it = learned.dispatch(qs[0], W, H)
        r.set_image_f32(False)
        r.reset(g)
        s = r.iterate(g, 3 + i, w, it, fuse=pack.FUSE)
        r.present(g, s)
        if float(np.mean(r.read_histogram()[..., 3] > 0)) < 0.05:
            continue
        w, it = learned.dispatch(a.ref_quality, W, H)
        r.reset(g)
        s = r.iterate(g, 0x9E3779B9 ^ i, w, it, fuse=pack.FUSE)
        r.present(g, s)
        ref = r.read_image()
        row = {"genome": g.name, "learned": {}, "plain": {}}
        for q in qs:
            w, it = learned.dispatch(q, W, H)
            for fz in fuses:
                r.set_image_f32(True)
                r.reset(g)
                s = r.iterate(g, 101 + i, w, it, fuse=fz)
                r.present(g, s, force_de_off=True)
                den.apply(r)
                row["learned"][_lkey(q, fz, fuses)] = round(learned.psnr(r.read_image(), ref), 3)
            r.set_image_f32(False)
            r.reset(g)
            s = r.iterate(g, 101 + i, w, it, fuse=pack.FUSE)
            r.present(g, s)
            row["plain"][str(q)] = round(learned.psnr(r.read_image(), ref), 3)
        rows.append(row)
        if len(rows) % 25 == 0:
            log(f"eval: {len(rows)} genomes scored ({time.time() - t1:.0f}s); learned "
                + ", ".join(f"{k} mean {np.mean([x['learned'][k] for x in rows]):.2f} dB"
                            for k in rows[0]["learned"]))
    den.destroy()
    r.destroy()
    summary = {}
    keys = [("learned", _lkey(q, fz, fuses), f"learned_q{q:g}" + (f"_fuse{fz}" if len(fuses) > 1 else ""))
            for q in qs for fz in fuses] + [("plain", str(q), f"plain_q{q:g}") for q in qs]
    for path, key, name in keys:
        v = np.array([x[path][key] for x in rows])
        v = v[np.isfinite(v)]
        summary[name] = {
            "mean_db": round(float(v.mean()), 3), "median_db": round(float(np.median(v)), 3),
            "p10_db": round(float(np.percentile(v, 10)), 3),
            "share_at_target": round(float(np.mean(v >= learned.TARGET_DB)), 4)}
    rep = {"genomes_scored": len(rows), "genomes_drawn": len(genomes), "shapes": a.shapes,
           "per_shape": a.per_shape, "ref_quality": a.ref_quality, "fuses": fuses,
           "size": [W, H], "target_db": learned.TARGET_DB, "summary": summary,
           "seconds": round(time.time() - t0, 1), "rows": rows}
    with open(a.report, "w") as f:
        json.dump(rep, f)
    log(f"eval: {json.dumps(summary)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
