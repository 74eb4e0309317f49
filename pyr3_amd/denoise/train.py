"""Train the learned denoiser with the hand-written BF16-MFMA training
kernels (include/pyr3nn.h, driven by net.HipNet) on one GPU.

Pairs are generated by this package's plain path from synthetic genomes of
the 4k workload's distribution (12 xforms + final from the whole catalog,
log-normal(0.7) affine scales), framed at 1920x1080 oversample 1:
  target  the plain render at quality `--ref-quality` (fuse 200, DE, tonemap)
  inputs  learned-path renders at the qualities `--qualities` (fuse
          learned.FUSE, tonemap without DE, independent seeds), as the
          network sees them: p3nn_features of the histogram + the base image
Genomes whose attractor covers less than 5% of the frame are skipped. Each
step trains on `--batch` random 256x256 crops (input, base, target) with
MSE on RGB and Adam (warm-up, then cosine decay). The kernel shapes of the
genomes compile in worker processes while the GPU renders the ready ones.

Writes `--out` (.npz: the parameters in the W[o][kp] layout of net.py and a
`meta` JSON string) and prints a progress line per phase and every 100 steps.

    python -m pyr3_amd.denoise.train --out weights.npz
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor, as_completed
from multiprocessing import get_context

import numpy as np

from .. import learned, pack, synth
from . import net

TRAIN_BASE = 20000    # genome seeds of the training set (the held-out set starts at 40000)
W, H = 1920, 1080


def log(msg):
    print(f"[train] {msg}", flush=True)


def shape_key(g) -> str:
    from ..kcache import key_of
    from ..renderer import pack_genome
    pg = pack_genome(g)
    return key_of(pg.xforms, pg.distrib.shape[0], pack.final_xform_slot(pg.expanded))


def compile_ahead(genomes, workers: int):
    """Process pool compiling the genomes' kernel shapes into the code-object
    cache; returns (pool, {future: index})."""
    from ..kcache import CACHE, _compile_key
    cache = os.environ.get("P3_KERNEL_CACHE") or str(CACHE)
    ex = ProcessPoolExecutor(max_workers=workers, mp_context=get_context("spawn"))
    futs = {}
    seen = {}
    for i, g in enumerate(genomes):
        k = shape_key(g)
        if k in seen:
            continue
        seen[k] = i
        futs[ex.submit(_compile_key, k, cache)] = k
    return ex, futs, seen


def learned_input(r, hn, g, quality: float, seed: int):
    """(features (1, H, W, 8) bf16, base (1, H, W, 4) f32, coverage) of one
    learned-path input render."""
    import torch
    w, it = learned.dispatch(quality, r.width, r.height)
    r.set_image_f32(True)
    r.reset(g)
    s = r.iterate(g, seed, w, it, fuse=learned.FUSE)
    r.present(g, s, force_de_off=True)
    hist = r.read_histogram()
    img = r.read_image_f32()
    _, k2 = pack.calibration(g.scale, s, g.tonemap.brightness, 1)
    feat = hn.features(torch.from_numpy(hist.view(np.int32)).cuda(), torch.from_numpy(img).cuda(),
                       k2)
    cover = float(np.mean(hist[..., 3] > 0))
    return feat, torch.from_numpy(img).cuda()[None], cover


def plain_target(r, g, quality: float, seed: int):
    import torch
    w, it = learned.dispatch(quality, r.width, r.height)
    r.set_image_f32(False)
    r.reset(g)
    s = r.iterate(g, seed, w, it, fuse=pack.FUSE)
    r.present(g, s)
    img = r.read_image()
    return torch.from_numpy(img[..., :3].astype(np.float32) / 255.0).cuda()[None]


def make_data(genomes, hn, qualities, ref_quality, workers):
    """Renders inputs and targets of every usable genome (kernels compiled
    in worker processes as the GPU goes)."""
    from ..renderer import Renderer
    ex, futs, first = compile_ahead(genomes, workers)
    ready = set()
    r = Renderer(W, H, 1)
    inputs, targets = [], []   # inputs: (feat, base, target index)
    t0 = time.time()
    todo = list(range(len(genomes)))
    try:
        for f in as_completed(futs):
            key, rc, _ = f.result()
            if rc != 0:
                raise RuntimeError(f"kernel compile failed for {key}")
            ready.add(key)
            for i in [i for i in todo if shape_key(genomes[i]) in ready]:
                todo.remove(i)
                g = genomes[i]
                feat, base, cover = learned_input(r, hn, g, qualities[0], 11 + i)
                if cover < 0.05:
                    continue
                targets.append(plain_target(r, g, ref_quality, 0x9E3779B9 + i))
                ti = len(targets) - 1
                inputs.append((feat, base, ti))
                for j, q in enumerate(qualities[1:]):
                    f2, b2, _ = learned_input(r, hn, g, q, 1000 * (j + 2) + i)
                    inputs.append((f2, b2, ti))
            if len(targets) and len(targets) % 8 == 0:
                log(f"data: {len(targets)} genomes rendered, {len(todo)} waiting "
                    f"({time.time() - t0:.0f}s)")
    finally:
        ex.shutdown()
        r.destroy()
    log(f"data: {len(targets)} usable genomes of {len(genomes)}, {len(inputs)} inputs "
        f"({time.time() - t0:.0f}s)")
    return inputs, targets


def batch_of(inputs, targets, rng, n: int, crop: int):
    import torch
    fs, bs, ts = [], [], []
    for _ in range(n):
        feat, base, ti = inputs[int(rng.integers(len(inputs)))]
        y = int(rng.integers(0, H - crop + 1))
        x = int(rng.integers(0, W - crop + 1))
        fs.append(feat[:, y:y + crop, x:x + crop])
        bs.append(base[:, y:y + crop, x:x + crop])
        ts.append(targets[ti][:, y:y + crop, x:x + crop])
    return (torch.cat(fs).contiguous(), torch.cat(bs).contiguous(), torch.cat(ts).contiguous())


def lr_at(step: int, steps: int, peak: float, warm: int = 100) -> float:
    if step < warm:
        return peak * (step + 1) / warm
    t = (step - warm) / max(1, steps - warm)
    return 1e-5 + (peak - 1e-5) * 0.5 * (1.0 + math.cos(math.pi * t))


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--genomes", type=int, default=96)
    ap.add_argument("--qualities", default="4,2,3,6")
    ap.add_argument("--ref-quality", type=float, default=4000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--crop", type=int, default=256)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--workers", type=int, default=16)
    ap.add_argument("--init", default=None, help="start from these weights (.npz)")
    a = ap.parse_args(argv)
    import torch
    t_start = time.time()
    qualities = [float(q) for q in a.qualities.split(",")]
    genomes = [synth.config2_genome(k, W, H) for k in range(TRAIN_BASE, TRAIN_BASE + a.genomes)]
    params = net.init_params(a.seed)
    if a.init:
        from .weights import load
        params = load(a.init)
    hn = net.HipNet(params)
    inputs, targets = make_data(genomes, hn, qualities, a.ref_quality, a.workers)
    rng = np.random.default_rng(a.seed)
    t0 = time.time()
    acc, n_acc = 0.0, 0
    history = []
    for step in range(a.steps):
        feat, base, target = batch_of(inputs, targets, rng, a.batch, a.crop)
        hn.forward(feat, base)
        loss = hn.loss_backward(target)
        hn.adam_step(lr_at(step, a.steps, a.lr))
        if step % 10 == 0:
            acc += loss.item()
            n_acc += 1
        if (step + 1) % 100 == 0:
            mse = acc / max(1, n_acc)
            history.append((step + 1, mse))
            log(f"step {step + 1}/{a.steps}: mse {mse:.3e} ({10 * math.log10(1 / mse):.2f} dB "
                f"on crops), {(time.time() - t0) / (step + 1) * 1e3:.1f} ms/step")
            acc, n_acc = 0.0, 0
    torch.cuda.synchronize()
    out = hn.params_numpy()
    meta = {"genomes": len(targets), "inputs": len(inputs), "qualities": qualities,
            "ref_quality": a.ref_quality, "fuse": learned.FUSE, "steps": a.steps,
            "batch": a.batch, "crop": a.crop, "lr": a.lr, "size": [W, H],
            "distribution": "4k workload genomes (synth.config2_genome) framed at 1920x1080",
            "genome_seeds": [TRAIN_BASE, TRAIN_BASE + a.genomes],
            "train_s": round(time.time() - t0, 1), "total_s": round(time.time() - t_start, 1),
            "loss_history": history, "kernels": "pyr3nn (hand-written BF16 MFMA)"}
    np.savez(a.out, meta=np.array(json.dumps(meta)), **out)
    log(f"wrote {a.out} ({net.param_count()} parameters) in {time.time() - t_start:.0f}s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
