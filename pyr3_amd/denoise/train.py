"""Train the denoiser on rendered pairs (runs on one GPU).

Pairs: seeded synthetic flames of the benchmark family (12 xforms + final,
flam3 core set; genome seeds disjoint from the ones the benchmark uses),
rendered by this package's renderer at `--width x --height`: a quality-4000
target and inputs at a few low qualities with independent seeds. Loss: MSE
on RGB (the benchmark scores PSNR). Writes `--out` (.npz: one float32 array
per weight, plus `meta` JSON) and prints a progress line per phase.

    python -m pyr3_amd.denoise.train --out denoiser.npz
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
import torch

from .. import pack, synth
from ..renderer import Renderer, pack_genome
from .model import Denoiser, features

GENOME_BASE = 1000   # bench_genome(k) seeds the benchmark searches start at k = 0


def render(r: Renderer, g, quality: float, seed: int, walkers: int = 131072):
    samples = int(round(quality * r.width * r.height))
    walkers = max(256, min(walkers, samples) // 256 * 256)
    iters = -(-samples // walkers)
    r.reset(g)
    s = r.iterate(g, seed, walkers, iters, packed=pack_genome(g))
    r.present(g, s)
    return r.read_image_f32(), r.read_histogram()[..., 3], r.read_image()


def log(msg):
    print(f"[train] {msg}", flush=True)


def make_pairs(n: int, width: int, height: int, qualities, ref_q: float, dev):
    from ..batch import precompile
    n_genomes = -(-n // len(qualities))
    cands = [synth.bench_genome(k, width, height)
             for k in range(GENOME_BASE, GENOME_BASE + n_genomes * 3 // 2 + 8)]
    t0 = time.time()
    log(f"compiling {len(cands)} genome shapes ...")
    precompile(cands)
    log(f"compiled in {time.time() - t0:.0f}s")
    r = Renderer(width, height, 1)
    r.set_image_f32(True)
    X, Y = [], []
    k = GENOME_BASE
    while len(Y) < n and k - GENOME_BASE < len(cands):
        g = cands[k - GENOME_BASE]
        k += 1
        ref, cnt, _ = render(r, g, 20, 7 + k)
        if np.mean(cnt > 0) < 0.05:
            continue
        ref, _, _ = render(r, g, ref_q, 0x9E3779B9 + k)
        ys = torch.from_numpy(np.ascontiguousarray(ref[..., :3].transpose(2, 0, 1))).to(dev)
        for q in qualities:
            img, cnt, _ = render(r, g, q, 1000 * k + int(q))
            rgb = torch.from_numpy(np.ascontiguousarray(img[..., :3].transpose(2, 0, 1))).to(dev)
            c = torch.from_numpy(cnt.astype(np.float32)[None]).to(dev)
            X.append(features(rgb[None], c[None])[0].to(torch.bfloat16))
            Y.append(ys.to(torch.bfloat16))
        if len(Y) % (10 * len(qualities)) == 0:
            log(f"{len(Y)} pairs from {k - GENOME_BASE} genomes ({time.time() - t0:.0f}s)")
    r.destroy()
    return torch.stack(X), torch.stack(Y)


def train(X, Y, steps: int, crop: int, batch: int, lr: float, dev, seed: int = 0):
    torch.manual_seed(seed)
    net = Denoiser().to(dev)
    opt = torch.optim.AdamW(net.parameters(), lr=lr, weight_decay=1e-4)
    sched = torch.optim.lr_scheduler.OneCycleLR(opt, max_lr=lr, total_steps=steps, pct_start=0.1)
    N, _, H, W = X.shape
    g = torch.Generator(device="cpu").manual_seed(seed)
    t0 = time.time()
    for step in range(steps):
        idx = torch.randint(0, N, (batch,), generator=g)
        ys = torch.randint(0, H - crop + 1, (batch,), generator=g)
        xs = torch.randint(0, W - crop + 1, (batch,), generator=g)
        xb = torch.stack([X[i, :, a:a + crop, b:b + crop] for i, a, b in zip(idx, ys, xs)])
        yb = torch.stack([Y[i, :, a:a + crop, b:b + crop] for i, a, b in zip(idx, ys, xs)])
        with torch.autocast("cuda", dtype=torch.bfloat16):
            out = net(xb)
        loss = torch.mean((out.float() - yb.float()) ** 2)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if step % 100 == 0 or step == steps - 1:
            log(f"step {step} loss {loss.item():.3e} psnr~{-10 * math.log10(max(loss.item(), 1e-12)):.2f} "
                f"({time.time() - t0:.0f}s)")
    return net


def psnr8(a: np.ndarray, b: np.ndarray) -> float:
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10 * math.log10(255.0 ** 2 / mse)


def to_rgba8(rgb: torch.Tensor) -> np.ndarray:
    """(3, H, W) float -> (H, W, 4) u8 like the tonemap store (round to nearest)."""
    v = (rgb.float().clamp(0, 1) * 255.0 + 0.5).floor().to(torch.uint8)
    a = torch.full_like(v[:1], 255)
    return torch.cat([v, a], 0).permute(1, 2, 0).cpu().numpy()


def evaluate(net, dev, width: int, height: int, ks, qualities, ref_q: float):
    r = Renderer(width, height, 1)
    r.set_image_f32(True)
    rows = []
    for k in ks:
        g = synth.bench_genome(k, width, height)
        _, _, ref8 = render(r, g, ref_q, 0x9E3779B9 + k)
        for q in qualities:
            img, cnt, img8 = render(r, g, q, 500 + int(q))
            rgb = torch.from_numpy(np.ascontiguousarray(img[..., :3].transpose(2, 0, 1))).to(dev)
            c = torch.from_numpy(cnt.astype(np.float32)[None]).to(dev)
            with torch.no_grad(), torch.autocast("cuda", dtype=torch.bfloat16):
                out = net(features(rgb[None], c[None]))[0]
            rows.append({"k": k, "q": q, "plain_db": round(psnr8(img8, ref8), 2),
                         "learned_db": round(psnr8(to_rgba8(out), ref8), 2)})
            log(str(rows[-1]))
    r.destroy()
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="denoiser.npz")
    ap.add_argument("--pairs-genomes", type=int, default=120)
    ap.add_argument("--width", type=int, default=960)
    ap.add_argument("--height", type=int, default=540)
    ap.add_argument("--qualities", default="3,6,12")
    ap.add_argument("--ref-quality", type=float, default=4000)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--crop", type=int, default=128)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-k", default="0")
    ap.add_argument("--eval-q", default="3,5,8,10,15")
    a = ap.parse_args(argv)
    dev = torch.device("cuda")
    qs = [float(v) for v in a.qualities.split(",")]
    t0 = time.time()
    X, Y = make_pairs(a.pairs_genomes * len(qs), a.width, a.height, qs, a.ref_quality, dev)
    log(f"data: {tuple(X.shape)} in {time.time() - t0:.0f}s")
    net = train(X, Y, a.steps, a.crop, a.batch, a.lr, dev)
    del X, Y
    rows = evaluate(net, dev, 1920, 1080, [int(v) for v in a.eval_k.split(",")],
                    [float(v) for v in a.eval_q.split(",")], a.ref_quality)
    arrays = {f"{n}.{p}": t.detach().float().cpu().numpy() for n, m in net.convs.items()
              for p, t in (("weight", m.weight), ("bias", m.bias))}
    arrays["out.weight"] = net.out.weight.detach().float().cpu().numpy()
    arrays["out.bias"] = net.out.bias.detach().float().cpu().numpy()
    meta = {"qualities": qs, "ref_quality": a.ref_quality, "steps": a.steps,
            "pairs_genomes": a.pairs_genomes, "size": [a.width, a.height],
            "genome_base": GENOME_BASE, "eval": rows}
    np.savez(a.out, meta=np.array(json.dumps(meta)), **arrays)
    log(f"wrote {a.out} ({time.time() - t0:.0f}s total)")


if __name__ == "__main__":
    main(sys.argv[1:])
