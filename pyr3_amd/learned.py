"""The learned path: a low-quality render corrected by the BF16-MFMA
denoiser, gated per genome by the plain path as judge.

A learned frame is one chaos dispatch at a low quality with a short fuse
(`FUSE` turns instead of the plain path's 200: the fuse's sequential turns
would otherwise dominate a frame of ~25 post-fuse turns per walker; the
network was trained on 16-turn inputs and scores the same on 8- and 4-turn
ones, measured on held-out genomes), the tonemap without DE into the f32
image, then the denoiser, which reads the histogram and that image and
writes the RGBA8 frame (pyr3_amd/denoise/net.py).

Fallback: the plain high-quality render stays the judge. `LearnedRenderer`
scores a genome's learned frame against its plain render at the reference
quality once (PSNR over RGB); a genome below the target (38 dB) is rendered
by the plain path from then on, at the plain quality.
"""
from __future__ import annotations

import math
from typing import Dict, Optional

import numpy as np

from . import pack
from .genome import Genome

FUSE = 8              # fuse of the learned path's input renders
QUALITY = 4           # default input quality (samples per pixel)
TARGET_DB = 38.0
REF_QUALITY = 4000


def dispatch(quality: float, width: int, height: int,
             walkers: int = pack.CHIP_WALKERS) -> tuple:
    """(walkers, iters) of a render of `quality` samples per pixel: at most
    one walker per sample, whole chaos blocks."""
    blk = pack.CHAOS_BLOCK
    samples = max(1, int(round(quality * width * height)))
    walkers = max(blk, min(walkers, samples))
    walkers = -(-walkers // blk) * blk
    return walkers, -(-samples // walkers)


def psnr(a: np.ndarray, b: np.ndarray) -> float:
    """PSNR in dB over the RGB channels of two RGBA8 images."""
    d = a[..., :3].astype(np.float64) - b[..., :3].astype(np.float64)
    mse = float(np.mean(d * d))
    return float("inf") if mse == 0 else 10.0 * math.log10(255.0 ** 2 / mse)


class LearnedRenderer:
    """Renders genomes at one size through the learned path, falling back to
    the plain path for genomes whose learned frame scored below the target
    against the plain reference render."""

    def __init__(self, width: int, height: int, quality: float = QUALITY, fuse: int = FUSE,
                 weights=None, target_db: float = TARGET_DB, ref_quality: float = REF_QUALITY,
                 walkers: int = pack.CHIP_WALKERS):
        from .renderer import Denoiser, Renderer
        self.r = Renderer(width, height, 1)
        self.den = Denoiser.from_npz(weights)
        self.quality, self.fuse, self.walkers = quality, fuse, walkers
        self.target_db, self.ref_quality = target_db, ref_quality
        self.verdicts: Dict[str, float] = {}   # genome key -> learned PSNR

    @staticmethod
    def _key(g: Genome) -> str:
        from .genome import genome_to_text
        return genome_to_text(g)

    def render_learned(self, g: Genome, seed: int) -> int:
        """Learned frame into the renderer's RGBA8 image; returns samples."""
        w, it = dispatch(self.quality, self.r.width, self.r.height, self.walkers)
        self.r.set_image_f32(True)
        self.r.reset(g)
        s = self.r.iterate(g, seed, w, it, fuse=self.fuse)
        self.r.present(g, s, force_de_off=True)
        self.den.apply(self.r)
        return s

    def render_plain(self, g: Genome, seed: int, quality: Optional[float] = None) -> int:
        """Plain frame (full fuse, DE, tonemap); returns samples."""
        q = quality if quality is not None else (g.quality or pack.DEFAULT_SPP)
        w, it = dispatch(q, self.r.width, self.r.height, self.walkers)
        self.r.set_image_f32(False)
        self.r.reset(g)
        s = self.r.iterate(g, seed, w, it, fuse=pack.FUSE)
        self.r.present(g, s)
        return s

    def check(self, g: Genome, seed: int = 1) -> float:
        """Scores the genome's learned frame against its plain render at the
        reference quality and records the verdict; returns the PSNR."""
        self.render_plain(g, seed ^ 0x5EED, self.ref_quality)
        ref = self.r.read_image()
        self.render_learned(g, seed)
        db = psnr(self.r.read_image(), ref)
        self.verdicts[self._key(g)] = db
        return db

    def render(self, g: Genome, seed: int, plain_quality: Optional[float] = None) -> str:
        """One frame of `g` into the renderer's image: learned when the
        genome's verdict (computed on first use) reaches the target, else
        plain. Returns the path taken ("learned" or "plain")."""
        key = self._key(g)
        if key not in self.verdicts:
            self.check(g, seed)
        if self.verdicts[key] >= self.target_db:
            self.render_learned(g, seed)
            return "learned"
        self.render_plain(g, seed, plain_quality)
        return "plain"

    def read_image(self) -> np.ndarray:
        return self.r.read_image()

    def close(self) -> None:
        self.den.destroy()
        self.r.destroy()
