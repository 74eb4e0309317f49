"""Renderer facade: the reference's stage interface over the C ABI.

Mirrors src/renderer.ts:163-262 (createRenderer / reset / iterate / present /
render / resize / readHistogram / destroy) with the same argument meaning:
``iterate`` queues one chaos dispatch that accumulates into the histogram,
``present`` runs DE (when the genome carries density settings and DE is not
forced off) and the tonemap into the renderer-owned RGBA8 image. The host
work per dispatch is the same as the reference's ChaosPass.dispatch
(src/chaos.ts:333-466): symmetry expansion, xform records, the pick table,
the uniform block. ISAAC seeding happens on the device.
"""
from __future__ import annotations

import ctypes as C
import random
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import curves, pack
from ._lib import check, load
from .genome import Genome, Tonemap


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _bp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


@dataclass
class PackedGenome:
    """The per-dispatch blobs for one genome (reusable across chunks)."""
    xforms: np.ndarray      # n_slots x 116 f32
    distrib: np.ndarray     # n_rows x 16384 u32
    expanded: Genome


def pack_genome(g: Genome) -> PackedGenome:
    e = pack.expand_genome_for_gpu(g)
    return PackedGenome(np.ascontiguousarray(pack.pack_xforms(e)),
                        np.ascontiguousarray(pack.pack_xform_distrib(e)), e)


class Renderer:
    def __init__(self, width: int, height: int, oversample: int = 1,
                 filter_radius: float = pack.DEFAULT_FILTER_RADIUS):
        self._lib = load()
        self._h = C.c_void_p()
        self._build(width, height, oversample, filter_radius, create=True)

    # -- lifecycle ------------------------------------------------------
    def _build(self, width, height, oversample, filter_radius, create):
        oversample = max(1, int(oversample))
        taps = np.ascontiguousarray(pack.spatial_kernel(filter_radius, oversample))
        if create:
            rc = self._lib.p3_create(width, height, oversample, _fp(taps), len(taps),
                                     C.byref(self._h))
            check(rc, None)
        else:
            check(self._lib.p3_resize(self._h, width, height, oversample, _fp(taps), len(taps)),
                  self._h)
        self.width, self.height, self.oversample = width, height, oversample
        self.super_w, self.super_h = width * oversample, height * oversample
        self.filter_radius = filter_radius
        self.taps = taps

    def resize(self, width: int, height: int, oversample: Optional[int] = None,
               filter_radius: Optional[float] = None) -> None:
        os_ = self.oversample if oversample is None else max(1, int(oversample))
        fr = self.filter_radius if filter_radius is None else filter_radius
        if (width, height, os_, fr) == (self.width, self.height, self.oversample,
                                        self.filter_radius):
            return
        self._build(width, height, os_, fr, create=False)

    def destroy(self) -> None:
        if self._h:
            self._lib.p3_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass

    # -- stage interface ------------------------------------------------
    def reset(self, genome: Genome) -> None:
        lut = np.ascontiguousarray(pack.pack_palette(genome).reshape(-1))
        check(self._lib.p3_reset(self._h, _fp(lut)), self._h)

    def chaos_uniforms(self, genome: Genome, seed: int, walkers: int, iters: int,
                       fuse: int = pack.FUSE, walker_jitter: Optional[float] = None,
                       trace: bool = False, color: Optional[pack.ColorOpts] = None) -> np.ndarray:
        jit = pack.DEFAULT_WALKER_JITTER if walker_jitter is None else walker_jitter
        return pack.chaos_uniforms(genome, self.super_w, self.super_h, self.oversample, iters,
                                   fuse, seed, walkers, jit, trace, color)

    def iterate(self, genome: Genome, seed: int, walkers: int, iters_per_walker: int,
                walker_jitter: Optional[float] = None, fuse: int = pack.FUSE,
                packed: Optional[PackedGenome] = None, isaac_states: Optional[np.ndarray] = None,
                color: Optional[pack.ColorOpts] = None) -> int:
        """One chaos dispatch; returns the samples it contributes (walkers * iters).
        `color` selects the splat colour mode and the colour-index capture."""
        pg = packed or pack_genome(genome)
        u = np.ascontiguousarray(self.chaos_uniforms(pg.expanded, seed, walkers,
                                                     iters_per_walker, fuse, walker_jitter,
                                                     color=color))
        n_slots, n_rows = pg.xforms.shape[0], pg.distrib.shape[0]
        if isaac_states is None:
            rc = self._lib.p3_iterate(self._h, _fp(pg.xforms), n_slots, _up(pg.distrib), n_rows,
                                      _up(u))
        else:
            st = np.ascontiguousarray(isaac_states, dtype=np.uint32)
            if st.size != walkers * 36:
                raise ValueError(f"isaac_states must hold {walkers} x 36 words")
            rc = self._lib.p3_iterate_states(self._h, _fp(pg.xforms), n_slots, _up(pg.distrib),
                                             n_rows, _up(u), _up(st))
        check(rc, self._h)
        return walkers * iters_per_walker

    def iterate_multi(self, genomes, seeds, walkers, iters_per_walker: int,
                      walker_jitter: Optional[float] = None, fuse: int = pack.FUSE,
                      packed=None) -> int:
        """Several dispatches of one genome shape (e.g. the sub-frames of a
        motion-blurred frame) in one batched launch; the histogram equals
        that of one `iterate` per (genome, seed, walkers). Returns samples."""
        n = len(genomes)
        if not (len(seeds) == len(walkers) == n):
            raise ValueError("genomes, seeds and walkers must have the same length")
        pgs = packed or [pack_genome(g) for g in genomes]
        us = [np.ascontiguousarray(self.chaos_uniforms(pg.expanded, sd, w, iters_per_walker, fuse,
                                                       walker_jitter))
              for pg, sd, w in zip(pgs, seeds, walkers)]
        n_slots, n_rows = pgs[0].xforms.shape[0], pgs[0].distrib.shape[0]
        if any(pg.xforms.shape[0] != n_slots or pg.distrib.shape[0] != n_rows for pg in pgs):
            raise ValueError("sub-dispatches differ in xform count")
        from .animate import packed_shape   # the batched kernel runs the first one's shape
        if len({packed_shape(pg) for pg in pgs}) != 1:
            raise ValueError("sub-dispatches differ in kernel shape (variation slots, "
                             "post-affine presence or final xform)")
        xfp = (C.POINTER(C.c_float) * n)(*[_fp(pg.xforms) for pg in pgs])
        dsp = (C.POINTER(C.c_uint32) * n)(*[_up(pg.distrib) for pg in pgs])
        ufp = (C.POINTER(C.c_uint32) * n)(*[_up(u) for u in us])
        check(self._lib.p3_iterate_multi(self._h, n, xfp, n_slots, dsp, n_rows, ufp), self._h)
        return int(sum(walkers)) * iters_per_walker

    def present(self, genome: Genome, total_samples: float, force_de_off: bool = False,
                transparent: bool = False) -> None:
        tm = genome.tonemap or Tonemap()
        k1, k2 = pack.calibration(genome.scale, total_samples, tm.brightness, self.oversample)
        de = None
        if genome.density is not None and not force_de_off:
            d = genome.density
            de = np.ascontiguousarray(pack.de_uniforms(self.super_w, self.super_h, d.max_rad,
                                                       d.min_rad, d.curve, k1, k2,
                                                       self.oversample))
        bg = genome.background or (0.0, 0.0, 0.0)
        mask = curves.active_mask(genome.channel_curves)
        if mask:
            self.set_curves(curves.bake_curves(genome.channel_curves))
        viz = np.ascontiguousarray(pack.viz_uniforms(self.width, self.height, k1, k2, tm,
                                                     self.oversample, len(self.taps), bg,
                                                     transparent, mask, genome.hsl_adjust))
        check(self._lib.p3_present(self._h, _up(de) if de is not None else None, _up(viz)),
              self._h)

    def render(self, genome: Genome, seed: Optional[int] = None, force_de_off: bool = False,
               transparent: bool = False, walker_jitter: Optional[float] = None,
               walkers: Optional[int] = None) -> int:
        """reset + one iterate sized by computeDispatch + present; returns samples.
        `walkers` pins the walker count (the reference's --walkers override:
        iters = ceil(samples / walkers)); the default is the reference's
        schedule of ~1,024 long walkers, which fills only a few CUs, so GPU
        callers pass a walker count of the order of 10^5."""
        if seed is None:
            seed = random.getrandbits(32)
        spp = genome.quality if genome.quality is not None else pack.DEFAULT_SPP
        walkers, iters, samples = pack.compute_dispatch(spp, self.width, self.height, walkers)
        self.reset(genome)
        self.iterate(genome, seed, walkers, iters, walker_jitter)
        self.present(genome, samples, force_de_off, transparent)
        return samples

    # -- readbacks ------------------------------------------------------
    def read_histogram(self):
        out = np.empty((self.super_h, self.super_w, 4), dtype=np.uint32)
        check(self._lib.p3_read_histogram(self._h, _up(out), out.size), self._h)
        return out

    def probe_atomics(self, blocks: int = 4096, per_thread: int = 256):
        """(atomic requests per second, requests per splat) of a uniform
        scatter over this renderer's histogram in the splat's request form
        (the histogram is overwritten: reset before rendering)."""
        out, per = C.c_double(), C.c_int()
        check(self._lib.p3_probe_atomics(self._h, blocks, per_thread, C.byref(out),
                                         C.byref(per)), self._h)
        return out.value, per.value

    def read_linear(self, k1: float, k2: float) -> np.ndarray:
        """(H, W, 4) f32 scene-linear RGBA from the histogram on the device
        (histogramToLinearRgba); k1, k2 as `pack.calibration` gives them."""
        out = np.empty((self.height, self.width, 4), dtype=np.float32)
        check(self._lib.p3_read_linear(self._h, float(k1), float(k2), _fp(out), out.size),
              self._h)
        return out

    def set_curves(self, lut: Optional[np.ndarray]) -> None:
        """Upload the (5, 256) channel-curves LUT (None clears it)."""
        if lut is None:
            check(self._lib.p3_set_curves(self._h, None, 0), self._h)
            return
        arr = np.ascontiguousarray(lut, dtype=np.float32).reshape(-1)
        check(self._lib.p3_set_curves(self._h, _fp(arr), arr.size), self._h)

    def write_histogram(self, hist: np.ndarray) -> None:
        h = np.ascontiguousarray(hist, dtype=np.uint32)
        check(self._lib.p3_write_histogram(self._h, _up(h), h.size), self._h)

    def kernel_ready(self, genome: Genome, packed: Optional[PackedGenome] = None) -> bool:
        """Whether this genome's specialised chaos kernel is already compiled."""
        pg = packed or pack_genome(genome)
        rc = self._lib.p3_chaos_ready(_fp(pg.xforms), pg.xforms.shape[0], pg.distrib.shape[0],
                                      pack.final_xform_slot(pg.expanded))
        if rc < 0:
            check(rc, self._h)
        return rc == 1

    def read_index(self):
        """(H_s, W_s) u32 colour-index sums of the capturing dispatches."""
        out = np.zeros((self.super_h, self.super_w), dtype=np.uint32)
        check(self._lib.p3_read_index(self._h, _up(out), out.size), self._h)
        return out

    def read_filtered(self):
        out = np.empty((self.super_h, self.super_w, 4), dtype=np.float32)
        check(self._lib.p3_read_filtered(self._h, _fp(out), out.size), self._h)
        return out

    def read_image(self):
        out = np.empty((self.height, self.width, 4), dtype=np.uint8)
        check(self._lib.p3_read_image(self._h, _bp(out), out.size), self._h)
        return out

    def read_image_f32(self):
        out = np.empty((self.height, self.width, 4), dtype=np.float32)
        check(self._lib.p3_read_image_f32(self._h, _fp(out), out.size), self._h)
        return out

    def set_trace(self, walkers: int, iters: int) -> None:
        check(self._lib.p3_set_trace(self._h, walkers, iters), self._h)
        self._trace_shape = (walkers, iters)

    def read_trace(self):
        w, i = self._trace_shape
        out = np.empty((w, i, 16), dtype=np.float32)
        check(self._lib.p3_read_trace(self._h, _fp(out), out.size), self._h)
        return out

    def set_image_f32(self, on: bool) -> None:
        """Also keep the pre-quantisation RGBA f32 image on each present."""
        check(self._lib.p3_set_image_f32(self._h, 1 if on else 0), self._h)

    def set_timing(self, on: bool) -> None:
        check(self._lib.p3_set_timing(self._h, 1 if on else 0), self._h)

    def read_timings(self):
        """Last kernel durations (ms): merge, chaos, density, tonemap, reset, denoise."""
        out = (C.c_float * 6)()
        check(self._lib.p3_read_timings(self._h, out), self._h)
        keys = ("merge", "chaos", "density", "tonemap", "reset", "denoise")
        return dict(zip(keys, list(out)))

    def sync(self) -> None:
        check(self._lib.p3_sync(self._h), self._h)


class Denoiser:
    """The learned denoiser on the GPU (BF16 MFMA kernels): corrects a
    renderer's low-quality frame presented without DE and with the f32 image
    enabled."""

    def __init__(self, params: np.ndarray):
        self._lib = load()
        p = np.ascontiguousarray(params, dtype=np.float32).reshape(-1)
        self._h = C.c_void_p()
        check(self._lib.p3_denoiser_create(_fp(p), p.size, C.byref(self._h)), None)

    @staticmethod
    def param_count() -> int:
        return int(load().p3_denoiser_param_count())

    @classmethod
    def from_npz(cls, path=None) -> "Denoiser":
        from .denoise import weights
        return cls(weights.flat_params(weights.load(path or weights.DEFAULT)))

    def apply(self, r: Renderer) -> None:
        """Denoise the renderer's last present (DE off, f32 image enabled)
        into its RGBA8 image."""
        check(self._lib.p3_denoise(r._h, self._h), r._h)

    def destroy(self) -> None:
        if self._h:
            self._lib.p3_denoiser_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def create_renderer(width: int, height: int, oversample: Optional[int] = None,
                    filter_radius: Optional[float] = None) -> Renderer:
    """src/renderer.ts:214 createRenderer."""
    return Renderer(width, height, oversample or 1,
                    pack.DEFAULT_FILTER_RADIUS if filter_radius is None else filter_radius)
