"""Batched rendering of many genomes (the 512-genome batch workload).

A batch is a list of jobs, each one output frame: a genome at a size and
quality, optionally a motion-blurred frame of N sub-frames (the sub-render
loop of src/animate-render.ts:73-171: walkers apportioned over the temporal
filter, sub-frame seed = seed + i * 0x9e3779b1, one present).

The MI355X-side design:
  * every distinct genome shape needs a specialised chaos kernel; the shapes
    of the whole batch are compiled up front by a pool of worker processes
    (hipRTC needs no GPU) into the persistent code-object cache, and jobs are
    rendered in the order their kernels become ready, so compilation overlaps
    GPU work;
  * the sub-frames of a motion-blurred frame share one shape and go out as
    one batched launch (`Renderer.iterate_multi`);
  * two renderers (two HIP streams) alternate, so one job's host packing and
    launches overlap the previous job's kernels; buffers only grow
    (`p3_resize`), so a batch of mixed sizes allocates a few times at most.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from concurrent.futures import ProcessPoolExecutor, as_completed
from dataclasses import dataclass, field
from multiprocessing import get_context
from typing import List, Optional, Sequence

import numpy as np

from . import pack
from .genome import Affine, Genome

SUBFRAME_SEED_STEP = 0x9E3779B1  # src/animate-render.ts:151


@dataclass
class Job:
    genome: Genome
    width: int
    height: int
    oversample: int = 1
    quality: float = 50.0
    seed: int = 1
    subframes: Optional[List[Genome]] = None   # motion blur: same shape as genome
    name: str = ""


@dataclass
class JobResult:
    name: str
    samples: int
    walkers: int
    iters: int
    subframes: int
    image: Optional[np.ndarray] = None


def apportion(total: int, weights: Sequence[float]) -> List[int]:
    """Largest-remainder split of `total` walkers over `weights`
    (src/animate-render.ts:135-140, apportionWalkers)."""
    s = float(sum(max(0.0, w) for w in weights))
    if total <= 0 or s <= 0:
        return [0] * len(weights)
    quotas = [total * max(0.0, w) / s for w in weights]
    base = [int(math.floor(q)) for q in quotas]
    rest = total - sum(base)
    order = sorted(range(len(weights)), key=lambda i: (-(quotas[i] - base[i]), i))
    for i in order[:rest]:
        base[i] += 1
    return base


def _shape_key(g: Genome) -> tuple:
    e = pack.expand_genome_for_gpu(g)
    xs = list(e.xforms) + ([e.finalxform] if e.finalxform is not None else [])
    return tuple((tuple(v.index for v in x.variations if v.active), x.post is not None)
                 for x in xs) + ((e.finalxform is not None),)


def _compile_worker(xforms: np.ndarray, n_rows: int, final_slot: int, cache: str) -> int:
    """Worker process: compile one shape into the code-object cache."""
    os.environ["P3_KERNEL_CACHE"] = cache
    from . import _lib
    lib = _lib.load()
    xf = np.ascontiguousarray(xforms, dtype=np.float32)
    return lib.p3_chaos_prepare(xf.ctypes.data_as(C.POINTER(C.c_float)), xf.shape[0], n_rows,
                                final_slot)


def _ready(pg) -> bool:
    """Whether the shape's kernel is compiled (in this process or the cache)."""
    from . import _lib
    xf = np.ascontiguousarray(pg.xforms, dtype=np.float32)
    rc = _lib.load().p3_chaos_ready(xf.ctypes.data_as(C.POINTER(C.c_float)), xf.shape[0],
                                    pg.distrib.shape[0], pack.final_xform_slot(pg.expanded))
    if rc < 0:
        raise ValueError(f"invalid genome shape for the chaos kernel (code {rc})")
    return rc == 1


def precompile(genomes: Sequence[Genome], workers: Optional[int] = None,
               cache_dir: Optional[str] = None) -> int:
    """Compile the chaos kernels of every distinct shape among `genomes` into
    the code-object cache with a pool of worker processes; returns the number
    of shapes."""
    cache_dir = cache_dir or os.environ.get("P3_KERNEL_CACHE") or os.path.join(
        os.path.dirname(os.path.abspath(__file__)), "kcache")
    os.environ["P3_KERNEL_CACHE"] = cache_dir
    shapes = {}
    for g in genomes:
        shapes.setdefault(_shape_key(g), pack_for(g))
    workers = workers or max(1, min(16, (os.cpu_count() or 2) - 1))
    with ProcessPoolExecutor(max_workers=workers, mp_context=get_context("spawn")) as ex:
        futs = [ex.submit(_compile_worker, pg.xforms, pg.distrib.shape[0],
                          pack.final_xform_slot(pg.expanded), cache_dir) for pg in shapes.values()]
        for f in futs:
            if f.result() != 0:
                raise RuntimeError("kernel specialisation failed")
    return len(shapes)


def compile_pool(workers: Optional[int] = None) -> ProcessPoolExecutor:
    """A pool of compile worker processes, every worker started before it
    returns. Create it before this process first touches the GPU: the
    workers are spawned (fresh interpreters), never forked from a process
    holding a GPU context."""
    workers = workers or max(1, min(16, (os.cpu_count() or 2) - 1))
    ex = ProcessPoolExecutor(max_workers=workers, mp_context=get_context("spawn"))
    for f in [ex.submit(os.getpid) for _ in range(4 * workers)]:
        f.result()
    return ex


class BatchRenderer:
    def __init__(self, n_streams: int = 2, compile_workers: Optional[int] = None,
                 cache_dir: Optional[str] = None, pool: Optional[ProcessPoolExecutor] = None):
        """Compiles every distinct shape in worker processes and renders each
        job as soon as its kernel is ready."""
        from .renderer import Renderer
        self._Renderer = Renderer
        self.cache_dir = cache_dir or os.environ.get("P3_KERNEL_CACHE") or os.path.join(
            os.path.dirname(os.path.abspath(__file__)), "kcache")
        os.environ["P3_KERNEL_CACHE"] = self.cache_dir
        self.compile_workers = compile_workers or max(1, min(16, (os.cpu_count() or 2) - 1))
        self.n_streams = max(1, n_streams)
        self._workers = pool     # caller-owned compile pool (see compile_pool), or None
        self._pool: List = []
        self.timings = {}

    def _renderer(self, slot: int, w: int, h: int, os_: int):
        while len(self._pool) <= slot:
            self._pool.append(None)
        r = self._pool[slot]
        if r is None:
            r = self._Renderer(w, h, os_)
            self._pool[slot] = r
        else:
            r.resize(w, h, os_)
        return r

    def render(self, jobs: Sequence[Job], read_images: bool = False,
               progress=None) -> List[JobResult]:
        """Render every job; `progress(k)` (optional) is called after every
        64th job."""
        t0 = time.perf_counter()
        # distinct shapes -> compile jobs (worker processes, no GPU)
        packed = []
        shapes = {}
        for j in jobs:
            pg_list = [pack_for(g) for g in (j.subframes or [j.genome])]
            packed.append(pg_list)
            key = _shape_key(j.genome)
            shapes.setdefault(key, pg_list[0])
        order: List[int] = []
        by_key = {}
        for i, j in enumerate(jobs):
            by_key.setdefault(_shape_key(j.genome), []).append(i)
        results: List[Optional[JobResult]] = [None] * len(jobs)
        n_done = 0

        def render_shape(key):
            nonlocal n_done
            for i in by_key[key]:
                results[i] = self._render_one(jobs[i], packed[i], n_done, read_images)
                order.append(i)
                n_done += 1
                if progress and n_done % 64 == 0:
                    progress(n_done)

        # shapes already compiled (this process or the code-object cache)
        # render at once; the others go to the compile workers
        pending = {key: pg for key, pg in shapes.items() if not _ready(pg)}
        t_sub = time.perf_counter()
        for key in shapes:
            if key not in pending:
                render_shape(key)
        if pending:
            own = self._workers is None
            ex = (ProcessPoolExecutor(max_workers=self.compile_workers,
                                      mp_context=get_context("spawn")) if own else self._workers)
            try:
                futs = {}
                for key, pg in pending.items():
                    futs[ex.submit(_compile_worker, pg.xforms, pg.distrib.shape[0],
                                   pack.final_xform_slot(pg.expanded), self.cache_dir)] = key
                for f in as_completed(futs):
                    if f.result() != 0:
                        raise RuntimeError(f"kernel specialisation failed for shape {futs[f]}")
                    render_shape(futs[f])
            finally:
                if own:
                    ex.shutdown()
        for r in self._pool:
            if r is not None:
                r.sync()
        self.timings = {"total_s": time.perf_counter() - t0,
                        "pack_s": t_sub - t0, "shapes": len(shapes), "jobs": len(jobs)}
        return results  # type: ignore[return-value]

    def _render_one(self, job: Job, pgs, n: int, read_images: bool) -> JobResult:
        r = self._renderer(n % self.n_streams, job.width, job.height, job.oversample)
        walkers, iters = gpu_dispatch(job.quality, job.width, job.height)
        r.reset(job.genome)
        if job.subframes:
            k = len(job.subframes)
            counts = apportion(walkers, [1.0] * k)
            seeds = [(job.seed + i * SUBFRAME_SEED_STEP) & 0xFFFFFFFF for i in range(k)]
            live = [i for i in range(k) if counts[i] > 0]
            samples = r.iterate_multi([job.subframes[i] for i in live], [seeds[i] for i in live],
                                      [counts[i] for i in live], iters,
                                      packed=[pgs[i] for i in live])
        else:
            samples = r.iterate(job.genome, job.seed, walkers, iters, packed=pgs[0])
        r.present(job.genome, samples)
        img = r.read_image() if read_images else None
        return JobResult(job.name, samples, walkers, iters, len(job.subframes or [0]), img)

    def close(self):
        for r in self._pool:
            if r is not None:
                r.destroy()
        self._pool = []


def gpu_dispatch(quality: float, width: int, height: int) -> tuple:
    """(walkers, iters) for `quality` samples per output pixel on one MI355X:
    about 768 post-fuse turns per walker (the 200-turn fuse stays a small
    share), whole chaos blocks, at most one block per CU (the autotuned
    optimum at 1080p)."""
    blk = pack.CHAOS_BLOCK
    samples = max(1, int(round(quality * width * height)))
    walkers = min(pack.CHIP_WALKERS, max(blk, samples // 768))
    walkers = max(blk, walkers // blk * blk)
    iters = -(-samples // walkers)
    return walkers, iters


def pack_for(g: Genome):
    from .renderer import pack_genome
    return pack_genome(g)


# ------------------------------------------------------------ workload ---

def _perturbed(g: Genome, rng: np.random.Generator, amount: float) -> Genome:
    """A second keyframe of the same shape: affine coefficients jittered."""
    import copy
    h = copy.deepcopy(g)
    for x in h.xforms:
        a = x.affine
        x.affine = Affine(*(v + float(rng.normal(0.0, amount)) for v in
                            (a.a, a.b, a.c, a.d, a.e, a.f)))
    return h


def _lerp_genome(g0: Genome, g1: Genome, t: float) -> Genome:
    """Sub-frame genome between two keyframes of one shape (affine lerp)."""
    import copy
    h = copy.deepcopy(g0)
    for x, x1 in zip(h.xforms, g1.xforms):
        a, b = x.affine, x1.affine
        x.affine = Affine(*((1 - t) * u + t * v for u, v in
                            zip((a.a, a.b, a.c, a.d, a.e, a.f), (b.a, b.b, b.c, b.d, b.e, b.f))))
    return h


def batch_workload(n: int = 512, seed: int = 3000, motion_fraction: float = 0.5,
                   subframes: int = 8) -> List[Job]:
    """The batch workload: n seeded genomes, long edge U{128..1024}, 2..16
    xforms from the flam3 core set, quality U(20, 200), the first
    `motion_fraction` of them motion-blurred with `subframes` sub-frames."""
    from .synth import flam3_core_pool, make_genome
    rng = np.random.default_rng(seed)
    pool = flam3_core_pool()
    jobs = []
    n_motion = int(round(n * motion_fraction))
    for i in range(n):
        long_edge = int(rng.integers(128, 1025))
        aspect = float(rng.uniform(0.5, 1.0))
        if rng.random() < 0.5:
            w, h = long_edge, max(16, int(round(long_edge * aspect)))
        else:
            w, h = max(16, int(round(long_edge * aspect))), long_edge
        n_xf = int(rng.integers(2, 17))
        counts = [int(c) for c in rng.integers(1, 4, n_xf)]
        q = float(rng.uniform(20, 200))
        g = make_genome(rng, n_xf, pool, counts, w, h, q, 1, post_fraction=0.3,
                        lognormal_sigma=0.5, name=f"batch-{i}")
        subs = None
        if i < n_motion:
            g1 = _perturbed(g, rng, 0.02)
            subs = [_lerp_genome(g, g1, (k + 0.5) / subframes) for k in range(subframes)]
        jobs.append(Job(g, w, h, 1, q, seed=int(rng.integers(0, 2 ** 32)), subframes=subs,
                        name=g.name))
    return jobs
