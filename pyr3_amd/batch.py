"""Batched rendering of many genomes (the 512-genome batch workload).

A batch is a list of jobs, each one output frame: a genome at a size and
quality, optionally a motion-blurred frame of N sub-frames (the sub-render
loop of src/animate-render.ts:73-171: walkers apportioned over the temporal
filter, sub-frame seed = seed + i * 0x9e3779b1, one present).

The MI355X-side design (p3_batch, pyr3_amd/csrc/p3_batch.cpp):
  * the jobs are split, in order, into launch groups of at most
    `max_shapes` distinct genome shapes; each group's shapes are compiled
    into ONE union chaos kernel (every shape's xform cases in one switch,
    selected per block by its sub-dispatch), by worker processes (hipRTC
    needs no GPU) into the persistent code-object cache;
  * each group is one persistent launch over all its jobs' sub-dispatches
    (motion-blur sub-frames included), each writing its own histogram, so
    many small genomes fill the chip together;
  * the whole batch (histogram fills, chaos launches, packed merges, DE and
    tonemap per frame) is captured once as a HIP graph over a few streams and
    replayed;
  * walkers per genome (through the post-fuse turns per walker) and the
    graph streams are picked by `autotune`, which times candidate plans on
    a sample of the batch.
"""
from __future__ import annotations

import ctypes as C
import math
import os
import time
from concurrent.futures import ProcessPoolExecutor
from dataclasses import dataclass
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


@dataclass
class BatchPlan:
    """Launch choices of a batch: post-fuse turns per walker (walkers per
    genome = its samples / turns, in whole chaos blocks per sub-dispatch),
    distinct genome shapes per union kernel, and graph streams."""
    turns: int = 1024
    max_shapes: int = 16
    lanes: int = 2


def job_dispatch(job: Job, turns: int) -> tuple:
    """(walkers, iters) of a job: every sub-dispatch gets the same whole
    number of chaos blocks, sized so each walker runs about `turns` post-fuse
    turns (few, long walkers keep the 200-turn fuse a small share; the batch
    fills the chip with many genomes at once), at most one chip's worth."""
    blk = pack.CHAOS_BLOCK
    k = len(job.subframes) if job.subframes else 1
    samples = max(1, int(round(job.quality * job.width * job.height * job.oversample ** 2)))
    per_sub = -(-samples // (k * max(1, turns)))
    per_sub = min(pack.CHIP_WALKERS, max(blk, -(-per_sub // blk) * blk))
    walkers = per_sub * k
    return walkers, -(-samples // walkers)


def _frame_blobs(job: Job, pgs, turns: int, fuse: int):
    """Per-sub-dispatch uniform blocks, walker count and samples of a job."""
    walkers, iters = job_dispatch(job, turns)
    k = len(pgs)
    counts = apportion(walkers, [1.0] * k)
    seeds = [(job.seed + i * SUBFRAME_SEED_STEP) & 0xFFFFFFFF for i in range(k)] \
        if job.subframes else [job.seed]
    sw, sh = job.width * job.oversample, job.height * job.oversample
    us = [np.ascontiguousarray(pack.chaos_uniforms(pg.expanded, sw, sh, job.oversample, iters,
                                                   fuse, sd, n, pack.DEFAULT_WALKER_JITTER,
                                                   False, None))
          for pg, sd, n in zip(pgs, seeds, counts)]
    return us, walkers, iters


def _present_blobs(job: Job, samples: int):
    """DE (or None) and tonemap uniform blocks of a job (Renderer.present)."""
    from . import curves
    from .genome import Tonemap
    g = job.genome
    if curves.active_mask(g.channel_curves):
        raise ValueError(f"{job.name}: channel curves are not available in a batch")
    tm = g.tonemap or Tonemap()
    k1, k2 = pack.calibration(g.scale, samples, tm.brightness, job.oversample)
    sw, sh = job.width * job.oversample, job.height * job.oversample
    de = None
    if g.density is not None:
        d = g.density
        de = np.ascontiguousarray(pack.de_uniforms(sw, sh, d.max_rad, d.min_rad, d.curve, k1, k2,
                                                   job.oversample))
    taps = np.ascontiguousarray(pack.spatial_kernel(pack.DEFAULT_FILTER_RADIUS, job.oversample))
    viz = np.ascontiguousarray(pack.viz_uniforms(job.width, job.height, k1, k2, tm,
                                                 job.oversample, len(taps),
                                                 g.background or (0.0, 0.0, 0.0), False, 0,
                                                 g.hsl_adjust))
    return de, viz, taps


def plan_groups(pgs_per_job, max_shapes: int) -> List[List[int]]:
    """Union kernels of a batch: for each launch group, the job index of
    each distinct shape's first job, in case order -- the grouping
    p3_batch_commit makes (jobs of at most 16 xforms in order, then the
    others; at most max_shapes shapes per group)."""
    from .kcache import key_of
    keys = [key_of(p[0].xforms, p[0].distrib.shape[0], pack.final_xform_slot(p[0].expanded))
            for p in pgs_per_job]
    wide = [p[0].distrib.shape[0] - 1 > 16 for p in pgs_per_job]
    order = [i for i in range(len(keys)) if not wide[i]] + [i for i in range(len(keys)) if wide[i]]
    groups: List[List[int]] = []
    seen: dict = {}
    cur: List[int] = []
    cur_wide = None
    for i in order:
        full = keys[i] not in seen and len(seen) >= max_shapes
        if cur and (full or wide[i] != cur_wide):
            groups.append(cur)
            cur, seen = [], {}
        if not cur:
            cur_wide = wide[i]
        if keys[i] not in seen:
            seen[keys[i]] = len(seen)
            cur.append(i)
    if cur:
        groups.append(cur)
    return groups


def union_key(pgs_per_job, group: List[int]) -> str:
    """Text key of a group's union kernel (the shapes in case order)."""
    from .kcache import key_of
    return "U|" + "|".join(key_of(pgs_per_job[i][0].xforms, pgs_per_job[i][0].distrib.shape[0],
                                  pack.final_xform_slot(pgs_per_job[i][0].expanded))
                           for i in group)


def _union_worker(blobs, cache: str) -> int:
    """Worker process: compile one union kernel into the code-object cache."""
    os.environ["P3_KERNEL_CACHE"] = cache
    os.environ.pop("P3_SHAPE_LOG", None)
    from . import _lib
    lib = _lib.load()
    xs = [np.ascontiguousarray(x, dtype=np.float32) for x, _, _ in blobs]
    n = len(xs)
    arr = (C.POINTER(C.c_float) * n)(*[x.ctypes.data_as(C.POINTER(C.c_float)) for x in xs])
    ns = (C.c_int * n)(*[x.shape[0] for x in xs])
    nr = (C.c_int * n)(*[b[1] for b in blobs])
    fi = (C.c_int * n)(*[b[2] for b in blobs])
    return lib.p3_union_prepare(n, arr, ns, nr, fi)


def precompile_unions(pgs_per_job, max_shapes: int, workers: Optional[int] = None,
                      cache_dir: Optional[str] = None, pool=None) -> int:
    """Compile every launch group's union kernel into the code-object cache
    with worker processes (hipRTC needs no GPU); returns the group count."""
    cache_dir = cache_dir or os.environ.get("P3_KERNEL_CACHE") or os.path.join(
        os.path.dirname(os.path.abspath(__file__)), "kcache")
    os.environ["P3_KERNEL_CACHE"] = cache_dir
    groups = plan_groups(pgs_per_job, max_shapes)
    todo = []
    for grp in groups:
        blobs = [(pgs_per_job[i][0].xforms, pgs_per_job[i][0].distrib.shape[0],
                  pack.final_xform_slot(pgs_per_job[i][0].expanded)) for i in grp]
        todo.append(blobs)
    workers = workers or max(1, min(16, (os.cpu_count() or 2)))
    ex = pool or ProcessPoolExecutor(max_workers=workers, mp_context=get_context("spawn"))
    try:
        for f in [ex.submit(_union_worker, b, cache_dir) for b in todo]:
            if f.result() != 0:
                raise RuntimeError("union kernel compile failed")
    finally:
        if pool is None:
            ex.shutdown()
    return len(groups)


class GraphBatch:
    """A committed batch (p3_batch): every job's frame rendered by one
    replay of a captured HIP graph -- per launch group one persistent chaos
    launch over all its genomes' sub-dispatches (a union kernel of their
    shapes), the packed merge, then DE and tonemap per frame."""

    def __init__(self, jobs: Sequence[Job], plan: BatchPlan, packed=None, fuse: int = pack.FUSE):
        from ._lib import check, load
        self._lib = load()
        self._check = check
        self._h = C.c_void_p()
        check(self._lib.p3_batch_create(C.byref(self._h)), None)
        self.jobs = list(jobs)
        self.plan = plan
        self.samples: List[int] = []
        self.walkers: List[int] = []
        self.iters: List[int] = []
        packed = packed or [[pack_for(g) for g in (j.subframes or [j.genome])] for j in jobs]
        try:
            for j, pgs in zip(self.jobs, packed):
                us, walkers, iters = _frame_blobs(j, pgs, plan.turns, fuse)
                samples = walkers * iters
                de, viz, taps = _present_blobs(j, samples)
                pal = np.ascontiguousarray(pack.pack_palette(j.genome).reshape(-1))
                n = len(pgs)
                xfs = [np.ascontiguousarray(p.xforms) for p in pgs]
                dss = [np.ascontiguousarray(p.distrib) for p in pgs]
                xfp = (C.POINTER(C.c_float) * n)(*[_fp(x) for x in xfs])
                dsp = (C.POINTER(C.c_uint32) * n)(*[_up(d) for d in dss])
                ufp = (C.POINTER(C.c_uint32) * n)(*[_up(u) for u in us])
                rc = self._lib.p3_batch_add(self._h, j.width, j.height, j.oversample, _fp(taps),
                                            len(taps), _fp(pal), n, xfp, xfs[0].shape[0], dsp,
                                            dss[0].shape[0], ufp,
                                            _up(de) if de is not None else None, _up(viz))
                if rc < 0:
                    self._raise(rc)
                self.samples.append(samples)
                self.walkers.append(walkers)
                self.iters.append(iters)
            rc = self._lib.p3_batch_commit(self._h, plan.max_shapes, plan.lanes)
            if rc:
                self._raise(rc)
        except Exception:
            self.close()
            raise

    def _raise(self, rc: int):
        from ._lib import P3Error
        msg = self._lib.p3_batch_last_error(self._h)
        raise P3Error(rc, msg.decode() if msg else "")

    def _ok(self, rc: int):
        if rc:
            self._raise(rc)

    def launch(self) -> None:
        self._ok(self._lib.p3_batch_launch(self._h))

    def sync(self) -> None:
        self._ok(self._lib.p3_batch_sync(self._h))

    def elapsed_ms(self) -> float:
        ms = C.c_float()
        self._ok(self._lib.p3_batch_elapsed(self._h, C.byref(ms)))
        return float(ms.value)

    def stats(self) -> dict:
        out = np.zeros(4, np.uint32)
        self._ok(self._lib.p3_batch_stats(self._h, _up(out)))
        return {"groups": int(out[0]), "shapes": int(out[1]), "sub_dispatches": int(out[2]),
                "chaos_blocks": int(out[3])}

    def read_image(self, i: int) -> np.ndarray:
        j = self.jobs[i]
        out = np.empty((j.height, j.width, 4), np.uint8)
        self._ok(self._lib.p3_batch_read_image(self._h, i, _bp(out), out.size))
        return out

    def read_histogram(self, i: int) -> np.ndarray:
        j = self.jobs[i]
        out = np.empty((j.height * j.oversample, j.width * j.oversample, 4), np.uint32)
        self._ok(self._lib.p3_batch_read_histogram(self._h, i, _up(out), out.size))
        return out

    def close(self) -> None:
        if self._h:
            self._lib.p3_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _up(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _bp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


TURN_CANDIDATES = (256, 512, 1024, 2048, 4096)
LANE_CANDIDATES = (1, 2, 4)


def autotune(jobs: Sequence[Job], packed=None, max_shapes: int = 16, sample: int = 64,
             repeats: int = 3, turns=TURN_CANDIDATES, lanes=LANE_CANDIDATES, log=None) -> tuple:
    """Pick the batch plan by measurement: the first `sample` jobs (whole
    launch groups of the full batch, so their union kernels are the batch's
    own) are committed at every candidate (turns per walker, then graph
    streams at the best turns) and replayed `repeats` times; the plan with
    the fewest device ms per sampled genome wins. Returns (plan, table of
    ms per candidate)."""
    packed = packed or [[pack_for(g) for g in (j.subframes or [j.genome])] for j in jobs]
    n = min(len(jobs), sample)
    sub, sub_p = list(jobs[:n]), packed[:n]
    table = {}

    def measure(plan: BatchPlan) -> float:
        gb = GraphBatch(sub, plan, packed=sub_p)
        try:
            gb.launch()
            gb.sync()
            t = []
            for _ in range(repeats):
                gb.launch()
                t.append(gb.elapsed_ms())
            return float(np.median(t))
        finally:
            gb.close()

    best = None
    for tv in turns:
        ms = measure(BatchPlan(tv, max_shapes, 2))
        table[f"turns={tv},lanes=2"] = round(ms, 3)
        if log:
            log(f"autotune turns={tv} lanes=2: {ms:.2f} ms for {n} genomes")
        if best is None or ms < best[0]:
            best = (ms, BatchPlan(tv, max_shapes, 2))
    for lv in lanes:
        if lv == 2:
            continue
        plan = BatchPlan(best[1].turns, max_shapes, lv)
        ms = measure(plan)
        table[f"turns={plan.turns},lanes={lv}"] = round(ms, 3)
        if log:
            log(f"autotune turns={plan.turns} lanes={lv}: {ms:.2f} ms")
        if ms < best[0]:
            best = (ms, plan)
    return best[1], table


class BatchRenderer:
    """Renders a list of jobs through one GraphBatch: the union kernels of
    the batch's launch groups are compiled up front by worker processes (or
    found in the code-object cache), then the batch is committed, launched
    once and read back."""

    def __init__(self, plan: Optional[BatchPlan] = None, compile_workers: Optional[int] = None,
                 cache_dir: Optional[str] = None, pool: Optional[ProcessPoolExecutor] = None):
        self.plan = plan or BatchPlan()
        self.cache_dir = cache_dir or os.environ.get("P3_KERNEL_CACHE") or os.path.join(
            os.path.dirname(os.path.abspath(__file__)), "kcache")
        os.environ["P3_KERNEL_CACHE"] = self.cache_dir
        self.compile_workers = compile_workers or max(1, min(16, (os.cpu_count() or 2)))
        self._workers = pool
        self.timings = {}

    def render(self, jobs: Sequence[Job], read_images: bool = False) -> List[JobResult]:
        t0 = time.perf_counter()
        packed = [[pack_for(g) for g in (j.subframes or [j.genome])] for j in jobs]
        t1 = time.perf_counter()
        groups = precompile_unions(packed, self.plan.max_shapes, self.compile_workers,
                                   self.cache_dir, self._workers)
        t2 = time.perf_counter()
        gb = GraphBatch(jobs, self.plan, packed=packed)
        try:
            t3 = time.perf_counter()
            gb.launch()
            gb.sync()
            t4 = time.perf_counter()
            res = []
            for i, j in enumerate(jobs):
                img = gb.read_image(i) if read_images else None
                res.append(JobResult(j.name, gb.samples[i], gb.walkers[i], gb.iters[i],
                                     len(j.subframes or [0]), img))
            self.timings = {"pack_s": t1 - t0, "compile_s": t2 - t1, "commit_s": t3 - t2,
                            "render_s": t4 - t3, "gpu_ms": gb.elapsed_ms(), "groups": groups,
                            **gb.stats()}
        finally:
            gb.close()
        return res

    def close(self):
        pass


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
