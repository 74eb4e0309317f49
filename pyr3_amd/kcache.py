"""Ahead-of-time compilation of the specialised chaos kernels.

Every genome shape (per xform: variation count, post-affine presence and the
catalog index of each slot; the final xform's slot and shape) runs its own
chaos kernel, compiled with hipRTC on first use and kept in the code-object
cache (``P3_KERNEL_CACHE``, by default ``pyr3_amd/kcache/``). hipRTC needs no
GPU, so the shapes a workload needs can be compiled on the build host by a
pool of worker processes and shipped with the tree instead of being compiled
one after another inside the first run.

A shape is named by the library's shape key, the text the C ABI keys its
module cache by (``p3::chaos_shape_key``):

    <nv><p|n>,<idx>,...;   once per regular xform, then
    F[<final slot>:<nv><p|n>,<idx>,...;]

The generated kernel source depends on nothing else, so a record blob with
those counts, flags and indices (zero weights and parameters) compiles to the
same code object as any genome of that shape. A process started with
``P3_SHAPE_LOG=<file>`` appends the key of each shape it compiles or loads,
which is how ``kernel_shapes.txt`` (the shapes of the smoke test, the GPU
tests and the benchmark) is collected. A batch's union kernel (several
shapes in one kernel, p3_batch) is keyed ``U|<shape>|<shape>|...`` with its
shapes in case order.

``python -m pyr3_amd.kcache [manifest ...]`` compiles the manifest (default
``kernel_shapes.txt``) into the cache.
"""
from __future__ import annotations

import ctypes as C
import hashlib
import os
import shutil
import sys
import time
from concurrent.futures import ProcessPoolExecutor
from multiprocessing import get_context
from pathlib import Path
from typing import Iterable, List, Tuple

import numpy as np

PKG = Path(__file__).resolve().parent
MANIFEST = PKG / "kernel_shapes.txt"
CACHE = PKG / "kcache"
STAMP = ".source-stamp"
XFORM_FLOATS = 116


def _parse_xform(text: str) -> Tuple[int, bool, List[int]]:
    head, *idx = text.split(",")
    nv, post = int(head[:-1]), head[-1] == "p"
    if head[-1] not in "pn" or len(idx) != nv:
        raise ValueError(f"bad xform shape {text!r}")
    return nv, post, [int(i) for i in idx]


def _record(nv: int, post: bool, idx: List[int]) -> np.ndarray:
    rec = np.zeros(XFORM_FLOATS, np.float32)
    rec[7] = nv
    rec[15] = 1.0 if post else 0.0
    for k, i in enumerate(idx):
        rec[20 + 4 * k] = i
    return rec


def blobs_from_key(key: str) -> Tuple[np.ndarray, int, int]:
    """(xform records, n_rows, final slot) of a record blob with this shape."""
    body, sep, fin = key.rpartition("F")
    if not sep:
        raise ValueError(f"bad shape key {key!r}")
    regular = [_parse_xform(t) for t in body.split(";") if t]
    recs = [_record(*x) for x in regular]
    final_slot = -1
    if fin:
        slot, _, shape = fin.partition(":")
        final_slot = int(slot)
        if final_slot != len(regular):
            raise ValueError(f"final slot {final_slot} != {len(regular)} regular xforms")
        recs.append(_record(*_parse_xform(shape.rstrip(";"))))
    return np.ascontiguousarray(np.stack(recs)), len(regular) + 1, final_slot


def key_of(xforms: np.ndarray, n_rows: int, final_slot: int) -> str:
    """The shape key of a record blob (Python twin of p3::chaos_shape_key)."""
    def one(rec):
        nv = int(rec[7]) if rec[7] < 8 else 8
        return (f"{nv}{'p' if rec[15] != 0 else 'n'}"
                + "".join(f",{int(rec[20 + 4 * k])}" for k in range(nv)) + ";")
    out = "".join(one(xforms[j]) for j in range(n_rows - 1)) + "F"
    if final_slot >= 0:
        out += f"{final_slot}:" + one(xforms[final_slot])
    return out


def read_manifest(paths: Iterable[Path]) -> List[str]:
    keys, seen = [], set()
    for p in paths:
        for line in Path(p).read_text().splitlines():
            line = line.strip()
            if line and not line.startswith("#") and line not in seen:
                seen.add(line)
                keys.append(line)
    return keys


def _compile_key(key: str, cache: str) -> Tuple[str, int, float]:
    """Worker process: compile one shape (or one union kernel, key
    ``U|<shape>|<shape>...``) into the cache."""
    os.environ["P3_KERNEL_CACHE"] = cache
    os.environ.pop("P3_SHAPE_LOG", None)
    from . import _lib
    lib = _lib.load()
    t0 = time.perf_counter()
    if key.startswith("U|"):
        parts = [blobs_from_key(k) for k in key.split("|")[1:]]
        n = len(parts)
        xs = [np.ascontiguousarray(p[0]) for p in parts]
        arr = (C.POINTER(C.c_float) * n)(*[x.ctypes.data_as(C.POINTER(C.c_float)) for x in xs])
        rc = lib.p3_union_prepare(n, arr, (C.c_int * n)(*[x.shape[0] for x in xs]),
                                  (C.c_int * n)(*[p[1] for p in parts]),
                                  (C.c_int * n)(*[p[2] for p in parts]))
        return key, rc, time.perf_counter() - t0
    xf, n_rows, final_slot = blobs_from_key(key)
    rc = lib.p3_chaos_prepare(xf.ctypes.data_as(C.POINTER(C.c_float)), xf.shape[0], n_rows,
                              final_slot)
    return key, rc, time.perf_counter() - t0


def source_stamp() -> str:
    """Hash of everything the generated kernel sources and their compile
    options come from: a change of it invalidates the whole cache."""
    h = hashlib.sha256()
    for name in ("p3_rtc_source.inc", "p3_rtc.cpp", "p3_var_draws.h"):
        h.update((PKG / "csrc" / name).read_bytes())
    return h.hexdigest()[:16]


def prune_if_stale(cache: Path = CACHE) -> bool:
    """Empty the cache when it was filled from other kernel sources (its code
    objects could never be hit again); returns True when it was emptied."""
    stamp = source_stamp()
    f = cache / STAMP
    if f.exists() and f.read_text().strip() == stamp:
        return False
    emptied = cache.exists() and any(cache.glob("*.co"))
    if cache.exists():
        shutil.rmtree(cache)
    cache.mkdir(parents=True)
    f.write_text(stamp + "\n")
    return emptied


def warm(keys: List[str], workers: int = 0, cache: Path = CACHE, verbose: bool = True) -> int:
    """Compile every shape of `keys` that is not cached yet; returns the number
    compiled. Worker processes are spawned (fresh interpreters, no GPU)."""
    cache.mkdir(parents=True, exist_ok=True)
    workers = workers or max(1, min(16, os.cpu_count() or 1))
    t0 = time.perf_counter()
    done = 0
    with ProcessPoolExecutor(max_workers=workers, mp_context=get_context("spawn")) as ex:
        futs = [ex.submit(_compile_key, k, str(cache)) for k in keys]
        for f in futs:
            key, rc, dt = f.result()
            if rc != 0:
                raise RuntimeError(f"kernel specialisation failed for shape {key}")
            done += 1
    if verbose:
        print(f"[pyr3_amd.kcache] {done} shapes ready in {cache} "
              f"({time.perf_counter() - t0:.1f}s, {workers} workers)", flush=True)
    return done


def main(argv: List[str]) -> int:
    paths = [Path(a) for a in argv] or [MANIFEST]
    prune_if_stale()
    warm(read_manifest(paths))
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
