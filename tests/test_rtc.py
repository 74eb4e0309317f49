"""Run-time chaos specialisation compiles for gfx950 without a device.

The chaos kernel is generated per genome shape and compiled with hipRTC; the
compile needs no GPU, so this runs on the build host: one shape holding every
implemented catalog entry (one xform per entry, pre_blur included, a post
affine on every other xform, a final xform), and the benchmark genome.
"""
import ctypes as C
import time

import numpy as np
import pytest

from helpers import genome, xform
from pyr3_amd import pack, synth
from pyr3_amd.variations_supported import SUPPORTED


def _prepare(g):
    from pyr3_amd import _lib
    lib = _lib.load()
    e = pack.expand_genome_for_gpu(g)
    xf = np.ascontiguousarray(pack.pack_xforms(e))
    n_rows = len(e.xforms) + 1
    fp = xf.ctypes.data_as(C.POINTER(C.c_float))
    rc = lib.p3_chaos_prepare(fp, xf.shape[0], n_rows, pack.final_xform_slot(e))
    n = C.c_size_t()
    lib.p3_chaos_source(fp, xf.shape[0], n_rows, pack.final_xform_slot(e), None, 0, C.byref(n))
    buf = C.create_string_buffer(n.value + 1)
    lib.p3_chaos_source(fp, xf.shape[0], n_rows, pack.final_xform_slot(e), buf, n.value + 1, None)
    return rc, buf.value.decode(), lib.p3_last_error(None)


def test_catalog_sample_compiles():
    """Every 9th catalog entry plus all direct-colour entries in one shape
    (the GPU tests compile and run the whole catalog)."""
    from pyr3_amd.catalog import DC_VARIATIONS
    idx = sorted(set(sorted(SUPPORTED)[::9]) | set(DC_VARIATIONS) | {97})
    xs = [xform(1.0, vars=((i, 1.0),), post=(1, 0, 0, 0, 1, 0) if k % 2 else None)
          for k, i in enumerate(idx)]
    g = genome(xs, final=xform(0.0, vars=((13, 1.0), (97, 0.1))))
    t0 = time.time()
    rc, src, err = _prepare(g)
    assert rc == 0, err
    for i in idx:
        if i != 97:
            assert f"apply_variation({i}, pam" in src
    assert "pre_blur_slot" in src and "post_affine(xf, pv)" in src
    assert "dc_color_for(" in src
    print(f"compiled {len(idx)} entries in {time.time() - t0:.1f}s")


def test_benchmark_genome_compiles():
    rc, src, err = _prepare(synth.bench_genome(0))
    assert rc == 0, err
    assert src.count("case ") >= 12


def test_unsupported_entry_is_rejected():
    rc, _, _ = _prepare(genome([xform(1.0, vars=((326, 1.0),))]))
    assert rc != 0


def _source(xf, n_rows, final_slot):
    from pyr3_amd import _lib
    lib = _lib.load()
    xf = np.ascontiguousarray(xf, dtype=np.float32)
    fp = xf.ctypes.data_as(C.POINTER(C.c_float))
    n = C.c_size_t()
    assert lib.p3_chaos_source(fp, xf.shape[0], n_rows, final_slot, None, 0, C.byref(n)) == 0
    buf = C.create_string_buffer(n.value + 1)
    lib.p3_chaos_source(fp, xf.shape[0], n_rows, final_slot, buf, n.value + 1, None)
    return buf.value


@pytest.mark.parametrize("k", range(3))
def test_shape_key_rebuilds_the_same_kernel(k):
    """A record blob rebuilt from a genome's shape key (counts, post flags,
    slot indices; zero weights) generates the genome's kernel source byte for
    byte, so compiling the key ahead of time fills the cache entry the
    genome's dispatch looks up. Genomes with a final xform, post-affines,
    several slots and pre_blur."""
    from pyr3_amd import kcache
    from pyr3_amd.renderer import pack_genome
    g = [synth.bench_genome(0), synth.config2_genome(1),
         genome([xform(1.0, vars=((97, 0.2), (4, 0.8)), post=(1, 0, 0, 0, 1, 0)),
                 xform(0.5, vars=((13, 1.0),))], final=xform(0.0, vars=((2, 1.0),)))][k]
    pg = pack_genome(g)
    fs = pack.final_xform_slot(pg.expanded)
    key = kcache.key_of(pg.xforms, pg.distrib.shape[0], fs)
    xf, n_rows, fs2 = kcache.blobs_from_key(key)
    assert (n_rows, fs2) == (pg.distrib.shape[0], fs)
    assert kcache.key_of(xf, n_rows, fs2) == key
    assert _source(xf, n_rows, fs2) == _source(pg.xforms, n_rows, fs)


def test_manifest_keys_parse():
    """Every line of the shipped manifest names a valid shape of implemented
    entries, or a union kernel (``U|shape|shape...``) of such shapes."""
    from pyr3_amd import kcache
    keys = kcache.read_manifest([kcache.MANIFEST])
    assert len(keys) >= 10
    assert any(k.startswith("U|") for k in keys)
    for line in keys:
        parts = line.split("|")[1:] if line.startswith("U|") else [line]
        assert parts
        for key in parts:
            xf, n_rows, fs = kcache.blobs_from_key(key)
            assert kcache.key_of(xf, n_rows, fs) == key
            for rec in xf:
                for s in range(int(min(rec[7], 8))):
                    assert int(rec[20 + 4 * s]) in SUPPORTED


def test_chaos_block_matches_the_host_schedule():
    """The host's walker schedules (whole chaos blocks, one block per CU) use
    the kernel's block size, and that block's LDS image fits one CU."""
    import re
    from conftest import ROOT
    src = (ROOT / "pyr3_amd" / "csrc" / "p3_args.h").read_text()
    blk = int(re.search(r"kChaosBlock = (\d+);", src).group(1))
    assert pack.CHAOS_BLOCK == blk and blk % 64 == 0
    assert pack.CHIP_WALKERS == 256 * blk
    # 192 B of ISAAC words, 36 B of walker scalars, 8 B of splat sums per
    # walker, plus the palette, the pick boundaries, two bucket rows and two
    # live counts and two probe sums
    lds = 236 * blk + 256 * 16 + 2 * 17 * 16 + 2 * 4 * (128 + 2) + 16
    assert lds <= 160 * 1024
    assert lds <= 160 * 1024
