"""Whole renders of flam3 genomes against flam3-C's own images.

Each fixture under tests/golden/flam3 is one of the reference's parity
corpus entries (fixtures/flam3-goldens/<id>: the Electric Sheep `.flam3`,
flam3-C's `golden.png`, `meta.json`). The genome goes through the `.flam3`
importer and the HIP render path (chaos at 131,072 walkers, DE, spatial
filter, tonemap) at the flame's own size, quality and supersample; the bar
is the reference's: R = mean |difference| over all RGBA bytes
(src/compare.ts:15-28) at most `thresholdR` (its own renders' R + 1, the
gate of src/parity.test.ts:139-150). The render seed is fixed; flam3-C's
and the reference's streams differ from it, so R measures the same noise
floor the reference's threshold was calibrated on.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from pyr3_amd import pack
from pyr3_amd.flam3 import load_genome

pytestmark = pytest.mark.gpu

FIXTURES = sorted(d.name for d in (GOLDEN / "flam3").iterdir() if d.is_dir())
WALKERS = 131072


def mean_abs_diff_rgba(a: np.ndarray, b: np.ndarray) -> float:
    assert a.shape == b.shape and a.shape[-1] == 4
    return float(np.abs(a.astype(np.int32) - b.astype(np.int32)).mean())


@pytest.mark.parametrize("fid", FIXTURES)
def test_flam3_golden_within_reference_threshold(fid):
    import sys
    sys.path.insert(0, str(GOLDEN))
    from png8 import read_png
    from pyr3_amd.renderer import Renderer
    d = GOLDEN / "flam3" / fid
    meta = json.loads((d / "meta.json").read_text())
    g, kind, rep = load_genome(str(d / f"{fid}.flam3"))
    assert kind == "flame" and not rep.dropped_variations
    W, H = g.size["width"], g.size["height"]
    os_ = g.oversample or 1
    radius = g.filter_radius if g.filter_radius is not None else pack.DEFAULT_FILTER_RADIUS
    print(f"[flam3 golden] {fid}: {W}x{H} os{os_} q{g.quality} xforms {len(g.xforms)}", flush=True)
    r = Renderer(W, H, os_, filter_radius=radius)
    r.render(g, seed=0x9E3779B9, walkers=WALKERS)
    img = r.read_image()
    r.destroy()
    gold = read_png(d / "golden.png")
    R = mean_abs_diff_rgba(img, gold)
    print(f"[flam3 golden] {fid}: R={R:.4f} (reference expectedR {meta['expectedR']}, "
          f"threshold {meta['thresholdR']})", flush=True)
    assert R <= meta["thresholdR"], (fid, R, meta["thresholdR"])


def test_render_cli_writes_png_with_genome(tmp_path):
    """python -m pyr3_amd.render on one fixture: the PNG decodes to a frame
    within the fixture's threshold and carries the genome it was made from."""
    import sys
    sys.path.insert(0, str(GOLDEN))
    from png8 import read_png
    from pyr3_amd.genome import genome_to_json
    from pyr3_amd.render import main
    fid = "244.57686"
    d = GOLDEN / "flam3" / fid
    out = tmp_path / "out.png"
    # the reference's parity rig renders with auto-exposure off
    assert main([str(d / f"{fid}.flam3"), str(out), "--seed", "7", "--no-auto-exposure"]) == 0
    img = read_png(out)
    meta = json.loads((d / "meta.json").read_text())
    assert mean_abs_diff_rgba(img, read_png(d / "golden.png")) <= meta["thresholdR"]
    g_png = load_genome(str(out))[0]
    assert genome_to_json(g_png) == genome_to_json(load_genome(str(d / f"{fid}.flam3"))[0])
