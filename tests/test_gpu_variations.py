"""Every implemented catalog entry on the GPU vs the oracle's f32 restatement.

One xform carrying the entry (weight 1, the reference test's injected
parameters, affine scaled x2 so inputs span [-2, 2]^2), one turn per walker,
4,096 walkers. The trace gives the pre-post-affine output pv_pre of turn 0
and the walker's draw counter: draws must match exactly (same ISAAC stream,
same consumption), and pv_pre within 1e-5 relative (scale max(|v|, 1)) on at
least 99.5% of walkers; the rest may only differ where the f32 evaluation is
ill-conditioned (tan/sin of a huge argument), which the bound caps.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import blobs, genome, xform
from pyr3_amd.variations_supported import SUPPORTED

pytestmark = pytest.mark.gpu

CASES = {c["index"]: c for c in json.loads((GOLDEN / "variation_cases.json").read_text()).values()}
INDICES = sorted(i for i in SUPPORTED if i != 97)


def _genome(idx):
    case = CASES.get(idx)
    params = case["params"] if case else [0.0] * 10
    aff = (2.0, 0.1, 0.05, -0.1, 2.0, -0.05)
    if case and case["runner"] == "runWithAffine" and idx in (16, 18, 30, 31):
        a, b, c, d, e, f = case["affine"]
        aff = (2.0, b, c, -0.1, 2.0, f)
    return genome([xform(1.0, aff, vars=((idx, 1.0),), params=[params])], scale=20.0)


@pytest.fixture(scope="module")
def R():
    from pyr3_amd.renderer import Renderer
    return Renderer


@pytest.mark.parametrize("idx", INDICES)
def test_variation_matches_oracle(orc, R, idx):
    g = _genome(idx)
    W = H = 32
    walkers = 4096
    r = R(W, H, 1)
    r.reset(g)
    r.set_trace(walkers, 1)
    r.iterate(g, 1234 + idx, walkers, 1, fuse=0)
    tg = r.read_trace()[:, 0]
    xf, ds, u, pal = blobs(g, W, H, iters=1, fuse=0, seed=1234 + idx, walkers=walkers)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, 1))
    to = to[:, 0]
    assert np.array_equal(tg[:, 2:4], to[:, 2:4]), "affine input"
    assert np.array_equal(tg[:, 10].view(np.uint32), to[:, 10].view(np.uint32)), "draw count"
    a, b = tg[:, 4:6].astype(np.float64), to[:, 4:6].astype(np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    same_inf = np.isinf(a) & (a == b)
    fin = np.isfinite(a) & np.isfinite(b)
    rel = np.where(fin, np.abs(a - b) / np.maximum(np.abs(b), 1.0), 0.0)
    ok = both_nan | same_inf | (fin & (rel <= 1e-5))
    frac = ok.all(axis=1).mean()
    assert frac >= 0.995, (idx, frac, float(np.nanmax(rel)))
    r.destroy()


def test_pre_blur_on_device(orc, R):
    g = genome([xform(1.0, (1, 0, 0, 0, 1, 0), vars=((97, 0.3), (0, 1.0)))], scale=20.0)
    walkers = 2048
    r = R(32, 32, 1)
    r.reset(g)
    r.set_trace(walkers, 3)
    r.iterate(g, 99, walkers, 3, fuse=0)
    tg = r.read_trace()
    xf, ds, u, pal = blobs(g, 32, 32, iters=3, fuse=0, seed=99, walkers=walkers)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, 3))
    assert np.array_equal(tg[..., 10].view(np.uint32), to[..., 10].view(np.uint32))
    assert np.allclose(tg[..., 4:6], to[..., 4:6], rtol=1e-5, atol=1e-6)
