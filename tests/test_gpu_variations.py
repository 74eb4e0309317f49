"""Every implemented catalog entry on the GPU vs the oracle's f32 restatement.

Entries are checked in groups of 16: one genome holds one xform per entry
(weight 1 each, the flam3 fixture parameters for V0..V98 and the reference's
defaults for V99..V325, affine scaled x2 so inputs span [-2, 2]^2), so a
group compiles one specialised kernel. 8,192 walkers run one turn each; the
trace gives, per walker, the picked xform, the affine input, the pre-post
output pv_pre and the draw counter. Per entry: the pick and the affine input
are bit-exact, the draw count is exact (same ISAAC stream, same
consumption), and pv_pre is within 1e-5 relative (scale max(|v|, 1)) on at
least 99.5% of its walkers. Where an entry misses that bar and draws no
random numbers, the share is taken over the walkers at which the entry is
well-conditioned in f32 (the oracle's f32 and f64 evaluations of that input
agree within 1e-6): an ulp of difference in a transcendental is amplified
near a pole or a Newton iteration's breakdown (mollweide, klein_j,
lichtenberg); epispiral, which also draws, is held to 98%. Entries whose value hangs on the last ulp of sin()
of a large argument (the sin-hash fract of circular, circular2, waves4) or on
a random integer times a rounded angle (hypertile2, phoenix_julia) are held
to their draw counts.
"""
import json

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import blobs, genome, xform
from pyr3_amd import catalog as cat
from pyr3_amd.variations_supported import SUPPORTED

pytestmark = pytest.mark.gpu

CASES = {c["index"]: c for c in json.loads((GOLDEN / "variation_cases.json").read_text()).values()}
INDICES = sorted(i for i in SUPPORTED if i != 97)
GROUP = 16
GROUPS = [INDICES[i:i + GROUP] for i in range(0, len(INDICES), GROUP)]
# value hangs on the last ulp of sin() of a large argument or on a random
# integer times a rounded angle: held to the draw count only
HASHY = {cat.VARIATION_INDEX[n] for n in ("circular", "circular2", "waves4", "hypertile2",
                                          "phoenix_julia")}
# randomised entry with a pole (1/cos(n theta)); its f32 conditioning cannot
# be checked per walker without the walker's draw: held to a 98% share
POLE = {cat.VARIATION_INDEX["epispiral"]}


def _rel_ok(a, b, tol):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    both_nan = np.isnan(a) & np.isnan(b)
    same_inf = np.isinf(a) & (a == b)
    fin = np.isfinite(a) & np.isfinite(b)
    rel = np.where(fin, np.abs(a - b) / np.maximum(np.abs(b), 1.0), 0.0)
    return (both_nan | same_inf | (fin & (rel <= tol))).all(axis=-1)


def _well_conditioned(orc, idx, params, pa):
    """Walkers whose input the oracle evaluates alike in f32 and f64."""
    out = []
    for x, y in pa:
        a = orc.variation(idx, float(x), float(y), 1.0, params, dtype=np.float64)
        b = orc.variation(idx, float(x), float(y), 1.0, params, dtype=np.float32)
        out.append(bool(_rel_ok([b[:2]], [a[:2]], 1e-6)[0]))
    return np.array(out)


def _params(idx):
    case = CASES.get(idx)
    return case["params"] if case else cat.VARIATION_DEFAULTS.get(cat.VARIATION_NAMES[idx], [])


def _xform(idx):
    case = CASES.get(idx)
    params = _params(idx)
    aff = (2.0, 0.1, 0.05, -0.1, 2.0, -0.05)
    if case and case["runner"] == "runWithAffine" and idx in (16, 18, 30, 31):
        a, b, c, d, e, f = case["affine"]
        aff = (2.0, b, c, -0.1, 2.0, f)
    return xform(1.0, aff, vars=((idx, 1.0),), params=[list(params)])


@pytest.fixture(scope="module")
def R():
    from pyr3_amd.renderer import Renderer
    return Renderer


@pytest.mark.parametrize("group", range(len(GROUPS)))
def test_variation_group_matches_oracle(orc, R, group):
    idx = GROUPS[group]
    g = genome([_xform(i) for i in idx], scale=20.0)
    W = H = 32
    walkers = 8192
    seed = 1234 + group
    r = R(W, H, 1)
    r.reset(g)
    r.set_trace(walkers, 1)
    r.iterate(g, seed, walkers, 1, fuse=0)
    tg = r.read_trace()[:, 0]
    xf, ds, u, pal = blobs(g, W, H, iters=1, fuse=0, seed=seed, walkers=walkers)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, 1))
    to = to[:, 0]
    r.destroy()
    assert np.array_equal(tg[:, 1], to[:, 1]), "pick"
    assert np.array_equal(tg[:, 2:4], to[:, 2:4]), "affine input"
    bad = []
    for j, v in enumerate(idx):
        m = to[:, 1] == j
        if m.sum() < 100:
            bad.append((v, "too few walkers", int(m.sum())))
            continue
        if not np.array_equal(tg[m, 10].view(np.uint32), to[m, 10].view(np.uint32)):
            bad.append((v, "draw count"))
            continue
        if v in HASHY:
            continue
        ok = _rel_ok(tg[m, 4:6], to[m, 4:6], 1e-5)
        if ok.mean() >= (0.98 if v in POLE else 0.995):
            continue
        if all(orc.variation(v, x, y, 1.0, _params(v))[2] == 0    # draws nothing
               for x, y in ((0.3, -0.2), (1.4, 0.7), (-0.9, 1.2))):
            cond = _well_conditioned(orc, v, _params(v), to[m, 2:4])
            if cond.mean() >= 0.5 and ok[cond].mean() >= 0.995:
                continue
        bad.append((v, cat.VARIATION_NAMES[v], float(ok.mean())))
    assert not bad, bad


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


def test_direct_colour_histogram_bit_exact(orc, R):
    """Direct-colour entries whose colour is exact f32 arithmetic (dc_linear,
    dc_gridout, dc_perlin) on linear maps: the splat colours, hence the whole
    histogram, are bit-exact. Covers the record's dc flag, the slot-weight
    gate, last-one-wins and the final xform's override."""
    V = cat.VARIATION_INDEX
    xs = [xform(1.0, (0.5, 0, 0, 0, 0.5, 0), vars=((0, 1.0), (V["dc_linear"], 0.3)), color=0.1),
          xform(1.0, (0.5, 0, 0.5, 0, 0.5, 0), vars=((0, 1.0), (V["dc_gridout"], 0.2)),
                params=[[], [6.0]], color=0.5),
          xform(1.0, (0.5, 0, 0.25, 0, 0.5, 0.5),
                vars=((V["dc_perlin"], 0.4), (0, 1.0), (V["dc_gridout"], 0.0)),
                params=[[3.0, 2.0, 0.25]], color=0.9)]
    fin = xform(1.0, (1, 0, 0, 0, 1, 0), vars=((0, 1.0), (V["dc_linear"], 0.1)))
    for final in (None, fin):
        g = genome(xs, scale=20.0, cx=0.5, cy=0.5, final=final)
        W = H = 40
        walkers, iters, fuse, seed = 1024, 60, 5, 77
        r = R(W, H, 1)
        r.reset(g)
        r.iterate(g, seed, walkers, iters, fuse=fuse)
        hg = r.read_histogram()
        r.destroy()
        xf, ds, u, pal = blobs(g, W, H, iters=iters, fuse=fuse, seed=seed, walkers=walkers)
        assert xf[:, 11].max() == 1.0  # dc flag packed
        ho, _ = orc.chaos(xf, ds, u, pal, trace_shape=(1, 1))
        assert np.array_equal(hg, ho), final is not None


def test_hash_entries_on_well_conditioned_inputs(orc, R):
    """Values of the five entries held to draw counts above, where their
    arithmetic is well conditioned: circular, circular2 and waves4 with
    inputs scaled to ~1e-6, where the sine inside their hash is its
    argument in f32 on both sides (circular2's seeds make its hash argument
    tiny too); hypertile2 and phoenix_julia at ordinary inputs, where only
    a 1-ulp difference of the large rotation angle's sum separates the two."""
    names = ("circular", "circular2", "waves4", "hypertile2", "phoenix_julia")
    tiny = {"circular", "circular2", "waves4"}
    xs = []
    for n in names:
        i = cat.VARIATION_INDEX[n]
        s = 1e-6 if n in tiny else 2.0
        xs.append(xform(1.0, (s, 0.0, 0.0, 0.0, s, 0.0), vars=((i, 1.0),), params=[list(_params(i))]))
    g = genome(xs, scale=20.0)
    W = H = 32
    walkers, seed = 8192, 4321
    r = R(W, H, 1)
    r.reset(g)
    r.set_trace(walkers, 1)
    r.iterate(g, seed, walkers, 1, fuse=0)
    tg = r.read_trace()[:, 0]
    r.destroy()
    xf, ds, u, pal = blobs(g, W, H, iters=1, fuse=0, seed=seed, walkers=walkers)
    _, to = orc.chaos(xf, ds, u, pal, trace_shape=(walkers, 1))
    to = to[:, 0]
    assert np.array_equal(tg[:, 1], to[:, 1])
    rates = {}
    for j, n in enumerate(names):
        m = to[:, 1] == j
        assert m.sum() >= 500
        assert np.array_equal(tg[m, 10].view(np.uint32), to[m, 10].view(np.uint32)), n
        rates[n] = float(_rel_ok(tg[m, 4:6], to[m, 4:6], 1e-4).mean())
    print("share within 1e-4:", rates)
    assert all(v >= 0.99 for v in rates.values()), rates
