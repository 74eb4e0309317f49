"""Oracle variation library vs the instrumented-flam3 f64 fixtures.

tests/golden/variations/*.json are the reference's fixture files (99
variations, 4,607 rows of flam3-C f64 outputs); tests/golden/variation_cases.json
holds the inputs the reference's test injects beside them (parameters, affine
b/c/e/f, RNG branch counts). The runners mirror src/variations.test.ts:
deterministic rows within abs 1e-6 (spiral 2e-6, power 3e-6) of flam3;
julia tries both branch bits; julian/juliascope/cpow/wedge_julia try each
floor(rand * |power|) branch; the RNG "smoke" variations must stay finite
under the reference's placeholder draws. The f64 instantiation is the pinned
one; the f32 instantiation (what the chaos oracle and the kernels run) must
agree with it to f32 accuracy on the same rows.
"""
import json
import math

import numpy as np
import pytest

from conftest import GOLDEN

CASES = json.loads((GOLDEN / "variation_cases.json").read_text())
TOL = {"spiral": 2e-6, "power": 3e-6}
PLACEHOLDER = [0.25, 0.5, 0.75, 0.5, 0.3]


def _close(a, b, abs_eps, rel=1e-12):
    if math.isnan(a) or math.isnan(b):
        return math.isnan(a) and math.isnan(b)
    if math.isinf(a) or math.isinf(b):
        return a == b
    return abs(a - b) <= max(abs_eps, rel * max(abs(a), abs(b), 1.0))


def _word(r01):
    return int(round(r01 * 268435455.0))


def _rows(name):
    return json.loads((GOLDEN / "variations" / f"{name}.json").read_text())["fixtures"]


def _branch_words(case, n):
    if case["runner"] == "runJulia":
        return [n]
    power = abs(case["params"][2] if case["index"] in (41,) else
                case["params"][2] if case["index"] == 96 else case["params"][0])
    return [_word((n + 0.5) / power)]


def _eval(orc, case, row, words, dtype=np.float64):
    params = list(case["params"])
    return orc.variation(case["index"], row["tx"], row["ty"], row["weight"], params,
                         case["affine"], words, dtype=dtype)


def _candidates(orc, case, row, dtype=np.float64):
    r = case["runner"]
    if r in ("runJulia", "runMultiBranchRng"):
        nb = 2 if r == "runJulia" else case["branches"]
        return [_eval(orc, case, row, _branch_words(case, n), dtype) for n in range(nb)]
    return [_eval(orc, case, row, [], dtype)]


NAMES = sorted(n for n, c in CASES.items() if c["runner"] != "runSmoke" and n != "pre_blur")
SMOKE = sorted(n for n, c in CASES.items() if c["runner"] == "runSmoke" and n != "pre_blur")


def test_fixture_set_is_complete():
    files = sorted(p.stem for p in (GOLDEN / "variations").glob("*.json"))
    assert files == sorted(CASES) and len(files) == 99
    assert sum(len(_rows(n)) for n in files) == 4607


@pytest.mark.parametrize("name", NAMES)
def test_f64_matches_flam3(orc, name):
    case = CASES[name]
    eps = TOL.get(name, 1e-6)
    bad = []
    for i, row in enumerate(_rows(name)):
        outs = _candidates(orc, case, row)
        if not any(_close(o[0], row["expected_x"], eps) and _close(o[1], row["expected_y"], eps)
                   for o in outs):
            bad.append((i, row["tx"], row["ty"], outs[0][:2], row["expected_x"], row["expected_y"]))
    assert not bad, f"{name}: {len(bad)} rows off, first {bad[:3]}"


@pytest.mark.parametrize("name", NAMES)
def test_f32_tracks_f64(orc, name):
    """The f32 restatement tracks the f64 one to f32 accuracy.

    Gap = |f32 - f64| / max(|f64|, 1). All rows must be below 5e-3 and at most
    2% above 1e-4: ill-conditioned rows (tan near a pole inside popcorn2)
    amplify f32 input rounding.
    """
    case = CASES[name]
    gaps = []
    for row in _rows(name):
        a = _candidates(orc, case, row, np.float64)
        b = _candidates(orc, case, row, np.float32)
        for (ax, ay, _), (bx, by, _) in zip(a, b):
            for u, v in ((ax, bx), (ay, by)):
                if math.isfinite(u) and abs(u) < 1e6:
                    gaps.append(abs(u - v) / max(abs(u), 1.0))
    g = np.array(gaps)
    assert g.max() < 5e-3 and np.mean(g > 1e-4) <= 0.02, f"{name}: worst gap {g.max():.2e}"


@pytest.mark.parametrize("name", SMOKE)
def test_rng_variations_finite_under_placeholder_draws(orc, name):
    case = CASES[name]
    words = [_word(r) for r in PLACEHOLDER[:case["rand_count"]]]
    for row in _rows(name)[:5]:
        x, y, used = _eval(orc, case, row, words)
        assert math.isfinite(x) and math.isfinite(y)
        assert used == case["rand_count"], (name, used)
