"""The generator's table of catalog entries that draw from the walker's RNG.

The table (pyr3_amd/csrc/p3_var_draws.h) lets the generated kernel skip,
during the fuse, a final xform that draws nothing. It is checked two ways:
against the device dispatch sources (an entry draws iff its case passes the
walker's Rng) and against the oracle's draw counters (entries the table
calls draw-free draw nothing at any of a spread of inputs; every other entry
draws at some input with its default or fixture parameters).
"""
import json
import re

import numpy as np
import pytest

from conftest import GOLDEN, ROOT
from pyr3_amd import catalog as cat

CSRC = ROOT / "pyr3_amd" / "csrc"


def _mask_table():
    text = (CSRC / "p3_var_draws.h").read_text()
    words = re.search(r"kVarDrawsMask\[11\] = \{([^}]*)\}", text).group(1)
    mask = [int(w.strip().rstrip("u"), 16) for w in words.split(",")]
    return {i: bool((mask[i >> 5] >> (i & 31)) & 1) for i in range(326) if i != 97}


def _cases(path, head):
    """catalog index -> statement text of its case in the named switch."""
    text = path.read_text()
    body = text[text.index("switch (idx)", text.index(head)):]
    body = body[:body.index("\n  }\n")]
    out, labels, buf = {}, [], []

    def flush():
        for lab in labels:
            out[lab] = "\n".join(buf)

    for line in body.split("\n")[1:]:
        if line.strip().startswith("default:"):
            break
        found = re.findall(r"case (\d+):", line)
        done = buf and "\n".join(buf).count("{") == "\n".join(buf).count("}")
        if found and (not buf or done):
            flush()
            labels, buf = [int(f) for f in found], []
            rest = re.sub(r"case \d+:", "", line).strip()
            if rest:
                buf.append(rest)
        else:
            buf.append(line.strip())
    flush()
    return out


def test_table_matches_dispatch_sources():
    cases = _cases(CSRC / "p3_variation_dispatch.h", "P3_INLINE float2 apply_variation(")
    cases.update(_cases(CSRC / "p3_variations_ext.h", "P3_INLINE float2 apply_variation_ext("))
    assert sorted(cases) == [i for i in range(326) if i != 97]
    derived = {i: "rng" in t for i, t in cases.items()}
    assert derived == _mask_table()
    assert sum(derived.values()) == 63


# entries that draw only away from their default parameters
DRAWING_PARAMS = {cat.VARIATION_INDEX["bcircle"]: [1.0, 0.5]}   # border != 0


def _params(idx, fixtures):
    if idx in DRAWING_PARAMS:
        return DRAWING_PARAMS[idx]
    case = fixtures.get(idx)
    if case:
        return case["params"]
    return cat.VARIATION_DEFAULTS.get(cat.VARIATION_NAMES[idx], [])


def test_table_matches_oracle_draw_counts(orc):
    fixtures = {c["index"]: c for c in
                json.loads((GOLDEN / "variation_cases.json").read_text()).values()}
    rng = np.random.default_rng(3)
    pts = [(float(x), float(y)) for x, y in rng.uniform(-2.0, 2.0, size=(24, 2))]
    pts += [(0.3, -0.2), (1.4, 0.7), (-0.9, 1.2), (0.01, 0.02)]
    table = _mask_table()
    wrong = []
    for idx, draws in table.items():
        p = _params(idx, fixtures)
        counts = [orc.variation(idx, x, y, 1.0, p)[2] for x, y in pts]
        if draws != any(c > 0 for c in counts):
            wrong.append((idx, cat.VARIATION_NAMES[idx], draws, max(counts)))
    assert not wrong, wrong
