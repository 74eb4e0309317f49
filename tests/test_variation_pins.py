"""Value pins of catalog entries V99..V325 from the reference's own tests (CPU).

The reference checks these entries in its per-variation GPU tests
(src/issue*.gpu.test.ts, src/post-flam3-batch*.gpu.test.ts, ...): the WGSL
function runs on a handful of points with explicit parameters and its f32
output is compared with an f64 oracle written in the test file, at the
test's toBeCloseTo precision (|a - b| < 10^-d / 2). Each pin below restates
one such test oracle in Python (reference file:line cited), evaluates it at
the test's points with the test's weight and parameters, and holds

  * the oracle's f64 instantiation to the restated test oracle within
    1e-9 relative (two transcriptions of one formula, from different
    sources: a slip in either shows), and
  * the oracle's f32 instantiation, the arithmetic the kernels run, to the
    test's own precision (the reference's GPU bar).

Parameters go in the WGSL function's argument order after the weight, the
order the oracle's apply_variation passes a record's parameter words in.
Where the reference test of an entry is a finiteness smoke test only (most
randomised entries), the pin restates that test (finite output at its points
and parameters for any RNG words) and adds the special cases the mathematics
fixes.
"""
import math

import numpy as np
import pytest

from pyr3_amd import catalog as cat

V = cat.VARIATION_INDEX
PI = math.pi
# TEST_POINTS shared by most per-variation files (e.g.
# src/post-flam3-batch1.gpu.test.ts:52-61)
P8 = [(0.5, 0.3), (-0.5, 0.7), (1.2, -0.4), (-1.0, -1.0), (0.1, 0.05), (2.0, 0.0),
      (0.0, 2.0), (-2.5, 1.7)]
# src/post-flam3-batch2b-b.gpu.test.ts:44-53 (off-axis variant)
P8B = [(0.5, 0.3), (-0.5, 0.7), (1.2, -0.4), (-1.0, -1.0), (0.1, 0.05), (2.0, 0.3),
       (0.3, 2.0), (-2.5, 1.7)]
WORDS = [int(v) for v in np.random.default_rng(77).integers(0, 2 ** 32, 256, dtype=np.uint64)]


def f32(v):
    return float(np.float32(v))


def run(orc, name, x, y, w, params, dtype=np.float64, words=()):
    return orc.variation(V[name], x, y, w, list(params), rng_words=words, dtype=dtype)


def check_pin(orc, name, oracle, w, params, points, digits, skip=None, f64_rel=1e-9, f32=True):
    """The test's bar on the f32 oracle and a tight f64 agreement (f64 only
    with f32=False: points where the f32 result is ill-conditioned)."""
    tol = 0.5 * 10.0 ** -digits
    n = 0
    for x, y in points:
        if skip is not None and skip(x, y):
            continue
        want = oracle(x, y, w, *params)
        g64 = run(orc, name, x, y, w, params)
        for k in range(2):
            assert abs(g64[k] - want[k]) <= f64_rel * max(1.0, abs(want[k])), \
                (name, x, y, k, g64[k], want[k])
        if f32:
            g32 = run(orc, name, np.float32(x).item(), np.float32(y).item(), w, params,
                      dtype=np.float32)
            for k in range(2):
                assert abs(g32[k] - want[k]) < tol, (name, "f32", x, y, k, g32[k], want[k])
        n += 1
    assert n > 0
    return n


def check_finite(orc, name, w, params, points, min_finite=None):
    """The reference's smoke bar: finite output at its points for its
    parameters, here for several RNG streams."""
    for s in range(4):
        words = WORDS[s * 32:(s + 1) * 32]
        fin = 0
        for x, y in points:
            o = run(orc, name, f32(x), f32(y), w, params, dtype=np.float32, words=words)
            fin += int(np.isfinite(o[0])) + int(np.isfinite(o[1]))
        need = 2 * len(points) if min_finite is None else min_finite
        assert fin >= need, (name, s, fin)


# ---------------------------------------------------------------- bubble2 --
def bubble2_ref(x, y, w, xs, ys):   # src/issue120-b2-bubble2.gpu.test.ts:32-35
    r = w / (0.25 * (x * x + y * y) + 1.0)
    return x * r * xs, y * r * ys


def test_bubble2(orc):
    # :87-101 default (1, 1), :103-117 anisotropic (2.0, 0.5), toBeCloseTo 5
    check_pin(orc, "bubble2", bubble2_ref, 1.0, (1.0, 1.0), P8, 5)
    check_pin(orc, "bubble2", bubble2_ref, 1.0, (2.0, 0.5), P8, 5)
    # :119-136 finite at (-1.5, 1.7); (1, 1) equals bubble (V20)
    check_finite(orc, "bubble2", 1.0, (-1.5, 1.7), P8)
    for x, y in P8:
        assert run(orc, "bubble2", x, y, 1.0, (1.0, 1.0))[:2] == \
            pytest.approx(orc.variation(V["bubble"], x, y, 1.0)[:2], rel=1e-12)


# -------------------------------------------------------- post-flam3 batch 1
def loonie2_ref(x, y, w, sides_f, star, circle):   # src/post-flam3-batch1.gpu.test.ts:132-163
    sides = max(1, min(16, int(math.trunc(sides_f))))
    a = 2 * PI / sides
    sa, ca = math.sin(a), math.cos(a)
    ss, cs = math.sin(-star * PI * 0.5), math.cos(star * PI * 0.5)
    sc, cc = math.sin(circle * PI * 0.5), math.cos(circle * PI * 0.5)
    vv = w * w
    xr, yr = x, y
    r2 = xr * cs + abs(yr) * ss
    cr = math.sqrt(x * x + y * y)
    for _ in range(min(sides - 1, 16)):
        xr, yr = xr * ca - yr * sa, xr * sa + yr * ca
        r2 = max(r2, xr * cs + abs(yr) * ss)
    r2 = r2 * cc + cr * sc
    r2 = r2 * r2 if sides > 2 else abs(r2) * r2
    if 0 < r2 < vv:
        r = w * math.sqrt(abs(vv / r2 - 1))
        return r * x, r * y
    if r2 < 0:
        r = w / math.sqrt(abs(vv / r2) - 1)
        return r * x, r * y
    return w * x, w * y


def test_loonie2(orc):
    check_pin(orc, "loonie2", loonie2_ref, 1.0, (4.0, 0.15, 0.25), P8, 3)   # :166-182
    check_finite(orc, "loonie2", 0.8, (2.0, 0.2, 0.3), P8)                  # :184-196


def epispiral_ref(x, y, w, n, holes):   # src/post-flam3-batch1.gpu.test.ts:203-209
    th = math.atan2(y, x)
    d = math.cos(n * th)
    t = -holes + 1 / d
    return w * t * math.cos(th), w * t * math.sin(th)


def test_epispiral(orc):
    # thickness 0 draws nothing (:212-232, toBeCloseTo 2, near-pole points skipped)
    def pole(x, y):
        return abs(math.cos(6.0 * math.atan2(y, x))) < 1e-3

    params = (6.0, 0.0, 1.0)
    for x, y in P8:
        if not pole(x, y):
            want = epispiral_ref(x, y, 1.0, 6.0, 1.0)
            g = run(orc, "epispiral", x, y, 1.0, params, words=WORDS)
            assert g[:2] == pytest.approx(want, rel=1e-9, abs=1e-9)
            g32 = run(orc, "epispiral", f32(x), f32(y), 1.0, params, np.float32, WORDS)
            assert abs(g32[0] - want[0]) < 5e-3 and abs(g32[1] - want[1]) < 5e-3
    check_finite(orc, "epispiral", 1.0, (4.0, 0.5, 1.0), P8, min_finite=len(P8) + 1)   # :234-250


def test_cpow2_cpow3_smoke_and_identity(orc):
    # src/post-flam3-batch1.gpu.test.ts:257-292: finite at the defaults
    check_finite(orc, "cpow2", 1.0, (1.0, 0.0, 1.0, 1.0), P8, min_finite=2 * len(P8) - 4)
    check_finite(orc, "cpow3", 1.0, (1.0, 1.0, 1.0, 1.0), P8, min_finite=2 * len(P8) - 4)


# ------------------------------------------------------- post-flam3 batch 2b-a
def loonie3_ref(x, y, w):   # src/post-flam3-batch2b-a.gpu.test.ts:132-146
    vv = w * w
    r2 = 2 * vv
    if x > 1e-30:
        q = (x * x + y * y) / x
        r2 = q * q
    if r2 < vv:
        r = w * math.sqrt(vv / r2 - 1)
        return r * x, r * y
    return w * x, w * y


def test_loonie3(orc):
    check_pin(orc, "loonie3", loonie3_ref, 1.0, (), P8, 3)   # :148-165
    check_finite(orc, "loonie3", 0.7, (), P8)                # :167-179


def test_batch2b_a_rng_smoke(orc):
    # src/post-flam3-batch2b-a.gpu.test.ts:186-317 (finite-output smoke)
    check_finite(orc, "juliaq", 1.0, (3.0, 2.0), P8, min_finite=2 * len(P8) - 4)
    check_finite(orc, "glynnia", 1.0, (), P8)
    check_finite(orc, "falloff", 1.0, (1.0, 0.5, 1.0, 1.0, 0.0, 0.0), P8)
    for typ in (0.0, 1.0, 2.0):
        check_finite(orc, "falloff2", 1.0, (1.0, typ, 1.0, 1.0, 0.0, 0.0, 0.5), P8)
    for inv in (0.0, 1.0):
        check_finite(orc, "falloff3", 1.0, (1.0, 1.0, 1.0, 0.0, 0.0, 0.5, inv), P8)


# ------------------------------------------------------- post-flam3 batch 2b-b
def collideoscope_ref(x, y, w, a_param, num_raw):   # src/post-flam3-batch2b-b.gpu.test.ts:111-135
    num = max(1, int(math.trunc(num_raw)))
    kn_pi, pi_kn = num / PI, PI / num
    ka_kn = PI * a_param / num
    a = math.atan2(y, x)
    r = w * math.sqrt(x * x + y * y)
    fm = math.fmod   # JS % keeps the dividend's sign, as fmod
    if a >= 0.0:
        alt = int(math.trunc(a * kn_pi))
        a = alt * pi_kn + (fm(ka_kn + a, pi_kn) if alt % 2 == 0 else fm(-ka_kn + a, pi_kn))
    else:
        alt = int(math.trunc(-a * kn_pi))
        a = -(alt * pi_kn + (fm(-ka_kn - a, pi_kn) if alt % 2 != 0 else fm(ka_kn - a, pi_kn)))
    return r * math.cos(a), r * math.sin(a)


def _circ_perimeter(x, y):
    ax, ay = abs(x), abs(y)
    if ax >= ay:
        return (ax + y if x >= ay else 5.0 * ax - y), ax
    return (3.0 * ay - x if y >= ax else 7.0 * ay + x), ay


def circlize_ref(x, y, w, hole):   # :175-200
    per, side = _circ_perimeter(x, y)
    if side == 0:
        return 0.0, 0.0
    r = w / (PI / 4) * side + hole
    a = (PI / 4) * per / side - PI / 4
    return r * math.cos(a), r * math.sin(a)


def circlize2_ref(x, y, w, hole):   # :244-268
    per, side = _circ_perimeter(x, y)
    if side == 0:
        return 0.0, 0.0
    r = w * (side + hole)
    a = (PI / 4) * per / side - PI / 4
    return r * math.cos(a), r * math.sin(a)


def eswirl_ref(x, y, w, in_p, out_p):   # :314-332
    tmp, tmp2 = y * y + x * x + 1.0, 2.0 * x
    r1 = math.sqrt(tmp + tmp2) if tmp + tmp2 > 0 else 0.0
    r2 = math.sqrt(tmp - tmp2) if tmp - tmp2 > 0 else 0.0
    xmax = max((r1 + r2) * 0.5, 1.0)
    mu = math.acosh(xmax)
    t = min(max(x / xmax, -1.0), 1.0)
    nu = math.acos(t)
    if y < 0:
        nu = -nu
    nw = nu + mu * out_p + in_p / (mu if mu != 0 else 1e-30)
    return w * math.cosh(mu) * math.cos(nw), w * math.sinh(mu) * math.sin(nw)


def petal_ref(x, y, w):   # :382-389
    a = math.cos(x)
    bx = (math.cos(x) * math.cos(y)) ** 3
    by = (math.sin(x) * math.cos(y)) ** 3
    return w * a * bx, w * a * by


def test_batch2b_b(orc):
    src = "post-flam3-batch2b-b"
    check_pin(orc, "collideoscope", collideoscope_ref, 1.0, (0.20, 1.0), P8B, 2)   # :138-154
    check_finite(orc, "collideoscope", 1.0, (0.5, 5.0), P8B)                        # :156-168
    check_pin(orc, "circlize", circlize_ref, 1.0, (0.40,), P8B, 3)                 # :203-219
    check_pin(orc, "circlize", circlize_ref, 0.8, (0.0,), P8B, 3)                  # :221-237
    check_pin(orc, "circlize2", circlize2_ref, 1.0, (0.0,), P8B, 3)                # :271-287
    check_pin(orc, "circlize2", circlize2_ref, 1.0, (0.5,), P8B, 3)                # :289-305
    # :335-361: parity where x^2 + y^2 > 0.5 (clear of the xmax = 1 cusp), 1e-2
    check_pin(orc, "eswirl", eswirl_ref, 1.0, (1.2, 0.2), P8B, 2,
              skip=lambda x, y: x * x + y * y <= 0.5)
    check_finite(orc, "eswirl", 1.0, (1.2, 0.0), P8B)                              # :363-375
    check_pin(orc, "petal", petal_ref, 1.0, (), P8B, 3)                            # :392-408
    check_finite(orc, "petal", 0.5, (), P8B)                                       # :410-422
    assert src


# ------------------------------------------------------- post-flam3 batch 2b-c
HALF_WORD = 134217727   # rand01 of this draw is 0.5 to 2e-9 (the test stubs rand01 = 0.5)


def bcircle_ref(x, y, w, scale):   # src/post-flam3-batch2b-c.gpu.test.ts:125-132
    if x == 0 and y == 0:
        return 0.0, 0.0
    xs, ys = x * scale, y * scale
    return (w * xs, w * ys) if math.sqrt(xs * xs + ys * ys) <= 1.0 else (0.0, 0.0)


def curl2_ref(x, y, w, c1, c2, c3):   # :196-208
    re = c3 * x ** 3 - 3 * c3 * x * y * y + c2 * x * x - c2 * y * y + c1 * x + 1.0
    im = 3 * c3 * x * x * y - c3 * y ** 3 + 2 * c2 * x * y + c1 * y
    r = w / (re * re + im * im)
    return (x * re + y * im) * r, (y * re - x * im) * r


def murl_ref(x, y, w, c_in, power_in):   # :258-271
    power = math.trunc(power_in)
    c = c_in / (power - 1) if power != 1 else c_in
    vp = w * (c + 1)
    a = math.atan2(y, x) * power
    r = c * (x * x + y * y) ** (power / 2.0)
    re, im = r * math.cos(a) + 1, r * math.sin(a)
    r1 = vp / (re * re + im * im + 1e-29)
    return r1 * (x * re + y * im), r1 * (y * re - x * im)


def stwins_ref(x, y, w, distort):   # :318-331
    sx, sy = x * w * 0.05, y * w * 0.05
    num = (sx * sx - sy * sy) * math.sin(2 * PI * distort * (sx + sy))
    den = sx * sx + sy * sy
    res = num / (den if den != 0 else 1.0)
    return w * x + res, w * y + res


def hexes_ref(x, y, w, cell, power, rotate, scale):   # :378-452
    if cell == 0:
        return 0.0, 0.0
    s3 = math.sqrt(3.0)
    ah, bh, ch, dh = 1 / 3, s3 / 3, -1 / 3, s3 / 3
    ac, bc, cc_, dc = 1.5, -1.5, s3 / 2, s3 / 2
    rs, rc = math.sin(rotate * 2 * PI), math.cos(rotate * 2 * PI)
    hx0 = math.floor((ah * x + bh * y) / cell)
    hy0 = math.floor((ch * x + dh * y) / cell)
    best, q = math.inf, 0
    for di in (-1, 0, 1):
        for dj in (-1, 0, 1):
            cx = (ac * (hx0 + di) + bc * (hy0 + dj)) * cell
            cy = (cc_ * (hx0 + di) + dc * (hy0 + dj)) * cell
            d2 = (cx - x) ** 2 + (cy - y) ** 2
            if d2 < best:
                best, q = d2, (di + 1) * 3 + (dj + 1)
    hx, hy = hx0 + q // 3 - 1, hy0 + q % 3 - 1

    def ctr(i, j):
        return (ac * i + bc * j) * cell, (cc_ * i + dc * j) * cell
    p0 = ctr(hx, hy)
    nbrs = [ctr(hx, hy + 1), ctr(hx + 1, hy + 1), ctr(hx + 1, hy), ctr(hx, hy - 1),
            ctr(hx - 1, hy - 1), ctr(hx - 1, hy)]

    def vor(ux, uy):
        m = -1.0e20
        for px, py in nbrs:
            dx, dy = px - p0[0], py - p0[1]
            if dx == 0 and dy == 0:
                m = max(m, 1.0)
                continue
            m = max(m, 2.0 * ((ux - p0[0]) * dx + (uy - p0[1]) * dy) / (dx * dx + dy * dy))
        return m
    l1 = vor(x, y)
    dxo, dyo = x - p0[0], y - p0[1]
    trg = (l1 + 1e-30) ** power * scale
    vx0, vy0 = dxo * rc + dyo * rs, -dxo * rs + dyo * rc
    l2 = vor(vx0 + p0[0], vy0 + p0[1])
    L = max(l1, l2)
    if L < 0.5:
        R = trg / l1
    elif L > 0.8:
        R = trg / l2
    else:
        R = ((trg / l1) * (0.8 - L) + (trg / l2) * (L - 0.5)) / 0.3
    return w * (vx0 * R + p0[0]), w * (vy0 * R + p0[1])


def test_batch2b_c(orc):
    half = [HALF_WORD] * 8

    def bc(x, y, w, scale, border):
        return bcircle_ref(x, y, w, scale)
    # :135-172 borderwidth 0 (disk path), scale 1 and 0.5, toBeCloseTo 3
    for scale in (1.0, 0.5):
        for x, y in P8B:
            want = bcircle_ref(x, y, 1.0, scale)
            got = run(orc, "bcircle", f32(x), f32(y), 1.0, (scale, 0.0), np.float32, half)
            assert abs(got[0] - want[0]) < 5e-4 and abs(got[1] - want[1]) < 5e-4, (x, y, scale)
    check_finite(orc, "bcircle", 1.0, (1.0, 0.5), P8B)                            # :174-189
    check_pin(orc, "curl2", curl2_ref, 1.0, (1.0, 0.0, 0.0), P8B, 3)               # :211-229
    check_pin(orc, "curl2", curl2_ref, 1.0, (0.5, 0.2, 0.1), P8B, 2)               # :231-251
    check_pin(orc, "murl", murl_ref, 1.0, (0.1, 1.0), P8B, 3)                      # :274-290
    check_pin(orc, "murl", murl_ref, 1.0, (0.5, 3.0), P8B, 1)                      # :292-311
    check_pin(orc, "stwins", stwins_ref, 1.0, (1.0,), P8B, 3)                      # :334-350
    check_pin(orc, "stwins", stwins_ref, 1.0, (0.5,), P8B, 3)                      # :352-368
    check_pin(orc, "hexes", hexes_ref, 1.0, (1.0, 1.0, 0.166, 1.0), P8B, 2)        # :455-480
    check_finite(orc, "hexes", 1.0, (0.5, 2.0, 0.166, 1.0), P8B)                   # :482-498


# ------------------------------------------------------- post-flam3 batch 2b-d
def xheart_ref(x, y, w, angle, ratio):   # src/post-flam3-batch2b-d.gpu.test.ts:121-135
    ang = PI / 4 + 0.5 * (PI / 4) * angle
    ca, sa = math.cos(ang), math.sin(ang)
    rat = 6 + 2 * ratio
    r2 = x * x + y * y + 4
    if r2 == 0:
        r2 = 1
    bx, by = 4 / r2, rat / r2
    xr = ca * (bx * x) - sa * (by * y)
    yr = sa * (bx * x) + ca * (by * y)
    return (w * xr, w * yr) if xr > 0 else (w * xr, -w * yr)


def xhyperbol_ref(x, y, w, m00, m01, m10, m11, m20, m21):   # :180-195
    r = 1 / (x * x + y * y + 1e-10)
    xi, yi = x * r, y * r
    re, im = m00 * xi + m01 * yi + m20, m10 * xi + m11 * yi + m21
    al = math.atan2(im, re) + 2 * PI
    rsq = re * re + im * im
    xo, yo = rsq * math.cos(al), rsq * math.sin(al)
    ri = w / (xo * xo + yo * yo + 1e-10)
    return xo * ri, yo * ri


def xcurl2_ref(x, y, w, c1, c2, c3):   # :245-254
    re = 1 + c1 * x + c2 * (x * x - y * y) + c3 * (x ** 3 - 3 * x)
    im = c1 * y + c2 * (2 * x * y) + c3 * (3 * x * y - 1)
    r = w / (re * re + im * im)
    return (x * re + y * im) * r, (y * re + x * im) * r


def xyrus_gridout_ref(x, y, w):   # :342-361
    rx = math.floor(x + 0.5) if x >= 0 else math.ceil(x - 0.5)
    ry = math.floor(y + 0.5) if y >= 0 else math.ceil(y - 0.5)
    dx = dy = 0
    if ry <= 0:
        if rx > 0:
            if -ry >= rx:
                dx = 1
            else:
                dy = 1
        elif ry <= rx:
            dx = 1
        else:
            dy = -1
    elif rx > 0:
        if ry >= rx:
            dx = -1
        else:
            dy = 1
    elif ry > -rx:
        dx = -1
    else:
        dy = -1
    return w * (x + dx), w * (y + dy)


def test_batch2b_d(orc):
    check_pin(orc, "xheart", xheart_ref, 1.0, (0.0, 0.0), P8B, 3)                  # :138-154
    check_pin(orc, "xheart", xheart_ref, 1.0, (1.0, 0.5), P8B, 3)                  # :156-172
    check_pin(orc, "xhyperbol", xhyperbol_ref, 1.0, (1, 0, 0, 1, 0, 0), P8B, 2)    # :198-216
    check_pin(orc, "xhyperbol", xhyperbol_ref, 1.0, (0.6, -0.8, 0.8, 0.6, 0, 0), P8B, 2)   # :218-236
    check_pin(orc, "xcurl2", xcurl2_ref, 1.0, (1.0, 0.0, 0.0), P8B, 3)             # :257-273
    check_pin(orc, "xcurl2", xcurl2_ref, 1.0, (0.3, 0.2, 0.1), P8B, 2)             # :275-294
    for prm in ((2.0, 1.0, 1.0, 0.5, 1.0, 1.0), (3.0, 1.0, 1.0, 0.3, 1.0, 1.0)):   # :304-335
        for s in range(4):
            for x, y in P8B:
                o = run(orc, "xtrb", f32(x), f32(y), 1.0, prm, np.float32, WORDS[32 * s:])
                assert not (math.isnan(o[0]) or math.isnan(o[1])), (prm, x, y)
    # :364-390: exact up to a whole-cell offset at the rounding boundary
    for x, y in P8B:
        want = xyrus_gridout_ref(x, y, 1.0)
        assert run(orc, "gridout", x, y, 1.0, ())[:2] == pytest.approx(want, abs=1e-12)
        g = run(orc, "gridout", f32(x), f32(y), 1.0, (), np.float32)
        for k in range(2):
            d = abs(g[k] - want[k])
            assert d < 1e-3 or abs(d - 1.0) < 1e-3
    check_finite(orc, "blur_circle", 1.0, (0.0,), P8B)                             # :397-418
    check_finite(orc, "blur_circle", 1.0, (0.3,), P8B)                             # :420-433


# ------------------------------------------------------------ issue 120 ----
def bipolar2_ref(x, y, w, shift, a, b, c, d, e, f1, g1, h):   # src/issue120-b1-bipolar2.gpu.test.ts:35-56
    hp = PI * 0.5
    x2y2 = (x * x + y * y) * g1
    t = x2y2 + a
    x2 = b * x
    yv = c * math.atan2(e * y, x2y2 - d) - hp * shift
    if yv > hp:
        yv = -hp + math.fmod(yv + hp, PI)
    elif yv < -hp:
        yv = hp - math.fmod(hp - yv, PI)
    fn, gn = t + x2, t - x2
    if gn == 0 or fn / gn <= 0:
        return 0.0, 0.0
    return w * f1 * (2 / PI) * math.log(fn / gn), w * (2 / PI) * yv * h


def test_bipolar2(orc):
    prm = (0.0, 1.0, 2.0, 0.5, 1.0, 2.0, 0.25, 1.0, 1.0)
    check_pin(orc, "bipolar2", bipolar2_ref, 1.0, prm, P8, 3)                       # :106-134
    assert run(orc, "bipolar2", 1.0, 0.1, 1.0, (0.0, -3.0) + prm[2:])[:2] == (0.0, 0.0)   # :136-157
    check_finite(orc, "bipolar2", 1.0, (0.5, 0.7, 1.5, 0.3, 0.8, 1.2, 0.4, 1.1, 0.9), P8)   # :159-173


def test_inverse_hyperbolics(orc):
    # src/issue120-b3-inverse-hyperbolic.gpu.test.ts:141-161: arcsinh at (0.5, 0.3)
    x, y = 0.5, 0.3
    zr, zi = x * x - y * y + 1.0, 2 * x * y
    rad = math.hypot(zr, zi)
    sr = math.sqrt(max(0.5 * (rad + zr), 0.0))
    si = (-1 if zi < 0 else 1) * math.sqrt(max(0.5 * (rad - zr), 0.0))
    ur, ui = x + sr, y + si
    want = (0.5 * math.log(ur * ur + ui * ui + 1e-20) * 2 / PI, math.atan2(ui, ur) * 2 / PI)
    assert run(orc, "arcsinh", x, y, 1.0, ())[:2] == pytest.approx(want, rel=1e-9)
    g = run(orc, "arcsinh", x, y, 1.0, (), np.float32)
    assert abs(g[0] - want[0]) < 5e-4 and abs(g[1] - want[1]) < 5e-4
    # :135-189 finite output of all six on the points (acosh / acosech draw)
    for name in ("acosh", "arcsinh", "arctanh", "acoth", "acosech", "arcsech2"):
        check_finite(orc, name, 1.0, (), P8)


def cell2_ref(x, y, w, size, a, snx, sny, ssx, ssy):   # src/issue120-b35-cell2.gpu.test.ts:32-52
    s = 1e-30 if abs(size) < 1e-30 else size
    inv = a / s
    cx, cy = math.floor(x * inv), math.floor(y * inv)
    dx, dy = x - cx * s, y - cy * s
    if cy >= 0:
        sx, sy = cx * snx, cy * sny
    else:
        sx, sy = cx * ssx, -ssy * cy
    return w * (dx + sx * s), -w * (dy + sy * s)


def test_cell2(orc):
    check_pin(orc, "cell2", cell2_ref, 1.0, (0.6, 1.0, 2.0, 2.0, 2.0, 2.0), P8, 4)   # :102-123
    check_pin(orc, "cell2", cell2_ref, 1.0, (0.5, 1.2, 1.7, 3.0, 0.8, 1.2), P8, 4)   # :125-146
    check_finite(orc, "cell2", 1.0, (0.0, 1.0, 2.0, 2.0, 2.0, 2.0), P8)              # :148-164


def curl_sp_ref(x, y, w, pw, c1, c2, sx, sy):   # src/issue120-b4-xyrus-lukout.gpu.test.ts:129-148 (pow 1)
    xp = abs(x) ** pw * math.copysign(1.0, x) if x != 0 else 0.0
    yp = abs(y) ** pw * math.copysign(1.0, y) if y != 0 else 0.0
    d = xp * xp - yp * yp
    s1 = c1 * xp + c2 * d
    re = abs(s1) * (1 if s1 > 0 else -1) + 1.0
    s2 = c1 * yp + 2.0 * c2 * xp * yp
    im = abs(s2) * (1 if s2 > 0 else -1)
    r = 1.0 / max(abs(re * re + im * im), 1e-30)
    return (xp * re + yp * im) * r * w, (yp * re - xp * im) * r * w


def murl2_ref(x, y, w, c, pw):   # :155-175
    p2, invp = pw * 0.5, 1.0 / pw
    cp1 = c + 1.0
    vp = abs(cp1) ** (2.0 * invp) * (1 if cp1 >= 0 else -1)
    a1 = math.atan2(y, x) * pw
    r0 = c * abs(x * x + y * y) ** p2
    re0, im0 = r0 * math.cos(a1) + 1.0, r0 * math.sin(a1)
    r1 = abs(re0 * re0 + im0 * im0) ** invp
    a2 = math.atan2(im0, re0) * 2.0 * invp
    re1, im1 = r1 * math.cos(a2), r1 * math.sin(a2)
    rl = vp / max(r1 * r1, 1e-30)
    return w * rl * (x * re1 + y * im1), w * rl * (y * re1 - x * im1)


def test_xyrus_lukout(orc):
    check_pin(orc, "curl_sp", curl_sp_ref, 1.0, (1.0, -0.01, 0.03, 0.0, 0.0), P8, 3)   # :125-149
    check_pin(orc, "murl2", murl2_ref, 1.0, (0.1, 3.0), P8, 3)                           # :151-176
    check_finite(orc, "lissajous", 1.0, (-3.14159, 3.14159, 3.0, 2.0, 0.0, 0.0, 0.0), P8)   # :178-183
    check_finite(orc, "spirograph", 1.0, (3.0, 2.0, 0.0, -1.0, 1.0, -1.0, 1.0, 0.0, 0.0), P8)   # :185-190
    check_finite(orc, "waffle", 1.0, (6.0, 0.5, 0.5, 0.0), P8)                           # :192-197


def test_glynnsim_smoke(orc):
    # src/issue120-b5-glynnsim.gpu.test.ts:122-144 (finite inside and outside the radius)
    check_finite(orc, "glynnSim1", 1.0, (1.0, 0.1, 110.0, 0.1, 1.5, 0.5), P8)
    check_finite(orc, "glynnSim2", 1.0, (1.0, 0.1, 0.5, 1.5, 110.0, 150.0), P8)
    check_finite(orc, "glynnSim3", 1.0, (1.0, 0.1, 0.5, 1.5), P8)


# src/issue120-b6-faber-novelties.gpu.test.ts:19-27 (P8 without (2, 0))
P7 = [(0.5, 0.3), (-0.5, 0.7), (1.2, -0.4), (-1.0, -1.0), (0.1, 0.05), (0.0, 2.0), (-2.5, 1.7)]


def eclipse_ref(x, y, w, shift):   # src/issue120-b6-faber-novelties.gpu.test.ts:113-128
    if abs(y) <= w:
        c2 = math.sqrt(max(w * w - y * y, 0.0))
        if abs(x) <= c2:
            xs = x + shift * w
            return (-w * x if abs(xs) >= c2 else w * xs), w * y
    return w * x, w * y


def barycentroid_ref(x, y, w, a, b, c, d):   # :148-164
    d00, d01, d02 = a * a + b * b, a * c + b * d, a * x + b * y
    d11, d12 = c * c + d * d, c * x + d * y
    den = d00 * d11 - d01 * d01
    if abs(den) < 1e-30:
        return x * w, y * w
    u = (d11 * d02 - d01 * d12) / den
    v = (d00 * d12 - d01 * d02) / den
    sg = lambda t: (t > 0) - (t < 0)   # noqa: E731
    return w * math.sqrt(u * u + x * x) * sg(u), w * math.sqrt(v * v + y * y) * sg(v)


def test_faber_novelties(orc):
    check_pin(orc, "flipy", lambda x, y, w: (w * x, w * y * (-1 if x > 0 else 1)), 1.0, (), P7, 5)   # :78-97
    check_pin(orc, "eclipse", eclipse_ref, 1.0, (0.0,), P7, 5)                             # :99-132
    check_pin(orc, "barycentroid", barycentroid_ref, 1.0, (1.0, 0.0, 0.0, 1.0), P7, 4)     # :134-168
    # :170-196 chunk mode 0 (a = c = 1, f = -1): passthrough inside the unit circle, else 0
    check_pin(orc, "chunk", lambda x, y, w, *q: (x, y) if x * x + y * y - 1.0 <= 0 else (0.0, 0.0),
              1.0, (1.0, 0.0, 1.0, 0.0, 0.0, -1.0, 0.0), P7, 5)


def test_folds_against_the_mandelbox_definitions(orc):
    # src/issue129-folds.gpu.test.ts:101-127 (box_fold limit 1, sphere_fold (0.5, 1),
    # mandelbox_step (2, 0.5, 1, 0, 0), kifs_fold (3, 0)); Tom Lowe's Mandelbox:
    # box fold reflects a coordinate about +-L, sphere fold scales by R^2/r_min^2
    # inside r_min and by R^2/r^2 inside R, a step is c + s * sphere(box(z)).
    def box(t, lim):
        return 2 * lim - t if t > lim else (-2 * lim - t if t < -lim else t)

    def sphere(x, y, rmin, rmax):
        r2 = x * x + y * y
        s = rmax ** 2 / rmin ** 2 if r2 < rmin ** 2 else (rmax ** 2 / r2 if r2 < rmax ** 2 else 1.0)
        return x * s, y * s
    pts = [(0.5, 0.5), (1.5, -1.5), (0.75, 0.0), (2.0, 0.0), (0.2, -0.1), (-1.3, 0.4), (2, 2),
           (-0.5, -0.5)]
    check_pin(orc, "box_fold", lambda x, y, w, L: (w * box(x, L), w * box(y, L)), 1.0, (1.0,),
              pts, 5)
    check_pin(orc, "sphere_fold", lambda x, y, w, a, b: tuple(w * v for v in sphere(x, y, a, b)),
              1.0, (0.5, 1.0), pts, 5)

    def mstep(x, y, w, s, rmin, rmax, cx, cy):
        u, v = sphere(box(x, 1.0), box(y, 1.0), rmin, rmax)
        return w * (s * u + cx), w * (s * v + cy)
    check_pin(orc, "mandelbox_step", mstep, 1.0, (2.0, 0.5, 1.0, 0.0, 0.0), pts, 4)
    check_pin(orc, "mandelbox_step", mstep, 0.7, (-1.5, 0.3, 1.2, 0.2, -0.1), pts, 4)
    # kaleidoscopic fold: |z| kept, angle folded into [offset, offset + pi/n]
    for n, off in ((3.0, 0.0), (5.0, 0.4)):
        for x, y in pts + P8:
            o = run(orc, "kifs_fold", x, y, 1.0, (n, off))
            r = math.hypot(x, y)
            assert math.hypot(o[0], o[1]) == pytest.approx(r, rel=1e-12, abs=1e-12)
            a = (math.atan2(o[1], o[0]) - off) % (2 * PI)
            assert a <= PI / n + 1e-9 or r < 1e-6


def test_attractors_against_their_definitions(orc):
    # src/issue130-attractors.gpu.test.ts:99-112 (standard_map k = 1, de_jong
    # (-2.24, 0.43, -0.65, -2.43), ikeda u = 0.9; finite at (0,0), (1,1), (-0.5,-0.5)):
    # Chirikov's standard map, Peter de Jong's attractor and the Ikeda map.
    pts = [(0.0, 0.0), (1.0, 1.0), (-0.5, -0.5)] + P8
    check_pin(orc, "standard_map", lambda x, y, w, k: (w * (x + k * math.sin(y)),
                                                       w * (y + x + k * math.sin(y))),
              1.0, (1.0,), pts, 5)
    check_pin(orc, "de_jong", lambda x, y, w, a, b, c, d: (w * (math.sin(a * y) - math.cos(b * x)),
                                                           w * (math.sin(c * x) - math.cos(d * y))),
              1.0, (-2.24, 0.43, -0.65, -2.43), pts, 5)

    def ikeda(x, y, w, u):
        t = 0.4 - 6.0 / (1.0 + x * x + y * y)
        return w * (1 + u * (x * math.cos(t) - y * math.sin(t))), w * u * (x * math.sin(t) + y * math.cos(t))
    check_pin(orc, "ikeda", ikeda, 1.0, (0.9,), pts, 5)


# ------------------------------------------------------------ issue 131 ----
def test_modular_reference_expectations(orc):
    # src/issue131-modular.gpu.test.ts:136-142: jacobi_theta -> (1, 0) for im_floor 5 (q -> 0)
    o = run(orc, "jacobi_theta", f32(0.0), f32(0.0), 1.0, (5.0,), np.float32)
    assert abs(o[0] - 1.0) < 5e-5 and abs(o[1]) < 5e-5
    # :144-150 lambda(i) = 1/2; :152-162 real and in (0, 1) on the imaginary axis
    o = run(orc, "modular_lambda", 0.0, f32(0.98), 1.0, (f32(0.02),), np.float32)
    assert abs(o[0] - 0.5) < 5e-4 and abs(o[1]) < 5e-4
    for y in (0.5, 1.5, 2.5):
        o = run(orc, "modular_lambda", 0.0, y, 1.0, (0.1,), np.float32)
        assert 0 < o[0] < 1 and abs(o[1]) < 1e-3
    # :165-172 j(i) = 1728, compressed to log(1729) on the real axis
    o = run(orc, "klein_j", 0.0, f32(0.98), 1.0, (f32(0.02),), np.float32)
    assert abs(o[0] - math.log(1729)) < 5e-3 and abs(o[1]) < 5e-3
    for x, y, fl in ((0.0, 0.0, 0.05), (0.9, 0.1, 0.05), (-0.5, 0.5, 0.3), (0.3, -0.7, 0.1)):   # :174-179
        assert np.isfinite(run(orc, "klein_j", x, y, 1.0, (fl,), np.float32)[:2]).all()
    # :183-200 gauss_map = frac(1 / x) per axis, in [0, 1) and finite at the origin
    o = run(orc, "gauss_map", f32(0.3), f32(-0.4), 1.0, (), np.float32)
    assert abs(o[0] - (1 / 0.3 - math.floor(1 / 0.3))) < 5e-5 and abs(o[1] - 0.5) < 5e-5
    for x, y in ((0.0, 0.0), (1.7, -2.3), (0.001, 0.05)):
        o = run(orc, "gauss_map", f32(x), f32(y), 1.0, (), np.float32)
        assert all(np.isfinite(v) and 0 <= v < 1 for v in o[:2])
    # :204-221 weierstrass_p even on the square lattice, finite near a pole
    lat = (1.0, 0.0, 0.0, 1.0)
    a = run(orc, "weierstrass_p", f32(0.37), f32(0.21), 1.0, lat, np.float32)
    b = run(orc, "weierstrass_p", f32(-0.37), f32(-0.21), 1.0, lat, np.float32)
    assert abs(a[0] - b[0]) < 5e-5 and abs(a[1] - b[1]) < 5e-5
    for x, y in ((0.5, 0.5), (0.99, 0.01), (0.01, 0.01), (1.3, 0.7)):
        assert np.isfinite(run(orc, "weierstrass_p", x, y, 1.0, lat, np.float32)[:2]).all()


# ------------------------------------------------------------ issue 132 ----
def nbody_ref(x, y, w, c1x, c1y, c2x, c2y, m1, m2, g, eps):   # src/issue132-exotic.gpu.test.ts:108-114
    e = max(eps, 5e-3)

    def pull(cx, cy, m):
        dx, dy = cx - x, cy - y
        r2 = dx * dx + dy * dy + e
        inv = g * m / (r2 * math.sqrt(r2))
        return inv * dx, inv * dy
    a, b = pull(c1x, c1y, m1), pull(c2x, c2y, m2)
    return w * (x + a[0] + b[0]), w * (y + a[1] + b[1])


def _hash01(n):   # :115
    h = n & 0xFFFFFFFF
    h ^= h >> 17
    h = (h * 0xED5AD4BB) & 0xFFFFFFFF
    h ^= h >> 11
    h = (h * 0xAC4C1B51) & 0xFFFFFFFF
    h ^= h >> 15
    return h / 4294967296.0


def _corner(ix, iy):   # :116
    ux = ((ix & 0xFFFFFFFF) * 0x9E3779B1) & 0xFFFFFFFF
    uy = ((iy & 0xFFFFFFFF) * 0x85EBCA77) & 0xFFFFFFFF
    return _hash01(ux ^ uy) * 2 - 1


def _vnoise(px, py):   # :117
    fx, fy = math.floor(px), math.floor(py)
    ix, iy = int(fx), int(fy)
    tx, ty = px - fx, py - fy
    ux, uy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
    c00, c10, c01, c11 = _corner(ix, iy), _corner(ix + 1, iy), _corner(ix, iy + 1), _corner(ix + 1, iy + 1)
    bot, top = c00 + (c10 - c00) * ux, c01 + (c11 - c01) * ux
    return bot + (top - bot) * uy


def curl_noise_ref(x, y, w, freq, amp):   # :118-125
    f = max(freq, 1e-3)
    sx, sy, h = x * f, y * f, 1e-2
    ddx = (_vnoise(sx + h, sy) - _vnoise(sx - h, sy)) / (2 * h)
    ddy = (_vnoise(sx, sy + h) - _vnoise(sx, sy - h)) / (2 * h)
    dx, dy = min(8.0, max(-8.0, ddy)), min(8.0, max(-8.0, -ddx))
    return w * (x + amp * dx), w * (y + amp * dy)


def test_exotic(orc):
    P = (-0.4, 0.0, 0.4, 0.0, 0.06, 0.06, 1.0, 0.05)
    check_pin(orc, "nbody_lensing", nbody_ref, 1.0, P,
              [(0.5, 0.3), (-0.7, 0.4), (1.0, -0.5), (0.2, 0.6), (0.0, 0.5)], 4)   # :127-145
    assert abs(run(orc, "nbody_lensing", 0.0, 0.5, 1.0, P, np.float32)[0]) < 1e-4  # :140-145
    check_pin(orc, "nbody_lensing", nbody_ref, 1.0, P[:4] + (0.0, 0.0) + P[6:],
              [(0.3, 0.4), (-0.6, 0.2)], 5)                                       # :146-155
    assert np.isfinite(run(orc, "nbody_lensing", -0.4, 0.0, 1.0,
                           (-0.4, 0.0, 0.4, 0.0, 1.0, 1.0, 4.0, 0.005), np.float32)[:2]).all()
    # curl of value noise (the noise hash mirrored for negative cells), :166-188
    check_pin(orc, "curl_noise", curl_noise_ref, 1.0, (2.5, 0.3),
              [(0.5, 0.3), (-0.7, 0.4), (-1.2, -0.9), (0.9, -0.5)], 3, f64_rel=1e-7)
    check_pin(orc, "curl_noise", curl_noise_ref, 1.0, (2.5, 0.0), [(0.5, 0.3), (-0.7, 0.4)], 6)
    for k in range(13):   # :189-201 bounded displacement
        x, y = math.cos(0.5 * k), math.sin(0.5 * k)
        o = run(orc, "curl_noise", x, y, 1.0, (2.5, 0.3), np.float32)
        assert math.hypot(o[0] - x, o[1] - y) < 3.4


# ---------------------------------------------------- literal expectations --
# (entry, (x, y), params, (expected x, expected y), toBeCloseTo digits for x
# and y, reference file:line). The reference feeds (x, y, p0, p1) rows to
# var_<entry>(p, 1.0, p0, p1) (e.g. src/issue133-blaschke.gpu.test.ts:53).
E = math.e
SQRT_PI = math.sqrt(PI)
LITERALS = [
    ("blaschke", (0.0, 0.0), (0.5, 0.0), (0.0, 0.0), 5, 5, "issue133-blaschke.gpu.test.ts:87-96"),
    ("blaschke", (0.5, 0.0), (0.5, 0.0), (0.0, 0.0), 5, 5, "issue133-blaschke.gpu.test.ts:87-96"),
    ("blaschke", (1.0, 0.0), (0.5, 0.0), (1.0, 0.0), 4, 4, "issue133-blaschke.gpu.test.ts:98-105"),
    ("cayley", (0.0, 0.0), (1.0,), (-1.0, 0.0), 4, 4, "issue133-cayley.gpu.test.ts:84-96"),
    ("cayley", (1.0, 0.0), (1.0,), (0.0, -1.0), 4, 4, "issue133-cayley.gpu.test.ts:84-96"),
    ("cayley", (-1.0, 0.0), (1.0,), (0.0, 1.0), 4, 4, "issue133-cayley.gpu.test.ts:84-96"),
    ("cayley", (0.0, 1.0), (1.0,), (0.0, 0.0), 5, 5, "issue133-cayley.gpu.test.ts:98-105"),
    ("cayley", (0.0, 0.0), (2.0,), (-1.0, 0.0), 4, 4, "issue133-cayley.gpu.test.ts:116-126"),
    ("cayley", (0.0, 0.0), (0.5,), (-1.0, 0.0), 4, 4, "issue133-cayley.gpu.test.ts:116-126"),
    ("complex_gamma", (1.0, 0.0), (1.0,), (1.0, 0.0), 2, 2, "issue133-complex-gamma.gpu.test.ts:109-118"),
    ("complex_gamma", (2.0, 0.0), (1.0,), (1.0, 0.0), 2, 2, "issue133-complex-gamma.gpu.test.ts:109-118"),
    ("complex_gamma", (3.0, 0.0), (1.0,), (2.0, 0.0), 2, 2, "issue133-complex-gamma.gpu.test.ts:120-129"),
    ("complex_gamma", (4.0, 0.0), (1.0,), (6.0, 0.0), 1, 1, "issue133-complex-gamma.gpu.test.ts:120-129"),
    ("complex_gamma", (0.5, 0.0), (1.0,), (SQRT_PI, 0.0), 1, 2, "issue133-complex-gamma.gpu.test.ts:131-137"),
    ("complex_gamma", (1.0, 1.0), (1.0,), (0.4980, -0.1549), 2, 2, "issue133-complex-gamma.gpu.test.ts:139-145"),
    ("complex_gamma", (-0.5, 0.0), (1.0,), (-2 * SQRT_PI, 0.0), 0, 0, "issue133-complex-gamma.gpu.test.ts:147-154"),
    ("complex_gamma", (1.0, 0.0), (0.3,), (0.3, None), 2, 2, "issue133-complex-gamma.gpu.test.ts:156-162"),
    ("lambert_w", (0.0, 0.0), (3,), (0.0, 0.0), 3, 3, "issue133-lambert-w.gpu.test.ts:101-107"),
    ("lambert_w", (1.0, 0.0), (3,), (0.5671, 0.0), 2, 3, "issue133-lambert-w.gpu.test.ts:109-115"),
    ("lambert_w", (E, 0.0), (3,), (1.0, 0.0), 2, 3, "issue133-lambert-w.gpu.test.ts:117-123"),
    ("lambert_w", (-0.3, 0.0), (3,), (-0.4894, 0.0), 2, 3, "issue133-lambert-w.gpu.test.ts:125-131"),
    ("lambert_w", (1.0, 1.0), (4,), (0.6569, 0.3254), 2, 2, "issue133-lambert-w.gpu.test.ts:133-139"),
    ("lambert_w", (100.0, 0.0), (4,), (3.385, 0.0), 1, 1, "issue133-lambert-w.gpu.test.ts:141-147"),
    ("newton", (1.0, 0.0), (3,), (1.0, 0.0), 4, 4, "issue133-newton.gpu.test.ts:167-176"),
    ("newton", (2.0, 0.0), (3,), (17 / 12, 0.0), 3, 4, "issue133-newton.gpu.test.ts:178-186"),
    ("newton", (1.0, 0.0), (4,), (1.0, 0.0), 4, 4, "issue133-newton.gpu.test.ts:196-202"),
]
# (entry, points, params, reference file:line): finite output required
FINITE = [
    ("blaschke", [(2.0, 0.0), (1.95, 0.01)], (0.5, 0.0), "issue133-blaschke.gpu.test.ts:118-125"),
    ("blaschke", [(0.3, 0.4)], (0.5, 0.3), "issue133-blaschke.gpu.test.ts:127-135"),
    ("blaschke", [(-0.6, 0.2)], (0.1, -0.4), "issue133-blaschke.gpu.test.ts:127-135"),
    ("blaschke", [(0.7, -0.5)], (0.8, 0.0), "issue133-blaschke.gpu.test.ts:127-135"),
    ("blaschke", [(0.2, 0.9)], (-0.3, 0.5), "issue133-blaschke.gpu.test.ts:127-135"),
    ("cayley", [(0.0, -1.0), (0.01, -0.99)], (1.0,), "issue133-cayley.gpu.test.ts:107-114"),
    ("complex_gamma", [(0.5, 0.5), (-0.3, 0.7), (1.2, -0.4), (0.1, 0.05)], (0.3,),
     "issue133-complex-gamma.gpu.test.ts:164-172"),
    ("lambert_w", P8[:6], (2,), "issue133-lambert-w.gpu.test.ts:149-159"),
    ("newton", [(0.0, 0.0), (1e-9, 0.0)], (3,), "issue133-newton.gpu.test.ts:188-194"),
    ("newton", [(0.5, 0.5), (-0.3, 0.7)], (3,), "issue133-newton.gpu.test.ts:204-211"),
    ("newton", [(1.2, -0.4)], (5,), "issue133-newton.gpu.test.ts:204-211"),
]


@pytest.mark.parametrize("case", LITERALS, ids=lambda c: f"{c[0]}@{c[6].split(':')[1]}")
def test_literal_expectation(orc, case):
    name, (x, y), params, want, dx, dy, src = case
    got = run(orc, name, f32(x), f32(y), 1.0, params, np.float32)
    for k, (w, d) in enumerate(zip(want, (dx, dy))):
        if w is not None:
            assert abs(got[k] - w) < 0.5 * 10.0 ** -d, (src, k, got[k], w)


@pytest.mark.parametrize("case", FINITE, ids=lambda c: f"{c[0]}@{c[3].split(':')[1]}")
def test_finite_expectation(orc, case):
    name, pts, params, src = case
    for x, y in pts:
        assert np.isfinite(run(orc, name, f32(x), f32(y), 1.0, params, np.float32)[:2]).all(), src


def test_reference_properties_issue133(orc):
    # src/issue133-blaschke.gpu.test.ts:107-116: |B(z)| = 1 on the unit circle
    for (x, y), a in (((0.0, 1.0), (0.5, 0.0)), ((-1.0, 0.0), (0.3, 0.2)), ((0.6, 0.8), (0.4, 0.1))):
        o = run(orc, "blaschke", x, y, 1.0, a, np.float32)
        assert abs(math.hypot(o[0], o[1]) - 1.0) < 5e-3
    # src/issue133-cayley.gpu.test.ts:128-139: upper half-plane into the unit disk
    for x, y in ((0.5, 1.0), (-0.3, 2.0), (1.0, 3.0)):
        o = run(orc, "cayley", x, y, 1.0, (1.0,), np.float32)
        assert math.hypot(o[0], o[1]) < 1.0
    # src/issue133-lambert-w.gpu.test.ts:161-183: w e^w = z after the iteration
    for x, y in ((2.0, 0.0), (0.7, 0.4)):
        wr, wi = run(orc, "lambert_w", x, y, 1.0, (4,), np.float32)[:2]
        ez = complex(wr, wi) * np.exp(complex(wr, wi))
        assert abs(ez.real - x) < 0.05 and abs(ez.imag - y) < 0.05
    # src/issue133-newton.gpu.test.ts:220-292: basin colours of z^n - 1
    c0 = orc.dc_color(V["newton"], 0.95, 0.0, (3.0,))
    c1 = orc.dc_color(V["newton"], -0.5, 0.85, (3.0,))
    c2 = orc.dc_color(V["newton"], -0.5, -0.85, (3.0,))
    assert c0[0] > c0[1] and c0[0] > c0[2] and c0[0] > 0.9
    assert c1[1] > c1[0] and c1[1] > c1[2]
    assert c2[2] > c2[0] and c2[2] > c2[1]

    def distinct(a, b):
        return any(abs(u - v) > 0.1 for u, v in zip(a, b))
    assert distinct(c0, c1) and distinct(c1, c2) and distinct(c0, c2)
    c4 = [orc.dc_color(V["newton"], x, y, (4.0,)) for x, y in ((0.95, 0), (0, 0.95), (-0.95, 0), (0, -0.95))]
    assert all(distinct(c4[i], c4[j]) for i in range(4) for j in range(i + 1, 4))


# ---------------------------------------------- issues 134, 135: textbook --
def test_map_projections_against_textbook(orc):
    """src/issue134-cartographic.gpu.test.ts:99-123 only asks for finite
    output at (0,0), (1,1), (-0.5,-0.5), (0,1.5), (0,pi/2); the entries are the
    textbook projections of (longitude x, latitude y) with a 1e-6 guard in
    each denominator, so they are held to the textbook formulas."""
    from scipy.optimize import brentq
    pts = [(0.0, 0.0), (1.0, 1.0), (-0.5, -0.5), (0.3, 1.2), (-1.1, 0.4), (0.7, -0.9)] + P8[:5]

    def mercator(x, y, w):
        return w * x, w * math.log(math.tan(PI / 4 + min(max(y, -1.5), 1.5) / 2))

    def lambert(x, y, w):   # azimuthal equal-area, centred on (0, 0)
        k = math.sqrt(2 / (1 + math.cos(y) * math.cos(x)))
        return w * k * math.cos(y) * math.sin(x), w * k * math.sin(y)

    def hammer(x, y, w):
        z = math.sqrt(1 + math.cos(y) * math.cos(x / 2))
        return w * 2 * math.sqrt(2) * math.cos(y) * math.sin(x / 2) / z, w * math.sqrt(2) * math.sin(y) / z

    def stereo(x, y, w):
        k = 2 / (1 + math.cos(y) * math.cos(x))
        return w * k * math.cos(y) * math.sin(x), w * k * math.sin(y)

    def mollweide(x, y, w):   # 2t + sin 2t = pi sin(lat)
        t = brentq(lambda t: 2 * t + math.sin(2 * t) - PI * math.sin(y), -PI / 2, PI / 2) \
            if abs(y) > 0 else 0.0
        return w * 2 * math.sqrt(2) / PI * x * math.cos(t), w * math.sqrt(2) * math.sin(t)
    for name, ref in (("mercator", mercator), ("lambert", lambert), ("hammer", hammer),
                      ("stereographic", stereo), ("mollweide", mollweide)):
        check_pin(orc, name, ref, 1.0, (), pts, 4, f64_rel=2e-5)
    for x, y in ((0.0, 1.5), (0.0, PI / 2)):
        assert np.isfinite(run(orc, "mercator", x, y, 1.0, (), np.float32)[:2]).all()
        assert np.isfinite(run(orc, "mollweide", x, y, 1.0, (), np.float32)[:2]).all()


def test_plane_curves_against_textbook(orc):
    """src/issue135-plane-curves.gpu.test.ts:106-140 (superellipse (1, 1, 2)
    maps (1,0), (0,1) to themselves; limacon (1, 0.5) at theta 0 -> 1.5;
    epicycloid k = 3 at theta 0 -> 3; catenary a = 1 at 0 -> (0, 1); tractrix
    finite at (1, 1)), and the curves' definitions at the angle of the input:
    the Lame curve |cos/a|^n + |sin/b|^n = r^-n, the limacon r = b + a cos t,
    the epicycloid ((k+1) cos t - cos (k+1)t, (k+1) sin t - sin (k+1)t), the
    catenary y = a cosh(x / a), the tractrix (x - tanh x, sech x)."""
    def superellipse(x, y, w, a, b, n):
        t = math.atan2(y, x)
        r = (abs(math.cos(t) / a) ** n + abs(math.sin(t) / b) ** n) ** (-1 / n)
        return w * r * math.cos(t), w * r * math.sin(t)

    def limacon(x, y, w, a, b):
        t = math.atan2(y, x)
        r = b + a * math.cos(t)
        return w * r * math.cos(t), w * r * math.sin(t)

    def epicycloid(x, y, w, k):
        t = math.atan2(y, x)
        return (w * ((k + 1) * math.cos(t) - math.cos((k + 1) * t)),
                w * ((k + 1) * math.sin(t) - math.sin((k + 1) * t)))
    pts = [(1.0, 0.0), (0.0, 1.0)] + P8
    check_pin(orc, "superellipse", superellipse, 1.0, (1.0, 1.0, 2.0), pts, 4, f64_rel=1e-6)
    check_pin(orc, "superellipse", superellipse, 1.0, (1.5, 0.7, 3.5), pts, 4, f64_rel=1e-6)
    check_pin(orc, "limacon", limacon, 1.0, (1.0, 0.5), pts, 5)
    check_pin(orc, "epicycloid", epicycloid, 1.0, (3.0,), pts, 5)
    check_pin(orc, "catenary", lambda x, y, w, a: (w * x, w * a * math.cosh(x / a)), 1.0, (1.0,),
              [(0.0, 0.0)] + P8, 4)
    check_pin(orc, "tractrix", lambda x, y, w: (w * (x - math.tanh(x)), w / math.cosh(x)), 1.0, (),
              [(1.0, 1.0)] + P8, 5)
    assert run(orc, "limacon", 1.0, 0.0, 1.0, (1.0, 0.5))[:2] == pytest.approx((1.5, 0.0))
    assert run(orc, "epicycloid", 1.0, 0.0, 1.0, (3.0,))[:2] == pytest.approx((3.0, 0.0))


# ------------------------------------------------------------ issue 137 ----
def _j0_ref(x):   # src/issue137-special-functions.gpu.test.ts:113-123 (A&S 9.4.1 / 9.4.3 in (3/x)^2)
    ax = abs(x)
    if ax < 3.0:
        t = x * x / 9.0
        return 1.0 + t * (-2.2499997 + t * (1.2656208 + t * (-0.3163866 + t * (0.0444479 + t * (
            -0.0039444 + t * 0.0002100)))))
    t = (3.0 / ax) ** 2
    amp = 0.79788456 + t * (-0.00000077 + t * (-0.00552740 + t * (-0.00009512 + t * (
        0.00137237 + t * (-0.00072805 + t * 0.00014476)))))
    ph = ax - 0.78539816 + t * (-0.04166397 + t * (-0.00003954 + t * (0.00262573 + t * (
        -0.00054125 + t * (-0.00029333 + t * 0.00013558)))))
    return amp / math.sqrt(ax) * math.cos(ph)


def _airy_ref(x):   # :124-131
    if x > 4.0:
        xi = 2 / 3 * x ** 1.5
        return math.exp(-xi) / (2 * 1.7724539 * x ** 0.25)
    if x < -5.0:
        xi = 2 / 3 * (-x) ** 1.5
        return math.sin(xi + 0.78539816) / (1.7724539 * (-x) ** 0.25)
    x3 = x ** 3
    f = tf = 1.0
    g = tg = x
    for k in range(1, 12):
        tf *= x3 / ((3 * k - 1) * (3 * k))
        f += tf
        tg *= x3 / ((3 * k) * (3 * k + 1))
        g += tg
    return 0.355028053887817 * f - 0.258819403792807 * g


def _struve_h1_ref(x):   # :132-138
    h = 0.5 * min(10.0, max(-10.0, x))
    term = h * h / (0.8862269 * 1.3293404)
    s = term
    for m in range(1, 16):
        term = term * -(h * h) / ((m + 0.5) * (m + 1.5))
        s += term
    return s


def _cornu_ref(x, y, w, freq):   # :214-222 (rational Fresnel fit)
    t = freq * x
    s, sg = abs(t), (-1 if t < 0 else 1)
    f = (1 + 0.926 * s) / (2 + 1.792 * s + 3.104 * s * s)
    g = 1 / (2 + 4.142 * s + 3.492 * s * s + 6.670 * s ** 3)
    a = PI * s * s / 2
    c = 0.5 + f * math.sin(a) - g * math.cos(a)
    sn = 0.5 - f * math.cos(a) - g * math.sin(a)
    return w * sg * c, w * (sg * sn + 0.25 * y)


def test_special_function_entries_against_the_reference_oracles(orc):
    P4 = [(0.5, 0.3), (-0.7, 0.4), (1.0, -0.5), (0.0, 0.6)]
    sweep = [(0.1 + 0.2 * k, 0.0) for k in range(15)]

    def radial(fn, gain):
        return lambda x, y, w, *q: tuple(w * gain(*q) * fn(math.hypot(x, y), *q) * v for v in (x, y))
    # bessel_j0, :140-170: J0(freq r) p; origin fixed; collapse near the first zero
    check_pin(orc, "bessel_j0", lambda x, y, w, f: (w * _j0_ref(f * math.hypot(x, y)) * x,
                                                    w * _j0_ref(f * math.hypot(x, y)) * y),
              1.0, (4.0,), P4 + sweep, 4)
    assert math.hypot(*run(orc, "bessel_j0", 2.4048 / 4, 0.0, 1.0, (4.0,), np.float32)[:2]) < 0.02
    # airy_radial, :172-207: 3 Ai(scale (r - shift)) p, bounded by 2 |p|
    check_pin(orc, "airy_radial", lambda x, y, w, sc, sh: tuple(
        w * 3 * _airy_ref(sc * (math.hypot(x, y) - sh)) * v for v in (x, y)), 1.0, (3.0, 1.5),
        P4 + sweep, 4)
    assert abs(_airy_ref(0) - 0.3550280) < 5e-6 and abs(_airy_ref(8)) < 1e-4
    for x, y in sweep:
        o = run(orc, "airy_radial", x, y, 1.0, (3.0, 1.5), np.float32)
        assert math.hypot(o[0], o[1]) < 2 * x + 1e-3
    # cornu_spiral, :209-248
    check_pin(orc, "cornu_spiral", _cornu_ref, 1.0, (1.6,), P4 + [(0.0, 0.8), (12.5, 0.0)], 4)
    o = run(orc, "cornu_spiral", 12.5, 0.0, 1.0, (1.6,), np.float32)
    assert abs(o[0] - 0.5) < 0.05 and abs(o[1] - 0.5) < 0.05
    # struve_h1, :250-287: 0.6 H1(freq r) p, factor below 0.7
    check_pin(orc, "struve_h1", lambda x, y, w, f: tuple(
        w * 0.6 * _struve_h1_ref(f * math.hypot(x, y)) * v for v in (x, y)), 1.0, (2.5,),
        P4 + sweep + [(0.0, 0.0)], 4)
    for x, y in sweep:
        o = run(orc, "struve_h1", x, y, 1.0, (2.5,), np.float32)
        assert math.hypot(o[0], o[1]) / x < 0.7
    assert radial


# ------------------------------------------------------------ issue 138 ----
def lorentz_ref(x, y, w, rap, ang):   # src/variations.ts:2844-2858
    c, s = math.cos(ang), math.sin(ang)
    rx, ry = x * c + y * s, -x * s + y * c
    xp, yp = rx * math.cosh(rap) + ry * math.sinh(rap), rx * math.sinh(rap) + ry * math.cosh(rap)
    return w * (xp * c - yp * s), w * (xp * s + yp * c)


def schwarzschild_ref(x, y, w, mass, eps):   # src/variations.ts:2861-2871
    a = mass / (math.hypot(x, y) + max(eps, 1e-6))
    return w * (x * math.cos(a) - y * math.sin(a)), w * (x * math.sin(a) + y * math.cos(a))


def dipole_ref(x, y, w, q, sep, step, ang):   # src/variations.ts:2874-2895
    d = sep * 0.5
    px, py = d * math.cos(ang), d * math.sin(ang)
    rp, rn = (x - px, y - py), (x + px, y + py)
    dp3 = (rp[0] ** 2 + rp[1] ** 2 + 1e-4) ** 1.5
    dn3 = (rn[0] ** 2 + rn[1] ** 2 + 1e-4) ** 1.5
    ex, ey = q * (rp[0] / dp3 - rn[0] / dn3), q * (rp[1] / dp3 - rn[1] / dn3)
    return w * (x + step * ex), w * (y + step * ey)


def pendulum_ref(x, y, w, magnets, radius, strength, damping):   # src/variations.ts:2898-2921
    n = max(3, min(6, int(math.trunc(magnets))))
    fx = fy = 0.0
    for k in range(n):
        th = 2.0 * PI * k / n
        rx, ry = radius * math.cos(th) - x, radius * math.sin(th) - y
        den = (rx * rx + ry * ry + 1e-4) ** 1.5
        fx += rx / den
        fy += ry / den
    return w * (x + strength * fx - damping * x), w * (y + strength * fy - damping * y)


def test_physics_entries_against_the_reference_oracles(orc):
    # src/issue138-physics.gpu.test.ts (toBeCloseTo 4; identity / radius properties at 5 / 4)
    P4a = [(0.0, 0.0), (0.5, 0.0), (0.3, 0.4), (-1.2, 0.7)]
    check_pin(orc, "lorentz_boost", lorentz_ref, 1.0, (0.5, 0.3), P4a + P8, 4)          # :111-131
    check_pin(orc, "lorentz_boost", lorentz_ref, 1.0, (0.0, 0.5),
              [(0.4, 0.6), (-0.3, 0.2), (1.1, -0.8)], 5)                                  # :133-143
    check_pin(orc, "schwarzschild_lensing", schwarzschild_ref, 1.0, (0.5, 0.05), P4a + P8, 4)   # :145-167
    for x, y in ((0.4, 0.6), (-0.3, 0.2), (1.1, -0.8)):                                   # :169-182
        o = run(orc, "schwarzschild_lensing", x, y, 1.0, (0.7, 0.05), np.float32)
        assert abs(math.hypot(o[0], o[1]) - math.hypot(x, y)) < 5e-5
    for x, y in ((1e-7, 0.0), (0.0, -1e-7), (3e-7, 4e-7)):                               # :184-196
        assert np.isfinite(run(orc, "schwarzschild_lensing", x, y, 1.0, (2.0, 0.0), np.float32)[:2]).all()
    check_pin(orc, "field_dipole", dipole_ref, 1.0, (1.0, 0.5, 0.2, 0.0),
              [(0.5, 0.3), (-0.7, 0.4), (1.0, -0.5), (0.0, 0.6)] + P8, 4, f64_rel=1e-8)   # :198-221
    for x, y in ((0.3, 0.05), (-0.3, -0.05)):                                             # :223-234
        assert np.isfinite(run(orc, "field_dipole", x, y, 1.0, (1.0, 0.5, 0.2, 0.0), np.float32)[:2]).all()
    check_pin(orc, "magnetic_pendulum", pendulum_ref, 1.0, (3.0, 1.0, 0.5, 0.1),
              [(0.0, 0.0), (0.5, 0.3), (-0.7, 0.8), (1.2, -0.5)] + P8, 4, f64_rel=1e-8)   # :236-257
    for x, y in ((0.4, 0.2), (-0.6, 0.7), (0.1, -1.0)):                                   # :259-267
        assert np.isfinite(run(orc, "magnetic_pendulum", x, y, 1.0, (6.0, 1.0, 0.5, 0.1), np.float32)[:2]).all()
    for hi, ref in ((15.0, 6.0), (1.0, 3.0)):                                             # :269-283
        a = run(orc, "magnetic_pendulum", 0.3, 0.4, 1.0, (hi, 1.0, 0.5, 0.1))
        b = run(orc, "magnetic_pendulum", 0.3, 0.4, 1.0, (ref, 1.0, 0.5, 0.1))
        assert a[:2] == b[:2]


# ------------------------------------------------------- issues 139 - 141 --
LITERALS += [
    ("tinkerbell", (0.5, 0.5), (0.9, -0.6, 2.0, 0.5), (0.15, 1.75), 4, 4, "issue139-flows.gpu.test.ts:101-109"),
    ("duffing", (1.0, 1.0), (0.1, 0.1, 0.1, 1.0), (1.1, 1.0 + 0.1 * (-0.1 + 0.1 * math.cos(1.0))), 4, 4,
     "issue139-flows.gpu.test.ts:111-120"),
    ("vanderpol", (0.5, 0.5), (0.1, 1.0), (0.55, 0.4875), 4, 4, "issue139-flows.gpu.test.ts:122-130"),
    ("rossler", (1.0, 0.0), (0.1, 0.2), (0.9, 0.1), 4, 4, "issue139-flows.gpu.test.ts:132-141"),
    ("bakers_map", (0.25, 0.25), (), (0.0, 0.375), 2, 2, "issue140-toral.gpu.test.ts:98-106"),
    ("tent_map", (0.0, 0.0), (), (0.5, 0.5), 2, 2, "issue140-toral.gpu.test.ts:108-113"),
    ("logistic_map", (0.0, 0.0), (3.9,), (0.475, 0.475), 2, 2, "issue140-toral.gpu.test.ts:115-121"),
]


def test_flows_and_toral_maps(orc):
    # the Tinkerbell map (src/issue139-flows.gpu.test.ts:101-109 states it)
    check_pin(orc, "tinkerbell", lambda x, y, w, a, b, c, d: (w * (x * x - y * y + a * x + b * y),
                                                              w * (2 * x * y + c * x + d * y)),
              1.0, (0.9, -0.6, 2.0, 0.5), P8, 4)
    # src/issue140-toral.gpu.test.ts:90-96: arnold_cat lands in [-0.5, 0.5]^2
    for x, y in [(10.5, -2.1), (0.1, 0.1)] + P8:
        o = run(orc, "arnold_cat", f32(x), f32(y), 1.0, (), np.float32)
        assert -0.5 <= o[0] <= 0.5 and -0.5 <= o[1] <= 0.5


def _scramble_ref(bits, extent, perm):   # src/issue141-digit-scramble.gpu.test.ts:121-126
    levels = 2 ** bits
    s = max(extent, 1e-4)
    mask = levels - 1

    def enc(c):
        nrm = (c + s) / (2 * s)
        return min(int((nrm - math.floor(nrm)) * levels), levels - 1)

    def dec(i):
        return (i + 0.5) / levels * 2 * s - s
    return lambda x, y, w, *q: (w * dec(perm(enc(x)) & mask), w * dec(perm(enc(y)) & mask))


def test_digit_scrambles(orc):
    def rev(bits):   # :127
        return lambda v: int(format(v, f"0{bits}b")[::-1], 2)
    pts3 = [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)]
    check_pin(orc, "radical_inverse", _scramble_ref(12, 1.0, rev(12)), 1.0, (1.0, 12.0),
              pts3 + [(0.13, 0.77)], 4)                                          # :141-152
    check_pin(orc, "gray_code", _scramble_ref(12, 1.0, lambda v: v ^ (v >> 1)), 1.0,
              (1.0, 12.0), pts3, 4)                                              # :172-184

    def morton(x, y, w, extent, bits_f):   # :196-212
        bits = int(bits_f)
        levels, s = 2 ** bits, max(extent, 1e-4)

        def enc(c):
            nrm = (c + s) / (2 * s)
            return min(int((nrm - math.floor(nrm)) * levels), levels - 1)

        def part(v):
            return sum(((v >> i) & 1) << (2 * i) for i in range(bits))

        def dec(i):
            return (i + 0.5) / levels * 2 * s - s
        code = part(enc(x)) | (part(enc(y)) << 1)
        return w * dec(code & (levels - 1)), w * dec((code >> bits) & (levels - 1))
    check_pin(orc, "morton_zorder", morton, 1.0, (1.0, 8.0), pts3, 4)
    for name, prm in (("radical_inverse", (1.0, 12.0)), ("morton_zorder", (1.0, 8.0))):   # :153-161, :214-221
        for k in range(13):
            v = -3 + 0.5 * k
            o = run(orc, name, f32(v), f32(-v), 1.0, prm, np.float32)
            assert all(np.isfinite(t) and abs(t) < 1.0 for t in o[:2])


# ------------------------------------------------------- issues 142, 143 --
GRID142 = [(0.5, 0.7), (-1.2, 0.4), (0.3, -2.1), (2.5, 1.1), (-3, 0.5), (1.7, -0.8), (0.05, 0.05),
           (4, 1.5)]   # src/issue142-number-theoretic.gpu.test.ts:85-87
GRID143 = [(0, 0), (0.5, 0.7), (-1.2, 0.4), (0.3, -2.1), (3, 3), (-3, -3), (8, -8), (0.001, 0.001)]
LITERALS += [
    ("collatz", (2.0, 0.0), (1.0, 0.0), (1.0, 0.0), 3, 3, "issue142-number-theoretic.gpu.test.ts:89-95"),
    ("collatz", (3.0, 0.0), (1.0, 0.0), (5.0, 0.0), 3, 3, "issue142-number-theoretic.gpu.test.ts:96-102"),
    ("digamma", (1.0, 0.0), (1.0, 0.0), (-0.5772157, 0.0), 3, 3, "issue142-number-theoretic.gpu.test.ts:116-122"),
    ("digamma", (2.0, 0.0), (1.0, 0.0), (0.4227843, 0.0), 3, 3, "issue142-number-theoretic.gpu.test.ts:123-129"),
    ("quasicrystal", (0.0, 0.0), (5.0, 3.0), (0.0, 0.0), 5, 5, "issue143-aperiodic.gpu.test.ts:84-91"),
]
FINITE += [
    ("collatz", GRID142, (1.0, 0.0), "issue142-number-theoretic.gpu.test.ts:103-113"),
    ("digamma", GRID142, (1.0, 0.0), "issue142-number-theoretic.gpu.test.ts:130-139"),
    ("quasicrystal", GRID143, (5.0, 3.0), "issue143-aperiodic.gpu.test.ts:92-103"),
    ("penrose", GRID143, (1.0, 0.2), "issue143-aperiodic.gpu.test.ts:115-126"),
]


def test_number_theoretic_and_aperiodic_properties(orc):
    def moved(name, p, a, b, tol):
        u = run(orc, name, *p, 1.0, a, np.float32)
        v = run(orc, name, *p, 1.0, b, np.float32)
        return math.hypot(u[0] - v[0], u[1] - v[1]) > tol
    # src/issue142-number-theoretic.gpu.test.ts:109-112, :135-138: scale / shift have effect
    assert moved("collatz", (0.6, 0.4), (1.0, 0.0), (1.5, 0.3), 1e-3)
    assert moved("digamma", (0.6, 0.4), (1.0, 0.0), (1.5, 0.3), 1e-3)
    # digamma on the real axis against scipy's psi (recurrence + asymptotic series)
    from scipy.special import digamma as psi
    for x in (0.7, 1.5, 3.2, 6.0):
        assert run(orc, "digamma", x, 0.0, 1.0, (1.0, 0.0))[0] == pytest.approx(psi(x), abs=1e-4)
    # src/issue143-aperiodic.gpu.test.ts:92-112: displacement at most 0.4; n = 5 and 7 differ
    agg = 0.0
    for x, y in GRID143:
        a = run(orc, "quasicrystal", f32(x), f32(y), 1.0, (5.0, 3.0), np.float32)
        b = run(orc, "quasicrystal", f32(x), f32(y), 1.0, (7.0, 3.0), np.float32)
        assert math.hypot(a[0] - x, a[1] - y) <= 0.4001
        agg += math.hypot(a[0] - b[0], a[1] - b[1])
    assert agg > 1e-3
    # :115-143 penrose: |out| <= 5 |in| + 5, lattice snap, scale has effect
    for x, y in GRID143:
        o = run(orc, "penrose", f32(x), f32(y), 1.0, (1.0, 0.2), np.float32)
        assert math.hypot(o[0], o[1]) <= 5 * math.hypot(x, y) + 5
    a = run(orc, "penrose", f32(0.30), f32(0.30), 1.0, (1.0, 0.2), np.float32)
    b = run(orc, "penrose", f32(0.3001), f32(0.3001), 1.0, (1.0, 0.2), np.float32)
    assert abs(a[0] - b[0]) < 5e-6 and abs(a[1] - b[1]) < 5e-6
    assert moved("penrose", (1.3, -0.7), (1.0, 0.2), (2.0, 0.2), 1e-4)


# ------------------------------------------------------------ issue 144 ----
def _cheb_t(v, n):   # src/issue144-polynomials.gpu.test.ts:112-118 (argument clamped to [-1, 1])
    c = max(-1.0, min(1.0, v))
    if n <= 0:
        return 1.0
    a, b, t = 1.0, c, c
    for _ in range(2, n + 1):
        t = 2 * c * b - a
        a, b = b, t
    return t


def _leg_p(v, n):   # :119-125
    c = max(-1.0, min(1.0, v))
    if n <= 0:
        return 1.0
    a, b, p = 1.0, c, c
    for k in range(1, n):
        p = ((2 * k + 1) * c * b - k * a) / (k + 1)
        a, b = b, p
    return p


def test_polynomial_warps(orc):
    check_pin(orc, "chebyshev", lambda x, y, w, nx, ny: (w * _cheb_t(x, int(nx)), w * _cheb_t(y, int(ny))),
              1.0, (4.0, 3.0), [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5), (1.4, -1.8)], 4)   # :127-137
    assert abs(run(orc, "chebyshev", f32(math.cos(0.4)), 0.0, 1.0, (3.0, 0.0), np.float32)[0]
               - math.cos(1.2)) < 5e-5                                                     # :138-143
    p = run(orc, "chebyshev", 0.6, 0.6, 1.0, (2.0, 3.0))
    n = run(orc, "chebyshev", -0.6, -0.6, 1.0, (2.0, 3.0))
    assert n[0] == pytest.approx(p[0]) and n[1] == pytest.approx(-p[1])                   # :144-151
    sweep = [(-1 + 0.25 * k, -1 + 0.25 * k) for k in range(9)]
    for x, y in sweep:                                                                     # :152-159
        o = run(orc, "chebyshev", f32(x), f32(y), 1.0, (12.0, 11.0), np.float32)
        assert abs(o[0]) <= 1 + 1e-5 and abs(o[1]) <= 1 + 1e-5
        o = run(orc, "legendre", f32(x), f32(y), 1.0, (9.0, 8.0), np.float32)           # :185-192
        assert abs(o[0]) <= 1 + 1e-5 and abs(o[1]) <= 1 + 1e-5
    check_pin(orc, "legendre", lambda x, y, w, nx, ny: (w * _leg_p(x, int(nx)), w * _leg_p(y, int(ny))),
              1.0, (5.0, 4.0), [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)], 4)                  # :162-172
    assert abs(run(orc, "legendre", 0.6, 0.0, 1.0, (2.0, 0.0), np.float32)[0] - 0.04) < 5e-5   # :173-178
    assert abs(run(orc, "legendre", 1.0, 0.0, 1.0, (5.0, 0.0), np.float32)[0] - 1.0) < 5e-5    # :179-184
    pts3 = [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)]
    check_pin(orc, "spherical_harmonic", lambda x, y, w, *q: (w * x, w * y), 1.0, (0.0, 0.0, 1.0),
              pts3, 4, f64_rel=1e-6)                                                       # :196-205
    for x, y in pts3:                                                                      # :206-215
        o = run(orc, "spherical_harmonic", f32(x), f32(y), 1.0, (3.0, 2.0, 1.0), np.float32)
        assert abs(o[0] * y - o[1] * x) < 1e-4
    assert run(orc, "spherical_harmonic", 0.0, 0.0, 1.0, (3.0, 2.0, 1.0), np.float32)[:2] == (0.0, 0.0)   # :216-222
    for k in range(5):                                                                     # :223-236
        r = 0.2 + 0.3 * k
        o = run(orc, "spherical_harmonic", f32(0.6 * r), f32(0.8 * r), 1.0, (6.0, 6.0, 2.0), np.float32)
        assert np.isfinite(o[:2]).all() and math.hypot(o[0], o[1]) <= 3 * r + 1e-4
    check_pin(orc, "fourier_warp", lambda x, y, w, *q: (w * x, w * y), 1.0, (4.0, 0.0, 1.0), pts3, 4)   # :239-248
    for k in range(16):                                                                    # :249-261
        x, y = math.cos(0.4 * k), math.sin(0.4 * k)
        o = run(orc, "fourier_warp", f32(x), f32(y), 1.0, (4.0, 0.4, 1.0), np.float32)
        assert abs(o[0] * y - o[1] * x) < 1e-4 and o[0] * x + o[1] * y >= 0
    pts = [(0.7, 0.4), (-0.5, 0.6), (0.3, -0.9)]                                           # :262-271
    assert any(abs(a - b) > 1e-4 for x, y in pts
               for a, b in zip(run(orc, "fourier_warp", x, y, 1.0, (4.0, 0.4, 1.0))[:2],
                               run(orc, "fourier_warp", x, y, 1.0, (4.0, 0.4, 7.0))[:2]))


# ------------------------------------------------------------ issue 145 ----
# escape-time single steps: var_<entry>(z, 1.0, c_re, c_im[, relax])
LITERALS += [
    ("burning_ship", (0.3, 0.4), (-0.5, -0.5), (-0.57, -0.26), 4, 4, "issue145-escape-time.gpu.test.ts:88-93"),
    ("magnet1", (1.0, 0.0), (0.0, 0.0), (1.0, 0.0), 4, 4, "issue145-escape-time.gpu.test.ts:102-107"),
    ("magnet1", (0.5, 0.0), (0.0, 0.0), (0.5625, 0.0), 4, 4, "issue145-escape-time.gpu.test.ts:108-114"),
    ("nova", (1.0, 0.0), (0.0, 0.0, 1.0), (1.0, 0.0), 4, 4, "issue145-escape-time.gpu.test.ts:115-119"),
    ("nova", (0.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0), 6, 6, "issue145-escape-time.gpu.test.ts:120-124"),
    ("halley", (1.0, 0.0), (0.0, 0.0), (1.0, 0.0), 4, 4, "issue145-escape-time.gpu.test.ts:125-129"),
]
_GRID145 = [(-1.2 + 0.3 * i, -1.2 + 0.3 * j) for i in range(9) for j in range(9)]
FINITE += [
    ("burning_ship", _GRID145, (-0.5, -0.5), "issue145-escape-time.gpu.test.ts:130-144"),
    ("magnet1", _GRID145, (0.5, 0.3), "issue145-escape-time.gpu.test.ts:130-144"),
    ("nova", _GRID145, (0.0, 0.0, 1.0), "issue145-escape-time.gpu.test.ts:130-144"),
    ("halley", _GRID145, (0.0, 0.0), "issue145-escape-time.gpu.test.ts:130-144"),
]


def test_escape_time_steps(orc):
    # one burning-ship step (|Re z| + i |Im z|)^2 + c (the escape-time test's closed form)
    def bship(x, y, w, cr, ci):
        a, b = abs(x), abs(y)
        return w * (a * a - b * b + cr), w * (2 * a * b + ci)
    check_pin(orc, "burning_ship", bship, 1.0, (-0.5, -0.5), P8 + [(0.3, 0.4)], 5)
    a = run(orc, "burning_ship", -0.3, 0.4, 1.0, (0.0, 0.0))                     # :94-100
    b = run(orc, "burning_ship", 0.3, -0.4, 1.0, (0.0, 0.0))
    assert a[:2] == pytest.approx(b[:2], abs=1e-12)
    # src/issue145-escape-color.gpu.test.ts: in-gamut, finite band colours; a fast
    # escaper and a converger get different hues
    for idx, c, relax in ((V["burning_ship"], (-0.5, -0.5), 0.0), (V["magnet1"], (0.5, 0.3), 0.0),
                          (V["nova"], (0.0, 0.0), 1.0), (V["halley"], (0.0, 0.0), 0.0)):
        rgb = orc.dc_color(idx, 0.4, -0.3, (c[0], c[1], relax))
        assert rgb is not None and all(np.isfinite(v) and 0.0 <= v <= 1.0 for v in rgb)
    fast = orc.dc_color(V["burning_ship"], 3.0, 3.0, (-0.5, -0.5, 0.0))
    slow = orc.dc_color(V["burning_ship"], 0.0, 0.0, (-0.5, -0.5, 0.0))
    assert sum(abs(u - v) for u, v in zip(fast, slow)) > 0.05
    assert all(0.0 <= v <= 1.0 for v in orc.dc_color(V["burning_ship"], 0.3, 0.4, (-0.5, -0.5, 0.0)))


# ------------------------------------------------------- issues 146, 147 --
def r32(orc, name, x, y, params):
    return run(orc, name, f32(x), f32(y), 1.0, params, np.float32)


def test_optics_entries(orc):
    # src/issue146-optics.gpu.test.ts V284 snell_refraction: |out| = |p|; n 1 keeps p;
    # total internal reflection flips the normal component
    for x, y in ((0.5, 0.3), (-0.7, 0.4), (0.9, -0.5), (0.2, 0.95)):
        o = r32(orc, "snell_refraction", x, y, (0.67, 1.0))
        assert abs(math.hypot(o[0], o[1]) - math.hypot(x, y)) < 5e-5
    for x, y in ((0.5, 0.3), (0.9, -0.4)):
        o = r32(orc, "snell_refraction", x, y, (1.0, 1.0))
        assert abs(o[0] - x) < 5e-5 and abs(o[1] - y) < 5e-5
    o = r32(orc, "snell_refraction", 0.95, 0.31, (2.0, 1.0))
    assert o[1] < 0 and abs(math.hypot(o[0], o[1]) - math.hypot(0.95, 0.31)) < 5e-5
    # V285 grin_lens: collinear with the input, origin fixed, bounded near the
    # focus, thin lens at large r
    for x, y in ((0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)):
        o = r32(orc, "grin_lens", x, y, (0.6, 0.15))
        assert abs(o[0] * y - o[1] * x) < 1e-4
    assert r32(orc, "grin_lens", 0.0, 0.0, (0.6, 0.15))[:2] == (0.0, 0.0)
    o = r32(orc, "grin_lens", 0.001, 0.0, (0.6, 0.15))
    assert np.isfinite(o[0]) and abs(o[0]) < 1.0
    assert abs(r32(orc, "grin_lens", 2.0, 0.0, (0.6, 0.15))[0] - 2.0) < 0.02
    # V286 caustic_fold: displacement <= amp sqrt 2, amp 0 identity, separable,
    # x-displacement -amp sin(freq x) at freq 8 (safe_sin native branch)
    for x, y in ((0.5, 0.3), (-0.7, 0.4), (0.9, -0.5), (1.2, 0.8)):
        o = r32(orc, "caustic_fold", x, y, (2.5, 0.35, 0.0))
        assert math.hypot(o[0] - x, o[1] - y) <= 0.35 * math.sqrt(2) + 1e-4
    for x, y in ((0.5, 0.3), (-0.7, 0.4)):
        o = r32(orc, "caustic_fold", x, y, (2.5, 0.0, 0.0))
        assert abs(o[0] - x) < 5e-6 and abs(o[1] - y) < 5e-6
    a, b = r32(orc, "caustic_fold", 0.5, 0.3, (2.5, 0.35, 0.0)), r32(orc, "caustic_fold", 0.5, -0.9, (2.5, 0.35, 0.0))
    assert abs((a[0] - 0.5) - (b[0] - 0.5)) < 5e-6
    assert abs(r32(orc, "caustic_fold", 5.0, 0.0, (8.0, 0.35, 0.0))[0] - (5.0 - 0.35 * math.sin(40.0))) < 5e-4


def _chladni_ref(x, y, w, n, m, step):   # src/issue147-waves.gpu.test.ts jsChladni
    an, am = PI * n, PI * m
    f = math.cos(an * x) * math.cos(am * y) - math.cos(am * x) * math.cos(an * y)
    gx = -an * math.sin(an * x) * math.cos(am * y) + am * math.sin(am * x) * math.cos(an * y)
    gy = -am * math.cos(an * x) * math.sin(am * y) + an * math.cos(am * x) * math.sin(an * y)
    s = -1 if f < 0 else 1
    gx, gy = gx * s, gy * s
    inv = 1 / (math.hypot(gx, gy) + 1e-4)
    return w * (x - step * gx * inv), w * (y - step * gy * inv)


def _moire_ref(x, y, w, freq, beat, angle, amp):   # jsMoire
    ca, sa = math.cos(angle), math.sin(angle)
    u, v = x * ca + y * sa, -x * sa + y * ca
    gx = math.sin(freq * PI * x) * math.sin((freq + beat) * PI * u)
    gy = math.sin(freq * PI * y) * math.sin((freq + beat) * PI * v)
    return w * (x + amp * gx), w * (y + amp * gy)


def test_wave_entries(orc):

    def node(x, y):
        return abs(math.cos(3 * PI * x) * math.cos(5 * PI * y) - math.cos(5 * PI * x) * math.cos(3 * PI * y)) < 0.15
    check_pin(orc, "chladni", lambda x, y, w, n, m, st: _chladni_ref(x, y, w, int(n), int(m), st), 1.0,
              (3.0, 5.0, 0.18), [(0.37, 0.62), (-0.41, 0.23), (0.55, -0.71), (0.18, 0.44)], 4, skip=node)
    for x, y in ((0.4, 0.7), (-0.6, 0.2)):   # (n, m) and (m, n) descend alike
        a, b = r32(orc, "chladni", x, y, (3.0, 5.0, 0.18)), r32(orc, "chladni", x, y, (5.0, 3.0, 0.18))
        assert abs(a[0] - b[0]) < 5e-5 and abs(a[1] - b[1]) < 5e-5
    for i in range(7):
        for j in range(4):
            x, y = -0.9 + 0.3 * i, -0.9 + 0.6 * j
            o = r32(orc, "chladni", x, y, (3.0, 5.0, 0.18))
            assert math.hypot(o[0] - x, o[1] - y) < 0.18 + 1e-3
    assert r32(orc, "chladni", 0.5, 0.3, (3.4, 5.0, 0.18))[:2] == \
        pytest.approx(r32(orc, "chladni", 0.5, 0.3, (2.6, 5.0, 0.18))[:2], abs=5e-6)
    # V288 standing_wave: amp 0 identity, origin node, modes gate, bounded by amp / (1 - decay)
    for x, y in ((0.5, 0.3), (-0.7, 0.4)):
        o = r32(orc, "standing_wave", x, y, (3.0, 1.0, 0.0, 0.6))
        assert abs(o[0] - x) < 5e-6 and abs(o[1] - y) < 5e-6
    assert r32(orc, "standing_wave", 0.0, 0.0, (3.0, 1.0, 0.35, 0.6))[:2] == (0.0, 0.0)
    m1, m4 = r32(orc, "standing_wave", 0.6, 0.4, (1.0, 1.0, 0.35, 0.6)), r32(orc, "standing_wave", 0.6, 0.4, (4.0, 1.0, 0.35, 0.6))
    assert abs(m1[0] - m4[0]) + abs(m1[1] - m4[1]) > 1e-4
    for i in range(5):
        for j in range(3):
            x, y = -0.9 + 0.45 * i, -0.9 + 0.9 * j
            o = r32(orc, "standing_wave", x, y, (3.0, 1.0, 0.35, 0.6))
            assert abs(o[0] - x) < 0.35 / 0.4 + 1e-3 and abs(o[1] - y) < 0.35 / 0.4 + 1e-3
            o = r32(orc, "moire", x, y, (4.0, 0.6, 0.4, 0.25))
            assert abs(o[0] - x) <= 0.25 + 1e-4 and abs(o[1] - y) <= 0.25 + 1e-4
    check_pin(orc, "moire", _moire_ref, 1.0, (4.0, 0.6, 0.4, 0.25),
              [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)] + P8, 4)
    check_pin(orc, "moire", _moire_ref, 1.0, (4.0, 0.6, 0.4, 0.0), [(0.5, 0.3)], 5)
    assert r32(orc, "moire", 0.0, 0.0, (4.0, 0.6, 0.4, 0.25))[:2] == (0.0, 0.0)


# ------------------------------------------------- issues 148, 149, 150 ----
LITERALS += [   # var_<entry>(p, 1.0, params) with src/issue149-spirals.gpu.test.ts:55-60's params
    ("droste", (1.0, 0.0), (2.0,), (1.0, 0.0), 2, 2, "issue149-spirals.gpu.test.ts:99-105"),
    ("droste", (0.0, 1.0), (2.0,), (0.0, 0.8408964), 2, 5, "issue149-spirals.gpu.test.ts:99-105"),
    ("logspiral", (1.0, 1.0), (2.0, 0.5), (2.0944111, 2.0944111), 5, 5, "issue149-spirals.gpu.test.ts:106-110"),
    ("fermat_spiral", (0.0, 1.0), (2.0,), (0.0, 2.5066282), 5, 5, "issue149-spirals.gpu.test.ts:111-119"),
    ("fermat_spiral", (0.0, -1.0), (2.0,), (0.0, -4.3416075), 5, 5, "issue149-spirals.gpu.test.ts:111-119"),
    ("fermat_spiral", (1.0, -1.0), (2.0,), (3.3159577, -3.3159577), 5, 5, "issue149-spirals.gpu.test.ts:111-119"),
    ("lituus", (0.0, 1.0), (2.0,), (0.0, 1.5957691), 5, 5, "issue149-spirals.gpu.test.ts:120-128"),
    ("lituus", (0.0, -1.0), (2.0,), (0.0, -0.9213177), 5, 5, "issue149-spirals.gpu.test.ts:120-128"),
    ("lituus", (1.0, -1.0), (2.0,), (0.6031440, -0.6031440), 5, 5, "issue149-spirals.gpu.test.ts:120-128"),
    ("hyperbolic_spiral", (0.0, 1.0), (2.0,), (0.0, 1.2732395), 5, 5, "issue149-spirals.gpu.test.ts:129-137"),
    ("hyperbolic_spiral", (0.0, -1.0), (2.0,), (0.0, -0.4244132), 5, 5, "issue149-spirals.gpu.test.ts:129-137"),
    ("hyperbolic_spiral", (1.0, -1.0), (2.0,), (0.2572333, -0.2572333), 5, 5, "issue149-spirals.gpu.test.ts:129-137"),
]


def test_orbital_entries(orc):
    # src/issue148-orbitals.gpu.test.ts: 1s monotone, 2s node at r = 2, bounded
    # by 1.5, angle kept; hydrogen_orbital with lobe_mix 0 is radial_shell, the
    # p orbital's node on the y axis
    def rad(o):
        return math.hypot(o[0], o[1])
    r = [rad(r32(orc, "radial_shell", x, 0.0, (1.0, 0.0, 1.0))) for x in (0.2, 0.8, 1.5)]
    assert r[0] > r[1] > r[2]
    assert rad(r32(orc, "radial_shell", 2.0, 0.0, (2.0, 0.0, 1.0))) < 0.02
    assert rad(r32(orc, "radial_shell", 0.5, 0.0, (2.0, 0.0, 1.0))) > 0.05
    for n, l in ((1, 0), (2, 1), (3, 0), (3, 2), (4, 3)):
        for k in range(11):
            assert rad(r32(orc, "radial_shell", 0.3 * k, 0.4 * k, (n, l, 1.0))) <= 1.5 + 1e-4
    for x, y in ((0.5, 0.3), (-0.7, 0.4), (0.2, -0.9)):
        o = r32(orc, "radial_shell", x, y, (3.0, 0.0, 1.0))
        if rad(o) >= 1e-4:
            assert abs(math.atan2(o[1], o[0]) - math.atan2(y, x)) < 5e-5
        h = r32(orc, "hydrogen_orbital", x, y, (3.0, 2.0, 0.0, 1.0, 0.0))
        s = r32(orc, "radial_shell", x, y, (3.0, 2.0, 1.0))
        assert abs(h[0] - s[0]) < 5e-6 and abs(h[1] - s[1]) < 5e-6
    assert rad(r32(orc, "hydrogen_orbital", 0.0, 0.5, (2.0, 1.0, 0.0, 1.0, 1.0))) < 0.02
    assert rad(r32(orc, "hydrogen_orbital", 0.5, 0.0, (2.0, 1.0, 0.0, 1.0, 1.0))) > 0.02
    for n, l, m in ((1, 0, 0), (2, 1, 0), (3, 2, 0), (4, 3, 0)):
        for i in range(9):
            for j in range(8):
                rr, a = 0.2 + 0.6 * i, 0.78 * j
                o = r32(orc, "hydrogen_orbital", rr * math.cos(a), rr * math.sin(a), (n, l, m, 1.0, 1.0))
                assert np.isfinite(rad(o)) and rad(o) <= 1.5 + 1e-4


def billiard_circle_ref(x, y, w, radius, step, angle):   # src/variations.ts:2475-2522
    rc = max(radius, 1e-5)
    px, py = x, y
    r0 = math.hypot(px, py)
    if r0 > rc:
        px, py, r0 = px / r0 * rc, py / r0 * rc, rc
    bx, by = (px / r0, py / r0) if r0 > 1e-6 else (1.0, 0.0)
    vx = bx * math.cos(angle) - by * math.sin(angle)
    vy = bx * math.sin(angle) + by * math.cos(angle)
    dpv = px * vx + py * vy
    t = -dpv + math.sqrt(max(dpv * dpv + rc * rc - (px * px + py * py), 0.0))
    fx, fy = px + step * vx, py + step * vy
    if 0.0 < t < step:
        hx, hy = px + t * vx, py + t * vy
        nx, ny = -hx / rc, -hy / rc
        d = vx * nx + vy * ny
        rx, ry = vx - 2 * d * nx, vy - 2 * d * ny
        fx, fy = hx + (step - t) * rx, hy + (step - t) * ry
    return w * fx, w * fy


def billiard_polygon_ref(x, y, w, sides_f, radius, step, angle):   # src/variations.ts:2752-2840
    sides = max(3, min(12, int(math.floor(sides_f + 0.5))))
    R = max(radius, 1e-5)
    inr = R * math.cos(PI / sides)
    px, py = x, y
    r0 = math.hypot(px, py)
    if r0 > inr:
        s = inr / max(r0, 1e-6)
        px, py = px * s, py * s
    vx, vy = math.cos(angle), math.sin(angle)
    vs = [(R * math.cos(2 * PI * k / sides), R * math.sin(2 * PI * k / sides)) for k in range(sides)]
    t_hit, seg = math.inf, -1
    for k in range(sides):
        (ax, ay), (bx, by) = vs[k], vs[(k + 1) % sides]
        dx, dy = bx - ax, by - ay
        det = -vx * dy + vy * dx
        if abs(det) > 1e-6:
            t = (-(ax - px) * dy + (ay - py) * dx) / det
            u = (vx * (ay - py) - vy * (ax - px)) / det
            if t >= 0 and 0 <= u <= 1 and t < t_hit:
                t_hit, seg = t, k
    fx, fy = px + step * vx, py + step * vy
    if 0.0 < t_hit < step and seg != -1:
        hx, hy = px + t_hit * vx, py + t_hit * vy
        (ax, ay), (bx, by) = vs[seg], vs[(seg + 1) % sides]
        ex, ey = bx - ax, by - ay
        el = max(math.hypot(ex, ey), 1e-5)
        nx, ny = -ey / el, ex / el
        if nx * (ax + bx) / 2 + ny * (ay + by) / 2 > 0:
            nx, ny = -nx, -ny
        d = vx * nx + vy * ny
        rx, ry = vx - 2 * d * nx, vy - 2 * d * ny
        fx, fy = hx + (step - t_hit) * rx, hy + (step - t_hit) * ry
    return w * fx, w * fy


def _reflect(px, py, vx, vy, step, t, nx, ny):
    """Flight of `step` from p along v, reflected off a wall met at 0 < t < step."""
    if not 0.0 < t < step:
        return px + step * vx, py + step * vy
    hx, hy = px + t * vx, py + t * vy
    d = vx * nx + vy * ny
    return hx + (step - t) * (vx - 2 * d * nx), hy + (step - t) * (vy - 2 * d * ny)


def billiard_stadium_ref(x, y, w, width, height, step, angle):   # src/variations.ts:2525-2636
    R, w2 = max(height / 2, 1e-5), max(width / 2, 1e-5)
    px, py = x, y
    if abs(px) > w2:
        cx = w2 if px > 0 else -w2
        d = math.hypot(px - cx, py)
        if d > R:
            px, py = cx + (px - cx) / d * R, py / d * R
    elif abs(py) > R:
        py = R if py > 0 else -R
    vx, vy = math.cos(angle), math.sin(angle)
    best, wall = math.inf, 0
    for k, (sgn, ok) in enumerate(((1, vy > 0), (-1, vy < 0)), start=1):   # top, bottom
        if ok:
            t = (sgn * R - py) / vy
            if t >= 0 and -w2 <= px + t * vx <= w2 and t < best:
                best, wall = t, k
    for k, c in ((3, w2), (4, -w2)):   # right, left caps
        dx = px - c
        dpv = dx * vx + py * vy
        disc = dpv * dpv + R * R - (dx * dx + py * py)
        if disc >= 0:
            t = -dpv + math.sqrt(disc)
            xx = px + t * vx
            if t >= 0 and (xx >= w2 - 1e-5 if k == 3 else xx <= -w2 + 1e-5) and t < best:
                best, wall = t, k
    nx = ny = 0.0
    if 0.0 < best < step:
        hx, hy = px + best * vx, py + best * vy
        if wall in (1, 2):
            ny = -1.0 if wall == 1 else 1.0
        else:
            c = w2 if wall == 3 else -w2
            rl = max(math.hypot(hx - c, hy), 1e-5)
            nx, ny = -(hx - c) / rl, -hy / rl
    fx, fy = _reflect(px, py, vx, vy, step, best, nx, ny)
    return w * fx, w * fy


def billiard_sinai_ref(x, y, w, length, radius, step, angle):   # src/variations.ts:2639-2750
    L2, R = max(length / 2, 1e-5), max(radius, 1e-5)
    px, py = min(max(x, -L2), L2), min(max(y, -L2), L2)
    r0 = math.hypot(px, py)
    assert r0 > 1e-6   # the origin start takes a hashed direction: not restated here
    if r0 < R:
        px, py = px / r0 * R, py / r0 * R
    vx, vy = math.cos(angle), math.sin(angle)
    best, wall = math.inf, 0
    if vx < 0:
        t = (-L2 - px) / vx
        if t >= 0 and -L2 <= py + t * vy <= L2 and t < best:
            best, wall = t, 1
    if vx > 0:
        t = (L2 - px) / vx
        if t >= 0 and -L2 <= py + t * vy <= L2 and t < best:
            best, wall = t, 2
    if vy < 0:
        t = (-L2 - py) / vy
        if t >= 0 and -L2 <= px + t * vx <= L2 and t < best:
            best, wall = t, 3
    if vy > 0:
        t = (L2 - py) / vy
        if t >= 0 and -L2 <= px + t * vx <= L2 and t < best:
            best, wall = t, 4
    dpv = px * vx + py * vy
    disc = dpv * dpv - (px * px + py * py - R * R)
    if disc >= 0:
        t1 = -dpv - math.sqrt(disc)
        if 0 <= t1 < best:
            best, wall = t1, 5
    nx, ny = {1: (1.0, 0.0), 2: (-1.0, 0.0), 3: (0.0, 1.0), 4: (0.0, -1.0)}.get(wall, (0.0, 0.0))
    if wall == 5:
        nx, ny = (px + best * vx) / R, (py + best * vy) / R
    fx, fy = _reflect(px, py, vx, vy, step, best, nx, ny)
    return w * fx, w * fy


def test_billiards(orc):
    # src/issue150-billiards.gpu.test.ts:128-206 (toBeCloseTo 4) against the
    # reference's f64 oracles (src/variations.ts), at the test's points and
    # parameters and at the shared points with other ones
    check_pin(orc, "billiard_circle", billiard_circle_ref, 1.0, (1.0, 0.5, 0.0),
              [(0.0, 0.0), (0.5, 0.0), (0.9, 0.1), (1.2, -0.5)], 4)
    # extra points and parameters in f64 only: a start clamped onto a wall meets it
    # at t ~ 0, where f32 rounding decides whether the flight reflects
    ex = dict(f64_rel=1e-7, f32=False)
    check_pin(orc, "billiard_circle", billiard_circle_ref, 1.0, (1.2, 0.7, 2.2), P8, 4, **ex)
    check_pin(orc, "billiard_polygon", billiard_polygon_ref, 1.0, (5.0, 1.0, 0.5, 0.7),
              [(0.0, 0.0), (0.5, -0.2), (-0.8, 0.8)], 4)
    check_pin(orc, "billiard_polygon", billiard_polygon_ref, 1.0, (7.0, 1.3, 0.9, -1.1), P8, 4, **ex)
    check_pin(orc, "billiard_stadium", billiard_stadium_ref, 1.0, (1.5, 1.0, 0.5, 0.5),
              [(0.0, 0.0), (0.7, 0.4), (1.0, -0.2), (-1.5, 1.2)], 4)
    check_pin(orc, "billiard_stadium", billiard_stadium_ref, 1.0, (1.5, 1.0, 0.5, 0.5), P8, 4, **ex)
    check_pin(orc, "billiard_stadium", billiard_stadium_ref, 1.0, (2.2, 1.4, 0.9, -2.4), P8, 4, **ex)
    check_pin(orc, "billiard_sinai", billiard_sinai_ref, 1.0, (2.0, 0.5, 0.5, 0.3),
              [(0.7, 0.7), (-0.9, -0.9), (0.3, 0.4)], 4)
    check_pin(orc, "billiard_sinai", billiard_sinai_ref, 1.0, (2.0, 0.5, 0.5, 0.3), P8, 4, **ex)
    check_pin(orc, "billiard_sinai", billiard_sinai_ref, 1.0, (2.4, 0.3, 0.8, 2.0), P8, 4, **ex)
    # the origin start (:170) is finite
    assert np.isfinite(r32(orc, "billiard_sinai", 0.0, 0.0, (2.0, 0.5, 0.5, 0.3))[:2]).all()


# ------------------------------------------- radial inverse CDFs (V292-V295) --
# The quantile functions are the textbook ones (Weibull, logistic, Cauchy,
# Pareto) applied to u = r^2 / (1 + r^2), clamped as the test's uMap does
# (src/issue151-cdf.gpu.test.ts:81).
def _umap(r2):
    return min(max(r2 / (1.0 + r2), 1e-4), 1.0 - 1e-4)


def _radial(x, y, w, rp):
    r = math.hypot(x, y)
    if r < 1e-9:
        return 0.0, 0.0
    return w * rp * x / r, w * rp * y / r


def weibull_cdf_ref(x, y, w, lam, k):
    u = _umap(x * x + y * y)
    return _radial(x, y, w, lam * (-math.log1p(-u)) ** (1.0 / max(abs(k), 0.05)))


def logistic_cdf_ref(x, y, w, mu, s):
    u = _umap(x * x + y * y)
    return _radial(x, y, w, max(mu + s * min(max(math.log(u / (1.0 - u)), -12.0), 12.0), 0.0))


def cauchy_cdf_ref(x, y, w, gamma, tail):
    u = _umap(x * x + y * y)
    cap = max(tail, 0.01)
    return _radial(x, y, w, min(max(gamma * math.tan(PI * (u - 0.5)), -cap), cap))


def pareto_cdf_ref(x, y, w, xm, alpha, tail):
    u = _umap(x * x + y * y)
    return _radial(x, y, w, min(xm * (1.0 - u) ** (-1.0 / max(abs(alpha), 0.05)), max(tail, 0.01)))


def test_radial_inverse_cdfs(orc):
    # src/issue151-cdf.gpu.test.ts:84-150 (toBeCloseTo 4 / 3)
    pts = [(0.0, 0.0), (0.5, 0.0), (0.0, 0.7), (0.3, 0.0), (0.9, 0.0)]
    check_pin(orc, "weibull_cdf", weibull_cdf_ref, 1.0, (0.6, 1.5), pts + P8, 4, f64_rel=1e-7)
    o = [r32(orc, "weibull_cdf", x, y, (0.6, 1.5)) for x, y in pts]
    assert math.hypot(*o[0][:2]) < 1e-6 and abs(o[1][1]) < 1e-5 and abs(o[2][0]) < 1e-5
    assert math.hypot(*o[4][:2]) > math.hypot(*o[3][:2])        # monotone in r
    check_pin(orc, "logistic_cdf", logistic_cdf_ref, 1.0, (0.0, 0.35),
              [(1.0, 0.0), (0.3, 0.0), (50.0, 0.0)] + P8, 3, f64_rel=1e-7)
    assert math.hypot(*r32(orc, "logistic_cdf", 1.0, 0.0, (0.0, 0.35))[:2]) < 1e-4
    assert math.hypot(*r32(orc, "logistic_cdf", 50.0, 0.0, (0.0, 0.35))[:2]) < \
        0.35 * math.log(9999) + 1e-3                                   # the logit cap
    check_pin(orc, "logistic_cdf", logistic_cdf_ref, 1.0, (0.4, -0.8), P8, 4, f64_rel=1e-7)
    check_pin(orc, "cauchy_cdf", cauchy_cdf_ref, 1.0, (0.25, 4.0),
              [(1.0, 0.0), (0.5, 0.0), (30.0, 0.0)] + P8, 4, f64_rel=1e-7)
    assert math.hypot(*r32(orc, "cauchy_cdf", 1.0, 0.0, (0.25, 4.0))[:2]) < 1e-4   # median
    assert abs(math.hypot(*r32(orc, "cauchy_cdf", 30.0, 0.0, (0.25, 4.0))[:2]) - 4.0) < 5e-5
    check_pin(orc, "pareto_cdf", pareto_cdf_ref, 1.0, (0.3, 1.6, 4.0),
              [(0.0, 0.0), (0.2, 0.0), (50.0, 0.0)] + P8, 4, f64_rel=1e-7)
    assert math.hypot(*r32(orc, "pareto_cdf", 0.0, 0.0, (0.3, 1.6, 4.0))[:2]) == 0.0
    assert math.hypot(*r32(orc, "pareto_cdf", 0.2, 0.0, (0.3, 1.6, 4.0))[:2]) >= 0.3 - 1e-4


# ---------------------------------------------------- wavelets (V296-V298) --
def morlet_ref(x, y, w, freq, sigma, amp):
    r = math.hypot(x, y)
    t = r / max(sigma, 0.05)
    return w * (1.0 + amp * math.exp(-0.5 * min(t * t, 80.0)) * math.cos(freq * r)) * x, \
        w * (1.0 + amp * math.exp(-0.5 * min(t * t, 80.0)) * math.cos(freq * r)) * y


def mexican_hat_ref(x, y, w, sigma, amp):
    t2 = min((x * x + y * y) / max(sigma, 0.05) ** 2, 80.0)
    s = 1.0 + amp * (1.0 - t2) * math.exp(-0.5 * t2)
    return w * s * x, w * s * y


def chirp_ref(x, y, w, alpha, amp, decay):
    r2 = x * x + y * y
    s = 1.0 + amp * math.exp(-decay * min(r2, 160.0)) * math.sin(alpha * r2)
    return w * s * x, w * s * y


def test_wavelets(orc):
    # src/issue152-wavelets.gpu.test.ts:81-140 (toBeCloseTo 4 / 5): the
    # Morlet, Ricker and linear-chirp wavelets as radial gains 1 + amp * psi
    pts = [(0.0, 0.0), (0.5, 0.3), (-0.7, 0.4), (3.0, 4.0)]
    check_pin(orc, "morlet", morlet_ref, 1.0, (6.0, 0.5, 0.35), pts + P8, 4, f64_rel=1e-7)
    for x, y in pts[1:3]:
        o = r32(orc, "morlet", x, y, (6.0, 0.5, 0.35))
        assert abs(o[0] * y - o[1] * x) < 1e-5                 # direction kept
    o = r32(orc, "morlet", 1.0, 0.0, (2.0e6, 0.5, 0.35))      # past the trig cliff
    assert np.isfinite(o[:2]).all() and abs(o[0]) <= 1.35 + 1e-3
    sig = 0.45
    hat_pts = [(0.0, 0.0), (sig, 0.0), (sig * math.sqrt(3), 0.0), (0.3, 0.5), (-0.3, -0.5)]
    check_pin(orc, "mexican_hat", mexican_hat_ref, 1.0, (sig, 0.4), hat_pts + P8, 5, f64_rel=1e-7)
    assert abs(r32(orc, "mexican_hat", sig, 0.0, (sig, 0.4))[0] - sig) < 5e-6   # zero crossing
    assert abs(r32(orc, "mexican_hat", sig * math.sqrt(3), 0.0, (sig, 0.4))[0]) < sig * math.sqrt(3)
    a = r32(orc, "mexican_hat", 0.3, 0.5, (sig, 0.4))
    b = r32(orc, "mexican_hat", -0.3, -0.5, (sig, 0.4))
    assert a[0] == -b[0] and a[1] == -b[1]                     # odd map, exactly
    check_pin(orc, "chirp", chirp_ref, 1.0, (8.0, 0.3, 0.6), [(0.0, 0.0)] + P8, 4, f64_rel=1e-7)
    check_pin(orc, "chirp", chirp_ref, 1.0, (8.0, 0.3, 4.0), [(2.0, 0.0)], 4)
    o = r32(orc, "chirp", 0.0, 0.0, (8.0, 0.3, 0.6))
    assert o[0] == 0.0 and o[1] == 0.0
    o = r32(orc, "chirp", 1.0, 0.0, (2.0e7, 0.3, 0.0))        # alpha r^2 past the cliff
    assert np.isfinite(o[:2]).all() and abs(o[0]) <= 1.3 + 1e-3


# -------------------------------------------- celestial mechanics (V299-V301) --
def kepler_ref(x, y, w, e_in, scale):
    """Two Newton steps on Kepler's equation M = E - e sin E from E = M, the
    polar angle as mean anomaly (src/issue153-celestial.gpu.test.ts:79-84)."""
    e = min(max(e_in, 0.0), 0.95)
    m = math.atan2(y, x)
    a = scale * (math.hypot(x, y) + 1e-6)
    ea = m
    for _ in range(2):
        ea -= (ea - e * math.sin(ea) - m) / max(1.0 - e * math.cos(ea), 0.05)
    return w * a * (math.cos(ea) - e), w * a * math.sqrt(max(1.0 - e * e, 0.0)) * math.sin(ea)


def three_body_ref(x, y, w, mu_in, step, coriolis):
    """A step along the gradient of the rotating-frame effective potential of
    the restricted three-body problem, softened by 1e-3, plus a Coriolis turn,
    the displacement capped at 2 (src/issue153-celestial.gpu.test.ts:85-92)."""
    mu = min(max(mu_in, 0.01), 0.5)
    r1x, r2x = x + mu, x - (1.0 - mu)
    d1 = (r1x * r1x + y * y + 1e-3) ** 1.5
    d2 = (r2x * r2x + y * y + 1e-3) ** 1.5
    gx = x - (1.0 - mu) * r1x / d1 - mu * r2x / d2
    gy = y - (1.0 - mu) * y / d1 - mu * y / d2
    dx, dy = step * (gx - coriolis * gy), step * (gy + coriolis * gx)
    dl = math.hypot(dx, dy)
    if dl > 2.0:
        dx, dy = dx * 2.0 / dl, dy * 2.0 / dl
    return w * (x + dx), w * (y + dy)


def hill_ref(x, y, w, kappa, phase, shear):
    """Hill's linearised epicycle propagator over phase*kappa plus the
    guiding-centre shear (src/issue153-celestial.gpu.test.ts:93-96)."""
    k = min(max(kappa, 0.2), 2.0)
    c, s = math.cos(phase * k), math.sin(phase * k)
    return w * (x * c - s / k * y), w * (2.0 * k * s * x + y * c + shear * x)


def test_celestial(orc):
    # src/issue153-celestial.gpu.test.ts:98-170
    circ = [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)]
    for x, y in circ:                                            # e = 0: the circle
        o = r32(orc, "kepler_orbit", x, y, (0.0, 1.0))
        assert abs(o[0] - x) < 5e-5 and abs(o[1] - y) < 5e-5
    check_pin(orc, "kepler_orbit", kepler_ref, 1.0, (0.5, 0.6),
              [(0.5, 0.0), (0.4, 0.6), (-0.8, 0.2)] + P8, 3, f64_rel=1e-7)
    o = r32(orc, "kepler_orbit", 0.5, 0.0, (0.5, 0.6))           # perihelion a(1-e)
    assert abs(o[0] - 0.15) < 5e-4 and abs(o[1]) < 1e-3
    check_pin(orc, "kepler_orbit", kepler_ref, 1.0, (0.9, 1.3), P8, 3, f64_rel=1e-7)
    check_pin(orc, "restricted_3body", three_body_ref, 1.0, (0.2, 0.15, 0.0),
              [(0.5, 0.3), (-0.4, 0.5), (0.2, -0.6)] + P8, 4, f64_rel=1e-7)
    check_pin(orc, "restricted_3body", three_body_ref, 1.0, (0.2, 0.15, 0.4),
              [(0.8, 0.0), (0.5, 0.3)] + P8, 4, f64_rel=1e-7)
    o = r32(orc, "restricted_3body", 0.5, 0.3, (0.2, 0.15, 0.4))
    assert math.hypot(o[0] - 0.5, o[1] - 0.3) <= 0.15 * 2.0 + 1e-4
    for x, y in [(0.5, 0.3), (-0.7, 0.4)]:                       # phase 0: shear only
        o = r32(orc, "hill_epicyclic", x, y, (1.0, 0.0, 0.5))
        assert abs(o[0] - x) < 5e-6 and abs(o[1] - (y + 0.5 * x)) < 5e-6
    check_pin(orc, "hill_epicyclic", hill_ref, 1.0, (1.0, 1.2, 0.5), [(0.5, 0.3), (1.0, 0.6)] + P8, 4)
    a = r32(orc, "hill_epicyclic", 0.5, 0.3, (1.0, 1.2, 0.5))
    b = r32(orc, "hill_epicyclic", 1.0, 0.6, (1.0, 1.2, 0.5))
    assert abs(b[0] - 2 * a[0]) < 5e-5 and abs(b[1] - 2 * a[1]) < 5e-5   # linear
    check_pin(orc, "hill_epicyclic", hill_ref, 1.0, (3.0, -2.2, -0.7), P8, 4)


# ------------------------------------- conformal maps (V317 SC, V318 Doyle) --
def sc_series_ref(x, y, w, sides):
    """Ten terms of the Schwarz-Christoffel disk -> regular n-gon map
    w(z) = int_0^z (1 - t^n)^(-2/n) dt as its binomial series
    sum_k (2/n)_k / k! z^(nk+1) / (nk+1), |z| clamped to 0.999
    (src/shaders/chaos.wgsl:7012-7020 describes the series)."""
    n = max(round(sides), 3)
    z = complex(x, y)
    if abs(z) > 0.999:
        z *= 0.999 / abs(z)
    acc, c = 0j, 1.0
    for k in range(10):
        acc += c / (n * k + 1) * z ** (n * k + 1)
        c *= (2.0 / n + k) / (k + 1)
    return w * acc.real, w * acc.imag


def _sc_integral(z, n):
    """The defining integral along the segment 0 -> z (Gauss-Legendre)."""
    t, wt = np.polynomial.legendre.leggauss(64)
    s = 0.5 * (t + 1.0)
    return complex(np.sum(0.5 * wt * z * (1.0 - (s * z) ** n) ** (-2.0 / n)))


def doyle_ref(x, y, w, dp, dq):
    pp = max(abs(dp), 1.0)
    u, v = 0.5 * math.log(x * x + y * y + 1e-20), math.atan2(y, x)
    pitch = dq / pp
    us = u + pitch * v
    rip = 0.12 * (math.sin(pp * v) + math.cos(pp * v - 2 * PI * us))
    vs = v + pitch * rip
    e = math.exp(min(max(us + rip, -20.0), 20.0))
    return w * e * math.cos(vs), w * e * math.sin(vs)


def test_conformal_maps(orc):
    # src/issue154-sc-doyle.gpu.test.ts:87-150
    grid = [(0.0, 0.0), (0.3, 0.4), (-0.6, 0.2), (0.5, -0.5), (0.99, 0.0), (1.5, 1.5),
            (8.0, -8.0), (0.001, 0.001)]
    for n in (5.0, 3.0, 8.0):
        check_pin(orc, "schwarz_christoffel", sc_series_ref, 1.0, (n,), grid + P8, 5, f64_rel=1e-7)
        # well inside the disk ten terms are the integral itself
        for x, y in [(0.3, 0.4), (-0.6, 0.2), (0.5, -0.5), (0.1, 0.05)]:
            g = run(orc, "schwarz_christoffel", x, y, 1.0, (n,))
            want = _sc_integral(complex(x, y), int(n))
            assert abs(complex(g[0], g[1]) - want) < 1e-6, (n, x, y)
    for x, y in grid:
        o = r32(orc, "schwarz_christoffel", x, y, (5.0,))
        assert np.isfinite(o[:2]).all() and math.hypot(o[0], o[1]) < 4
    o = r32(orc, "schwarz_christoffel", 0.01, 0.02, (5.0,))
    assert abs(o[0] - 0.01) < 5e-5 and abs(o[1] - 0.02) < 5e-5           # w(z) ~ z
    dgrid = [(0.5, 0.7), (-1.2, 0.4), (0.3, -2.1), (3.0, 3.0), (-3.0, -3.0), (8.0, -8.0),
             (0.05, 0.05), (1.0, 0.0)]
    check_pin(orc, "doyle", doyle_ref, 1.0, (2.0, 1.0), dgrid + P8, 3, f64_rel=1e-7)
    check_pin(orc, "doyle", doyle_ref, 1.0, (3.7, -1.4), dgrid + P8, 3, f64_rel=1e-7)
    assert abs(r32(orc, "doyle", 1.0, 0.0, (2.0, 0.0))[1]) < 5e-5        # pitch 0 on the axis
    a = r32(orc, "doyle", 0.6, 0.8, (2.0, 1.0))
    b = r32(orc, "doyle", 0.6, 0.8, (2.001, 1.0))
    assert math.hypot(a[0] - b[0], a[1] - b[1]) < 0.05                   # continuous in p


# ------------------------------------------------- knots (V302, V303) --
def torus_knot_ref(x, y, w, pp, qq, radius, tube):
    phi = math.atan2(y, x)
    pf, qf = max(round(pp), 1.0), max(round(qq), 1.0)
    rr = radius + tube * min(max(math.hypot(x, y), 0.0), 1.0) * math.cos(qf * phi)
    return w * rr * math.cos(pf * phi), w * rr * math.sin(pf * phi)


def braid_ref(x, y, w, strands, twist, crossings):
    n, cr = max(round(strands), 2.0), max(round(crossings), 1.0)
    r, th = math.hypot(x, y), math.atan2(y, x)
    a = (th + PI) / (2 * PI)
    lane = math.floor(a * n)
    fr = a * n - lane
    sgn = -1.0 if int(lane) & 1 else 1.0
    nt = th + twist * sgn * math.sin(PI * fr) * math.sin(cr * PI * fr) * min(r, 1.0)
    return w * r * math.cos(nt), w * r * math.sin(nt)


def test_knots(orc):
    # src/issue155-knots.gpu.test.ts: torus_knot bounded by radius + tube, on
    # the radius circle at tube 0, p-fold symmetric; braid_warp keeps |p|
    pts = [(x * 0.5, float(y)) for x in range(-4, 5) for y in range(-2, 3)]
    check_pin(orc, "torus_knot", torus_knot_ref, 1.0, (3.0, 2.0, 0.6, 0.3), pts + P8, 4, f64_rel=1e-7)
    assert abs(math.hypot(*r32(orc, "torus_knot", 0.0, 0.0, (3.0, 2.0, 0.6, 0.3))[:2]) - 0.6) < 5e-4
    for x, y in pts:
        assert math.hypot(*r32(orc, "torus_knot", x, y, (3.0, 2.0, 0.6, 0.3))[:2]) <= 0.95
    for x, y in [(0.5, 0.3), (-0.7, 0.4), (0.9, -0.5)]:
        assert abs(math.hypot(*r32(orc, "torus_knot", x, y, (3.0, 2.0, 0.6, 0.0))[:2]) - 0.6) < 5e-4
    pa = (0.8 * math.cos(0.4), 0.8 * math.sin(0.4))
    pb = (0.8 * math.cos(0.4 + 2 * PI / 3), 0.8 * math.sin(0.4 + 2 * PI / 3))
    a, b = r32(orc, "torus_knot", *pa, (3.0, 2.0, 0.6, 0.0)), r32(orc, "torus_knot", *pb, (3.0, 2.0, 0.6, 0.0))
    assert abs(a[0] - b[0]) < 5e-4 and abs(a[1] - b[1]) < 5e-4
    ring = [(r * math.cos(0.4 * i), r * math.sin(0.4 * i)) for i in range(16) for r in (0.3, 0.8, 1.5)]
    check_pin(orc, "braid_warp", braid_ref, 1.0, (3.0, 0.6, 2.0), ring + P8, 4, f64_rel=1e-7)
    check_pin(orc, "braid_warp", braid_ref, 1.0, (2.0, 2.0, 8.0), ring + P8, 4, f64_rel=1e-7)
    for x, y in ring:
        o = r32(orc, "braid_warp", x, y, (3.0, 0.6, 2.0))
        assert abs(math.hypot(o[0], o[1]) - math.hypot(x, y)) < 5e-5
    th = 2 * PI / 3 - PI                                                 # a lane edge
    o = r32(orc, "braid_warp", 0.7 * math.cos(th), 0.7 * math.sin(th), (3.0, 0.6, 2.0))
    assert abs(o[0] - 0.7 * math.cos(th)) < 5e-4 and abs(o[1] - 0.7 * math.sin(th)) < 5e-4
    o = r32(orc, "braid_warp", 0.0, 0.0, (3.0, 0.6, 2.0))
    assert o[0] == 0.0 and o[1] == 0.0


# -------------------------------------------------- airy_caustic (V304) --
def airy_ai_ref(x):
    """Ai(x): the Maclaurin series c1 f(x) - c2 g(x) with twelve terms on
    [-5, 4] (A&S 10.4.2), the leading asymptotic terms outside (A&S 10.4.59
    and 10.4.60); src/issue216-airy-caustic.gpu.test.ts:90-97."""
    if x > 4.0:
        xi = 2.0 / 3.0 * x ** 1.5
        return math.exp(-xi) / (2.0 * math.sqrt(PI) * x ** 0.25)
    if x < -5.0:
        xi = 2.0 / 3.0 * (-x) ** 1.5
        return math.sin(xi + PI / 4) / (math.sqrt(PI) * (-x) ** 0.25)
    x3 = x ** 3
    f = tf = 1.0
    g = tg = x
    for k in range(1, 12):
        tf *= x3 / ((3 * k - 1) * (3 * k))
        tg *= x3 / ((3 * k) * (3 * k + 1))
        f += tf
        g += tg
    return 0.355028053887817 * f - 0.258819403792807 * g


def airy_caustic_ref(x, y, w, scale, r0, amp):   # :98-103
    env = min(max(1.0 + amp * airy_ai_ref(scale * (math.hypot(x, y) - r0)), 0.05), 3.0)
    return w * env * x, w * env * y


def test_airy_caustic(orc):
    from scipy.special import airy
    # the series and the asymptotic forms against Ai itself
    # (twelve terms truncate to 1e-11 on [-3, 3], 3.4e-4 at x = -5)
    for x in np.linspace(-3.0, 3.0, 25):
        assert abs(airy_ai_ref(x) - airy(x)[0]) < 1e-10, x
    for x in np.linspace(-5.0, 4.0, 37):
        assert abs(airy_ai_ref(x) - airy(x)[0]) < 4e-4, x
    for x in (4.5, 6.0, 9.0):
        assert abs(airy_ai_ref(x) / airy(x)[0] - 1.0) < 0.02, x   # 5/(72 xi) relative
    for x in (-5.5, -8.0, -12.0):
        assert abs(airy_ai_ref(x) - airy(x)[0]) < 0.01, x
    # src/issue216-airy-caustic.gpu.test.ts:106-143
    pts = [(0.3, 0.0), (0.0, 0.7), (0.9, 0.4), (1.1, 0.0), (1.5, -0.6)]
    check_pin(orc, "airy_caustic", airy_caustic_ref, 1.0, (4.5, 1.1, 2.2), pts + P8, 4, f64_rel=1e-7)
    check_pin(orc, "airy_caustic", airy_caustic_ref, 1.0, (2.0, 0.4, -1.3), P8, 4, f64_rel=1e-7)
    for x in (0.05, 0.5, 1.0, 2.0, 5.0):
        m = math.hypot(*r32(orc, "airy_caustic", x, 0.0, (4.5, 1.1, 2.2))[:2])
        assert 0.05 * x - 1e-4 <= m <= 3.0 * x + 1e-4
    assert abs(r32(orc, "airy_caustic", 6.0, 0.0, (4.5, 1.1, 2.2))[0] - 6.0) < 5e-3


# ---------------------------------------------- copulas (V315, V316) --
def winitzki_erfinv(x):
    """Winitzki's closed-form erfinv with a = 0.147, argument clamped to
    +-0.999999 (src/issue218-inverse-cdf.gpu.test.ts:81-87)."""
    xc = min(max(x, -0.999999), 0.999999)
    a = 0.147
    ln1 = math.log(1.0 - xc * xc)
    t1 = 2.0 / (PI * a) + 0.5 * ln1
    inner = math.sqrt(max(t1 * t1 - ln1 / a, 0.0)) - t1
    return math.copysign(math.sqrt(max(inner, 0.0)), xc) if xc != 0 else 0.0


def _sig_clamped(t):
    return min(max(1.0 / (1.0 + math.exp(-t)), 1e-4), 1.0 - 1e-4)


def copula_gaussian_ref(x, y, w, strength, rho_in):
    s = max(strength, 1e-3)
    zx = math.sqrt(2.0) * winitzki_erfinv(2.0 * _sig_clamped(s * x) - 1.0)
    zy = math.sqrt(2.0) * winitzki_erfinv(2.0 * _sig_clamped(s * y) - 1.0)
    rho = min(max(rho_in, -0.95), 0.95)
    return w * x, w * (rho * zx + math.sqrt(max(1.0 - rho * rho, 0.0)) * zy) / s


def _clayton_c(u, v, th):
    return (u ** -th + v ** -th - 1.0) ** (-1.0 / th)


def copula_clayton_ref(x, y, w, strength, theta_in):
    s = max(strength, 1e-3)
    u, v = _sig_clamped(s * x), _sig_clamped(s * y)
    th = min(max(theta_in, 0.05), 8.0)
    base = max(u ** -th + v ** -th - 1.0, 1e-6)
    h = min(max(u ** (-th - 1.0) * base ** (-1.0 - 1.0 / th), 1e-4), 1.0 - 1e-4)
    return w * x, w * min(max(math.log(h / (1.0 - h)), -12.0), 12.0) / s


def test_copulas(orc):
    from scipy.special import erfinv
    # Winitzki's approximation stays within its published 2e-3 relative error
    # off the clamp, and is odd
    for x in np.linspace(-0.99, 0.99, 45):
        if x != 0:
            assert abs(winitzki_erfinv(x) / erfinv(x) - 1.0) < 2e-3, x
    assert winitzki_erfinv(0.0) == 0.0
    assert abs(winitzki_erfinv(0.8427008) - 1.0) < 5e-3
    # the Clayton conditional h(v|u) is dC/du of the Clayton copula C(u, v)
    for u, v, th in [(0.3, 0.6, 2.0), (0.8, 0.2, 0.7), (0.5, 0.5, 5.0)]:
        d = 1e-6
        dc = (_clayton_c(u + d, v, th) - _clayton_c(u - d, v, th)) / (2 * d)
        h = u ** (-th - 1.0) * (u ** -th + v ** -th - 1.0) ** (-1.0 - 1.0 / th)
        assert abs(dc - h) < 1e-6, (u, v, th)
    # src/issue217-copula.gpu.test.ts:88-140
    grid = [(0.0, 0.0), (0.5, 0.7), (-1.2, 0.4), (0.3, -2.1), (3.0, 3.0), (-3.0, -3.0),
            (8.0, -8.0), (0.001, 0.001)]
    for prm in ((1.0, 0.0), (1.0, 0.6), (2.3, -0.8)):
        check_pin(orc, "copula_gaussian", copula_gaussian_ref, 1.0, prm, grid + P8, 3, f64_rel=1e-7)
    for x, y in grid:
        o = r32(orc, "copula_gaussian", x, y, (1.0, 0.6))
        assert abs(o[0] - f32(x)) < 5e-6 and abs(o[1]) < 8
    o = r32(orc, "copula_gaussian", 2.0, -1.0, (1.0, 0.0))
    want = math.sqrt(2.0) * erfinv(2.0 * _sig_clamped(-1.0) - 1.0)
    assert abs(o[1] - want) < 5e-3                       # decoupled at rho 0
    for prm in ((1.0, 2.0), (1.0, 50.0), (0.6, 0.3)):
        check_pin(orc, "copula_clayton", copula_clayton_ref, 1.0, prm, grid + P8, 3, f64_rel=1e-6)
        for x, y in grid:
            o = r32(orc, "copula_clayton", x, y, prm)
            assert np.isfinite(o[:2]).all() and abs(o[1]) <= 12.001 / prm[0]


# ------------------------------- gaussian_cdf, levy_cdf (V307, V308) --
def gaussian_cdf_ref(x, y, w, mu, sigma):
    u = _umap(x * x + y * y)
    return _radial(x, y, w, max(mu + sigma * math.sqrt(2.0) * winitzki_erfinv(2.0 * u - 1.0), 0.0))


def levy_cdf_ref(x, y, w, c, tail):
    u = _umap(x * x + y * y)
    e = winitzki_erfinv(1.0 - u)
    return _radial(x, y, w, min(max(c, 1e-4) / max(2.0 * e * e, 1e-6), max(tail, 0.01)))


def test_normal_and_levy_cdfs(orc):
    # src/issue218-inverse-cdf.gpu.test.ts:88-118
    check_pin(orc, "gaussian_cdf", gaussian_cdf_ref, 1.0, (0.9, 0.5),
              [(1.0, 0.0), (0.0, 0.0), (0.5, 0.0)] + P8, 3, f64_rel=1e-6)
    assert abs(math.hypot(*r32(orc, "gaussian_cdf", 1.0, 0.0, (0.9, 0.5))[:2]) - 0.9) < 5e-4
    assert math.hypot(*r32(orc, "gaussian_cdf", 0.0, 0.0, (0.9, 0.5))[:2]) < 1e-6
    check_pin(orc, "levy_cdf", levy_cdf_ref, 1.0, (0.35, 3.0), [(0.0, 0.0), (40.0, 0.0)] + P8, 4,
              f64_rel=1e-6)
    assert math.hypot(*r32(orc, "levy_cdf", 0.0, 0.0, (0.35, 3.0))[:2]) == 0.0
    assert abs(math.hypot(*r32(orc, "levy_cdf", 40.0, 0.0, (0.35, 3.0))[:2]) - 3.0) < 5e-5
    # the Levy quantile c / (2 erfinv(1 - u)^2) with the exact erfinv, where
    # Winitzki's relative error 2e-3 bounds the gap
    from scipy.special import erfinv
    for x, y in P8[:5]:
        u = _umap(x * x + y * y)
        exact = 0.35 / (2.0 * erfinv(1.0 - u) ** 2)
        if exact < 3.0:
            g = math.hypot(*run(orc, "levy_cdf", x, y, 1.0, (0.35, 3.0))[:2])
            assert abs(g / exact - 1.0) < 5e-3, (x, y)


# ----------------------------------------------------- lichtenberg (V314) --
def _lich_vnoise(px, py, period):
    fx, fy = math.floor(px), math.floor(py)
    ix, iy = int(fx), int(fy)
    tx, ty = px - fx, py - fy
    ux, uy = tx * tx * (3 - 2 * tx), ty * ty * (3 - 2 * ty)
    ix0, ix1 = ix % period, (ix + 1) % period
    c00, c10 = _corner(ix0, iy), _corner(ix1, iy)
    c01, c11 = _corner(ix0, iy + 1), _corner(ix1, iy + 1)
    bot, top = c00 + (c10 - c00) * ux, c01 + (c11 - c01) * ux
    return bot + (top - bot) * uy


def _lich_field(x, y, freq, branches, radial, octaves):
    r = math.hypot(x, y)
    up, vp = (math.atan2(y, x) / (2 * PI) + 0.5) * branches, r * freq
    ui, vi = x * freq, y * freq
    u, v = ui + (up - ui) * radial, vi + (vp - vi) * radial
    period = max(round(branches), 1)
    s = tot = 0.0
    amp = 1.0
    for _ in range(octaves):
        s += amp * _lich_vnoise(u, v, period)
        tot += amp
        amp *= 0.5
        u, v, period = 2 * u, 2 * v, 2 * period
    return s / tot


def lichtenberg_ref(x, y, w, freq_in, br_in, rad_in, det_in, str_in):
    """One Newton step toward the zero set of an angularly periodic value-noise
    field, central differences h = 1e-2, step capped at 2
    (src/issue219-lichtenberg.gpu.test.ts:100-160)."""
    freq, br = max(freq_in, 1e-3), min(max(round(br_in), 1), 16)
    rad, octs, st = min(max(rad_in, 0.0), 1.0), min(max(round(det_in), 1), 4), max(str_in, 0.0)
    if math.hypot(x, y) < 1e-4:
        return w * x, w * y
    h = 1e-2
    f = lambda a, b: _lich_field(a, b, freq, br, rad, octs)
    f0 = f(x, y)
    gx, gy = (f(x + h, y) - f(x - h, y)) / (2 * h), (f(x, y + h) - f(x, y - h)) / (2 * h)
    g2 = max(gx * gx + gy * gy, 1e-3)
    sx, sy = -f0 / g2 * gx, -f0 / g2 * gy
    sl = math.hypot(sx, sy)
    if sl > 2.0:
        sx, sy = sx * 2.0 / sl, sy * 2.0 / sl
    return w * (x + st * sx), w * (y + st * sy)


def test_lichtenberg(orc):
    prm = (1.5, 5.0, 0.8, 3.0, 0.5)
    pts = [(0.5, 0.3), (-0.7, 0.4), (-1.2, -0.9), (0.9, -0.5)]
    # (the oracle carries the field's constants in f32, which moves the
    # Newton step by a few 1e-7)
    check_pin(orc, "lichtenberg", lichtenberg_ref, 1.0, prm, pts, 3, f64_rel=1e-5)
    check_pin(orc, "lichtenberg", lichtenberg_ref, 1.0, prm, P8, 3, f64_rel=1e-5, f32=False)
    check_pin(orc, "lichtenberg", lichtenberg_ref, 1.0, (0.7, 11.0, 0.2, 4.0, 1.3), P8, 3,
              f64_rel=1e-5, f32=False)
    for x, y in pts[:2]:
        o = r32(orc, "lichtenberg", x, y, (1.5, 5.0, 0.8, 3.0, 0.0))
        assert abs(o[0] - x) < 5e-6 and abs(o[1] - y) < 5e-6
    for k in range(16):
        x, y = 1.3 * math.cos(0.4 * k), 1.3 * math.sin(0.4 * k)
        o = r32(orc, "lichtenberg", x, y, prm)
        assert np.isfinite(o[:2]).all() and math.hypot(o[0] - x, o[1] - y) <= 0.5 * 2.0 + 1e-3
    o = r32(orc, "lichtenberg", 0.0, 0.0, prm)
    assert o[0] == 0.0 and o[1] == 0.0


# --------------------------------------- elliptic_E, elliptic_K (V305, V306) --
def ellip_e_ref(m):   # A&S 17.3.36 with m1 = 1 - m
    m1 = min(max(1.0 - m, 0.0), 1.0)
    lg = math.log(1.0 / max(m1, 1e-6))
    return 1.0 + m1 * (0.4630151 + m1 * 0.1077812) + m1 * (0.2452727 + m1 * 0.0412496) * lg


def ellip_k_ref(m):   # A&S 17.3.34, m1 floored at 1e-3
    m1 = min(max(1.0 - m, 1e-3), 1.0)
    lg = math.log(1.0 / m1)
    return 1.3862944 + m1 * (0.1119723 + m1 * 0.0725296) + (0.5 + m1 * (0.1213478 + m1 * 0.0288729)) * lg


def elliptic_e_var_ref(x, y, w, scale):
    r2 = x * x + y * y
    return _radial(x, y, w, scale * ellip_e_ref(r2 / (1.0 + r2)))


def elliptic_k_var_ref(x, y, w, scale, tail):
    r2 = x * x + y * y
    return _radial(x, y, w, min(scale * ellip_k_ref(r2 / (1.0 + r2)), max(tail, 0.01)))


def test_elliptic(orc):
    from scipy.special import ellipe, ellipk
    # the Hastings fits keep their published error bounds (4e-5 for E,
    # 3e-5 for K) against the complete elliptic integrals of parameter m
    for m in np.linspace(0.0, 0.999, 60):
        assert abs(ellip_e_ref(m) - ellipe(m)) < 4e-5, m
        assert abs(ellip_k_ref(m) - ellipk(m)) < 3e-5, m
    # src/issue220-elliptic.gpu.test.ts:84-120
    pts = [(0.0, 0.0), (0.5, 0.0), (0.0, 0.7), (2.0, 0.0)]
    check_pin(orc, "elliptic_E", elliptic_e_var_ref, 1.0, (0.8,), pts + P8, 4, f64_rel=1e-7)
    o = [r32(orc, "elliptic_E", x, y, (0.8,)) for x, y in pts]
    assert math.hypot(*o[0][:2]) < 1e-6 and abs(o[1][1]) < 1e-5 and abs(o[2][0]) < 1e-5
    assert math.hypot(*o[3][:2]) <= 0.8 * PI / 2 + 1e-3
    check_pin(orc, "elliptic_K", elliptic_k_var_ref, 1.0, (0.5, 1.5), [(0.5, 0.0), (40.0, 0.0)] + P8,
              4, f64_rel=1e-7)
    assert abs(math.hypot(*r32(orc, "elliptic_K", 40.0, 0.0, (0.5, 1.5))[:2]) - 1.5) < 5e-5


# ----------------------------------------------------------- peano (V309) --
def peano_ref(x, y, w, extent, trits_f):
    """Per-axis base-3 fixed-point code, Peano reflected-ternary digit flip,
    decode to the cell centre (src/issue221-peano.gpu.test.ts:78-86)."""
    t = int(min(max(trits_f, 2.0), 15.0))
    cells, s = 3 ** t, max(extent, 1e-4)

    def enc(c):
        nrm = (c + s) / (2 * s)
        return min(int((nrm - math.floor(nrm)) * cells), cells - 1)

    def scr(i):
        flip, out = False, 0
        for k in range(t - 1, -1, -1):
            d = (i // 3 ** k) % 3
            e = 2 - d if flip else d
            out += e * 3 ** k
            flip ^= bool(e & 1)
        return out

    def dec(i):
        return (i + 0.5) / cells * 2 * s - s
    return w * dec(scr(enc(x))), w * dec(scr(enc(y)))


def test_peano(orc):
    # the digit flip permutes the cells (a bijection on 3^t indices) and, at
    # two trits, is the serpentine order of the Peano curve's first level
    prm = lambda ext, t: (ext, float(t))
    for t in (2, 3, 4):
        cells = 3 ** t
        xs = [((k + 0.5) / cells) * 2 - 1 for k in range(cells)]
        outs = sorted(round((peano_ref(x, 0.0, 1.0, 1.0, t)[0] + 1) / 2 * cells - 0.5) for x in xs)
        assert outs == list(range(cells))
    xs = [((k + 0.5) / 9) * 2 - 1 for k in range(9)]
    order = [round((peano_ref(x, 0.0, 1.0, 1.0, 2)[0] + 1) / 2 * 9 - 0.5) for x in xs]
    assert order == [0, 1, 2, 5, 4, 3, 6, 7, 8]
    # src/issue221-peano.gpu.test.ts:88-125
    pts = [(-0.8, 0.3), (0.0, 0.0), (0.55, -0.91), (0.99, 0.99), (-0.4, -0.6)]
    check_pin(orc, "peano", peano_ref, 1.0, prm(1.0, 5), pts + P8, 5, f64_rel=1e-7)
    check_pin(orc, "peano", peano_ref, 1.0, prm(1.0, 2), [(x, 0.0) for x in xs], 5)
    ext = 1.35
    for x, y in [(5.0, -7.0), (0.01, 0.01), (-3.3, 2.2), (100.0, 100.0)]:
        o = r32(orc, "peano", x, y, prm(ext, 6))
        assert abs(o[0]) <= ext and abs(o[1]) <= ext
    check_pin(orc, "peano", peano_ref, 1.0, prm(ext, 6), P8, 5, f64_rel=1e-7)


# ------------------------------------------ parallel weight invariant (V213) --
def test_parallel_weight_invariant(orc):
    # src/issue233-parallel-weight.gpu.test.ts:27-97: with zero widths both
    # branches reduce to out_x = tilesize * x and out_y = height * y +- w * move,
    # so out_x ignores w and out_y(2) - 2 out_y(1) = -height * y, for any draws
    prm = (0.0, 0.5, 10.0, 1.0, 0.5, 3.0, 0.0, 0.5, 10.0, 1.0)
    for s in range(8):
        words = WORDS[s * 32:(s + 1) * 32]
        for x, y in [(0.7, 0.4), (-0.6, 0.3), (0.9, -0.5), (1.2, 0.2)]:
            a = run(orc, "parallel", f32(x), f32(y), 1.0, prm, np.float32, words)
            b = run(orc, "parallel", f32(x), f32(y), 2.0, prm, np.float32, words)
            assert abs(a[0] - 0.5 * x) < 5e-6 and abs(b[0] - 0.5 * x) < 5e-6
            assert abs(b[1] - 2 * a[1] + 0.5 * y) < 5e-6
            assert min(abs(a[1] - (0.5 * y + 3.0)), abs(a[1] - (0.5 * y - 1.0))) < 5e-6
    assert run(orc, "parallel", 0.7, 0.4, 0.0, prm)[:2] == (0.0, 0.0)   # w = 0 is a no-op


# --------------------------------------- bwraps, crackle smoke (V107, V108) --
def test_bwraps_crackle_smoke(orc):
    # src/issue243-bwraps-crackle.gpu.test.ts:53-110 at its seven points
    pts = [(0.5, 0.3), (-0.5, 0.7), (1.2, -0.4), (-1.0, -1.0), (0.1, 0.05), (0.0, 2.0), (-2.5, 1.7)]
    check_finite(orc, "bwraps", 1.0, (1.0, 0.0, 1.0, -1.04, 0.71), pts)
    check_finite(orc, "crackle", 1.0, (1.0, 0.2, 0.0, 1.0), pts)
    check_finite(orc, "crackle", 1.0, (1.0, 0.2, 1.0, 1.0), pts)
    # bwraps with zero twists and zero space: every point is a bubble-wrap
    # cell centre plus its offset scaled by gain inside the cell's radius;
    # the identity holds outside the bubble
    o = run(orc, "bwraps", -1.0, -1.0, 1.0, (1.0, 0.0, 1.0, 0.0, 0.0))
    assert np.isfinite(o[:2]).all()


# --------------------------------------------------- parameter domains --
def test_parameter_domains(orc):
    # src/issue246-param-domain.gpu.test.ts:77-115
    pts = [(0.5, 0.3), (-0.6, 0.7), (1.1, -0.4), (0.2, 0.9)]
    m2 = [run(orc, "atan", f32(x), f32(y), 1.0, (2.0, 1.0), np.float32) for x, y in pts]
    m5 = [run(orc, "atan", f32(x), f32(y), 1.0, (5.0, 1.0), np.float32) for x, y in pts]
    assert all(o[0] == 0.0 and o[1] == 0.0 for o in m5)          # out-of-range mode
    assert any(abs(v) > 1e-6 for o in m2 for v in o[:2])
    for x, y in pts:
        w2 = run(orc, "shredrad", f32(x), f32(y), 1.0, (4.0, 2.0), np.float32)
        w1 = run(orc, "shredrad", f32(x), f32(y), 1.0, (4.0, 1.0), np.float32)
        assert abs(w2[0] - w1[0]) < 5e-6 and abs(w2[1] - w1[1]) < 5e-6   # width clamped to 1
    assert any(abs(run(orc, "shredrad", f32(x), f32(y), 1.0, (4.0, 2.0), np.float32)[k] -
                   run(orc, "shredrad", f32(x), f32(y), 1.0, (4.0, 0.5), np.float32)[k]) > 1e-4
               for x, y in pts for k in range(2))
    # Y_l^m uses |m| (the associated Legendre domain): m = -2 matches m = +2
    # through hydrogen_orbital, the caller whose m runs over -l..l
    for x, y in pts + P8:
        a = run(orc, "hydrogen_orbital", x, y, 1.0, (3.0, 2.0, -2.0, 0.5, 1.0))
        b = run(orc, "hydrogen_orbital", x, y, 1.0, (3.0, 2.0, 2.0, 0.5, 1.0))
        assert a[:2] == b[:2]
    assert any(abs(v) > 1e-6 for x, y in pts
               for v in run(orc, "hydrogen_orbital", x, y, 1.0, (3.0, 2.0, 2.0, 0.5, 1.0))[:2])


# ------------------------------------------- direct colours (V99 - V102) --
# Ken Perlin's permutation of 0..255 ("Improving Noise", SIGGRAPH 2002, and
# his reference implementation), which src/noise-perlin-oracle.ts:12-29 uses
PERLIN_P = [
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69,
    142, 8, 99, 37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252, 219,
    203, 117, 35, 11, 32, 57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168, 68, 175,
    74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211, 133, 230,
    220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216, 80, 73, 209, 76,
    132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100, 109, 198, 173,
    186, 3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82, 85, 212, 207, 206,
    59, 227, 47, 16, 58, 17, 182, 189, 28, 42, 223, 183, 170, 213, 119, 248, 152, 2, 44, 154, 163,
    70, 221, 153, 101, 155, 167, 43, 172, 9, 129, 22, 39, 253, 19, 98, 108, 110, 79, 113, 224, 232,
    178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238, 210, 144, 12, 191, 179, 162,
    241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31, 181, 199, 106, 157, 184, 84, 204,
    176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205, 93, 222, 114, 67, 29, 24, 72, 243, 141,
    128, 195, 78, 66, 215, 61, 156, 180]


def perlin2d_ref(x, y):
    """Improved noise in 2-D: quintic fade, gradients +-u +-2v chosen by the
    hash's low three bits (src/noise-perlin-oracle.ts:43-70)."""
    perm = PERLIN_P + PERLIN_P

    def grad(h, u, v):
        h &= 7
        a, b = (u, v) if h < 4 else (v, u)
        return (-a if h & 1 else a) + (-2 * b if h & 2 else 2 * b)

    fx, fy = math.floor(x), math.floor(y)
    X, Y = int(fx) & 255, int(fy) & 255
    xf, yf = x - fx, y - fy
    fade = lambda t: t * t * t * (t * (t * 6 - 15) + 10)
    u, v = fade(xf), fade(yf)
    A, B = (perm[X] + Y) & 255, (perm[X + 1] + Y) & 255
    g00, g10 = grad(perm[A], xf, yf), grad(perm[B], xf - 1, yf)
    g01, g11 = grad(perm[A + 1], xf, yf - 1), grad(perm[B + 1], xf - 1, yf - 1)
    lo, hi = g00 + (g10 - g00) * u, g01 + (g11 - g01) * u
    return lo + (hi - lo) * v


def perlin_fbm_ref(x, y, octaves, scale):   # src/noise-perlin-oracle.ts:72-90
    tot = mx = 0.0
    amp, fr = 1.0, scale
    for _ in range(max(1, min(8, int(octaves)))):
        tot += perlin2d_ref(x * fr, y * fr) * amp
        mx += amp
        amp *= 0.5
        fr *= 2
    return tot / mx


def _hue_of(rgb):
    import colorsys
    return colorsys.rgb_to_hls(*rgb)


def test_direct_colours(orc):
    import colorsys
    assert sorted(PERLIN_P) == list(range(256))
    # dc_linear (src/dc-variations.gpu.test.ts:36-41, toBeCloseTo 5)
    for x, y in [(0.0, 0.0), (0.5, 0.5), (-0.5, -0.5), (1.0, 1.0), (-1.0, 1.0), (5.0, 5.0), (-5.0, -5.0)]:
        want = [min(max(v, 0.0), 1.0) for v in (0.5 + 0.5 * x, 0.5 + 0.5 * y, 0.5 - 0.25 * (x + y))]
        got = orc.dc_color(V["dc_linear"], x, y)
        assert all(abs(g - w) < 5e-6 for g, w in zip(got, want)), (x, y)
    # dc_perlin: HSL(fract(0.5 + 0.5 n + seed), 1, 0.55) of the fBm n, read
    # back through the textbook HSL inverse (src/dc-variations.gpu.test.ts)
    samples = [(0.0, 0.0, 1.0, 3, 0.0), (0.5, 0.5, 1.0, 3, 0.0), (1.5, 2.7, 1.0, 3, 0.25),
               (-3.3, 4.8, 2.0, 4, 0.5), (10.0, 10.0, 0.5, 2, 0.0), (5.5, -3.3, 0.5, 4, 0.0),
               (-7.7, 4.2, 3.0, 3, 0.1), (0.31, -0.72, 1.3, 8, 0.7)]
    fps = set()
    for x, y, sc, octv, seed in samples:
        rgb = orc.dc_color(V["dc_perlin"], x, y, (sc, float(octv), seed))
        n = perlin_fbm_ref(f32(x), f32(y), octv, f32(sc))
        hue = (0.5 + 0.5 * n + seed) % 1.0
        want = colorsys.hls_to_rgb(hue, 0.55, 1.0)
        assert all(abs(g - w) < 2e-4 for g, w in zip(rgb, want)), (x, y, rgb, want)
        assert all(0.0 <= v <= 1.0 for v in rgb)
        fps.add(tuple(round(v * 10) for v in rgb))
    assert len(fps) > 1
    r0 = orc.dc_color(V["dc_perlin"], 0.0, 0.0, (1.0, 3.0, 0.0))
    assert r0[0] < r0[1] and r0[0] < r0[2]                         # hue 0.5: cyan
    assert perlin_fbm_ref(0.0, 0.0, 3, 1.0) == 0.0
    # the noise vanishes on the lattice and stays within its bound
    for x, y in [(3.0, -2.0), (17.0, 5.0), (-128.0, 255.0)]:
        assert perlin2d_ref(x, y) == 0.0
    g = np.random.default_rng(5).uniform(-50, 50, (400, 2))
    assert max(abs(perlin2d_ref(a, b)) for a, b in g) <= 1.5
    # dc_gridout: one colour per cell of the n x n lattice, from a hash of
    # the cell (src/dc-variations.gpu.test.ts:203-262)
    c = [orc.dc_color(V["dc_gridout"], x, y, (4.0,)) for x, y in
         [(0.1, 0.1), (0.2, 0.2), (0.3, 0.1), (0.1, 0.3), (0.5, 0.5)]]
    assert c[0] == c[1] and c[2] != c[0] and c[3] != c[0]
    assert all(0.0 <= v <= 1.0 for rgb in c for v in rgb)
    for x, y, n in [(0.1, 0.1, 4.0), (-0.7, 0.35, 4.0), (2.3, -1.9, 7.0), (0.0, 0.0, 1.0)]:
        cx, cy = math.floor(f32(x) * f32(n)), math.floor(f32(y) * f32(n))
        mixed = ((((cx * 73856093) ^ (cy * 19349663)) & 0xFFFFFFFF) + 0x9E3779B9) & 0xFFFFFFFF
        want = (_hash01(mixed), _hash01(mixed ^ 0xDEADBEEF), _hash01(mixed ^ 0x13579BDF))
        got = orc.dc_color(V["dc_gridout"], x, y, (n,))
        assert all(abs(a - b) < 1e-7 for a, b in zip(got, want)), (x, y, got, want)
    # dc_cylinder: HSL(fract(0.5 + 0.5 sin x), 0.9, clamp(0.5 + 0.25 tanh(y / 2), 0.2, 0.85))
    for x, y in P8 + [(1e3, -1e3), (0.0, 40.0)]:
        hue = (0.5 + 0.5 * math.sin(x)) % 1.0 if abs(x) <= 1e6 else None
        lit = min(max(0.5 + 0.25 * math.tanh(y * 0.5), 0.2), 0.85)
        want = colorsys.hls_to_rgb(hue, lit, 0.9)
        got = orc.dc_color(V["dc_cylinder"], x, y)
        assert all(abs(a - b) < 2e-4 for a, b in zip(got, want)), (x, y, got, want)


# ------------------------------------- quadratic and hopping maps (V323 - V325) --
def sprott_poly_ref(x, y, w, a1, a2, a3, a4, a5, b1, b2, b3, b4, b5):
    """Sprott's general 2-D quadratic map without constant terms
    (src/sprott-poly.cpu.test.ts:18-27)."""
    return (w * (a1 * x + a2 * x * x + a3 * x * y + a4 * y + a5 * y * y),
            w * (b1 * x + b2 * x * x + b3 * x * y + b4 * y + b5 * y * y))


def hopalong_ref(x, y, w, a, b, c):
    """Barry Martin's Hopalong map x' = y - sgn(x) sqrt|b x - c|, y' = a - x."""
    s = (x > 0) - (x < 0)
    return w * (y - s * math.sqrt(abs(b * x - c))), w * (a - x)


def gumowski_mira_ref(x, y, w, a, b):
    """The Gumowski-Mira map x' = b y + G(x), y' = -x + G(x'),
    G(t) = a t + 2 (1 - a) t^2 / (1 + t^2)."""
    g = lambda t: a * t + 2.0 * (1.0 - a) * t * t / (1.0 + t * t)
    xp = b * y + g(x)
    return w * xp, w * (-x + g(xp))


def test_quadratic_and_hopping_maps(orc):
    # src/sprott-poly.gpu.test.ts:60-90, src/hopalong.gpu.test.ts:55-70,
    # src/gumowski-mira.gpu.test.ts:55-68 (toBeCloseTo 4)
    for co, pts in [((0.5, -0.3, 0.2, 0.9, -0.4, 0.1, 0.6, -0.7, 0.3, 0.8),
                     [(0.4, -0.2), (-0.7, 0.5), (0.1, 0.9)]),
                    ((0.31, 0.88, -0.42, 0.17, -0.93, 0.55, -0.21, 0.64, 0.39, -0.76),
                     [(0.3, 0.7), (-0.5, -0.1), (0.9, -0.6)])]:
        check_pin(orc, "sprott_poly", sprott_poly_ref, 1.0, co, pts + P8, 4)
        check_pin(orc, "sprott_poly", sprott_poly_ref, 2.5, co, pts, 4)    # weight outside
    pts = [(0.3, -0.7), (-1.2, 0.4), (0.0, 0.9), (-0.05, -0.05)]
    check_pin(orc, "hopalong", hopalong_ref, 1.0, (1.0, 2.0, 0.5), pts + P8, 4)
    check_pin(orc, "hopalong", hopalong_ref, 0.7, (-0.4, 1.3, -2.0), pts + P8, 4)
    o = r32(orc, "hopalong", 0.0, 0.9, (1.0, 2.0, 0.5))
    assert o[0] == f32(0.9) and o[1] == 1.0                             # sgn(0) = 0
    pts = [(0.3, -0.7), (-1.2, 0.4), (0.0, 0.9), (2.5, -1.1)]
    check_pin(orc, "gumowski_mira", gumowski_mira_ref, 1.0, (0.31, 0.99), pts + P8, 4)
    check_pin(orc, "gumowski_mira", gumowski_mira_ref, 1.0, (-0.8, 0.05), pts + P8, 4)
    # G(0) = 0: the origin is a fixed point; a = 1 makes G the identity
    assert run(orc, "gumowski_mira", 0.0, 0.0, 1.0, (0.31, 0.99))[:2] == (0.0, 0.0)
    o = run(orc, "gumowski_mira", 0.3, -0.7, 1.0, (1.0, 0.5))
    assert abs(o[0] - (0.5 * -0.7 + 0.3)) < 1e-12 and abs(o[1] - (-0.3 + o[0])) < 1e-12


# ------------------------------- JWildfire ports, V152 - V219 (WGSL restated) --
# The reference's only tests of these entries are finite-output smokes at
# seven points (src/issue121-Lall.gpu.test.ts:50-98, 197-240 and
# src/issue121-Lall-tail.gpu.test.ts:50-72, 139-160); their definitions are
# the WGSL bodies in src/shaders/chaos.wgsl, restated here independently of
# the oracle's C++ and compared at those points and more.
LALL_PTS = [(0.5, 0.3), (-0.5, 0.7), (1.2, -0.4), (-1.0, -1.0), (0.1, 0.05), (0.0, 2.0),
            (-2.5, 1.7)]


def _sel(f, t, c):   # WGSL select(f, t, cond)
    return t if c else f


def ennepers_ref(x, y, w):
    return w * (x - x ** 3 / 3.0) + x * y * y, w * (y - y ** 3 / 3.0) + y * x * x


def circus_ref(x, y, w, scale):
    s1 = 1e30 if scale == 0 else 1.0 / scale
    r, a = math.hypot(x, y), math.atan2(y, x)
    rs = r * scale if r <= 1.0 else r * s1
    return w * rs * math.cos(a), w * rs * math.sin(a)


def clifford_js_ref(x, y, w, a, b, c, d):
    return (w * (math.sin(a * y) + c * math.cos(a * x)),
            w * (math.sin(b * x) + d * math.cos(b * y)))


def devil_warp_ref(x, y, w, a, b, effect, warp, rmin, rmax):
    r2 = 1.0 / max(x * x + y * y, 1e-30)
    ba, bb = x * x + r2 * b * y * y, y * y + r2 * a * x * x
    pa = ba ** warp if ba > 0 else 0.0
    pb = bb ** warp if bb > 0 else 0.0
    r = effect * min(max(pa - pb, rmin), rmax)
    return x * (1 + r), y * (1 + r)


def cardioid_ref(x, y, w, a):
    th = math.atan2(y, x)
    r = w * math.sqrt(max(x * x + y * y + math.sin(a * th) + 1.0, 0.0))
    return r * math.cos(th), r * math.sin(th)


def bcollide_ref(x, y, w, num_p, a):
    num = max(1.0, num_p)
    bcn_pi, pi_bcn, bca_bcn = num / PI, PI / num, PI * a / num
    xp1, xm1, y2 = x + 1, x - 1, y * y
    tau = 0.5 * (math.log(max(xp1 * xp1 + y2, 1e-30)) - math.log(max(xm1 * xm1 + y2, 1e-30)))
    sr = PI - math.atan2(y, xp1) - math.atan2(y, 1 - x)
    alt = int(sr * bcn_pi)
    off = bca_bcn if alt & 1 == 0 else -bca_bcn
    fold = sr + off - math.floor((sr + off) / pi_bcn) * pi_bcn
    sig = alt * pi_bcn + fold
    t = math.cosh(tau) - math.cos(sig)
    t = 1e-30 if abs(t) < 1e-30 else t
    return w * math.sinh(tau) / t, w * math.sin(sig) / t


def bsplit_ref(x, y, w, sx, sy):
    ax = x + sx
    s = math.sin(ax)
    if abs(s) < 1e-6:
        return 0.0, 0.0
    c = math.cos(ax)
    tn = s / c
    tn = math.copysign(1e6, tn) if abs(c) < 1e-6 else tn
    return w / tn * math.cos(y + sy), w / s * (-y + sy)


def bulge_ref(x, y, w, n):
    r = max(math.hypot(x, y), 1e-30)
    s = r ** n / r
    return w * x * s, w * y * s


def corners_ref(x, y, w, cx, cy, mx, my, xp, yp, xya, log_mode, log_base):
    xs, ys = x * x, y * y
    if log_mode == 0:
        ex, ey = xs ** (xp + xya) * mx, ys ** (yp + xya) * my
    else:
        lb = math.log(max(abs(log_base), 1.000001))
        ex = (math.log(xs * mx + 3.0) / lb) ** (xp + 2.25 + xya) - 1.33
        ey = (math.log(ys * my + 3.0) / lb) ** (yp + 2.25 + xya) - 1.33
    return (w * ex + cx if x > 0 else -w * ex - cx), (w * ey + cy if y > 0 else -w * ey - cy)


def fibonacci2_ref(x, y, w, sc, sc2):
    ff, fl = 0.447213595, 0.481211825
    a, b = y * fl, -(x * PI + y * fl)
    e1, e2 = sc * math.exp(sc2 * (x * fl)), sc * math.exp(sc2 * -(x * fl - y * PI))
    return (w * (e1 * math.cos(a) - e2 * math.cos(b)) * ff,
            w * (e1 * math.sin(a) - e2 * math.sin(b)) * ff)


def hypertile_ref(x, y, w, p_p, q, n):
    pa, qa = 2 * PI / max(3.0, p_p), 2 * PI / max(3.0, q)
    den = math.cos(pa) + math.cos(qa)
    r2 = (1 - math.cos(pa)) / den + 1 if abs(den) > 1e-30 else 1.0
    r = 1 / math.sqrt(max(r2, 1e-30)) if r2 > 0 else 1.0
    re, im = r * math.cos(n * pa), r * math.sin(n * pa)
    a, b = x + re, y - im
    c, d = re * x - im * y + 1, re * y + im * x
    vr = w / max(c * c + d * d, 1e-30)
    return vr * (a * c + b * d), vr * (b * c - a * d)


def idisc_ref(x, y, w):
    a = PI / (math.hypot(x, y) + 1)
    v = math.atan2(y, x) * w / PI
    return v * math.cos(a), v * math.sin(a)


def hole_ref(x, y, w, a, inside_p):
    al = math.atan2(y, x)
    de = max(al / PI + 1, 1e-30) ** a
    s = x * x + y * y
    r = w * de / max(s + de, 1e-30) if int(inside_p) != 0 else w * math.sqrt(max(s + de, 0.0))
    return r * math.cos(al), r * math.sin(al)


def kaleidoscope_ref(x, y, w, pull, rotate, line_up, off_x, off_y):
    c45, s45 = math.cos(45.0), math.sin(45.0)   # 45 radians, as the reference has it
    ox = w * ((rotate * x * c45 - y * s45 + line_up) + off_x)
    oy = (w * ((rotate * y * c45 + x * s45 + pull + line_up) + off_y) if y > 0
          else w * (rotate * y * c45 + x * s45 - pull - line_up))
    return ox, oy


def layered_spiral_ref(x, y, w, radius):
    t = x * x + y * y + 1e-30
    return w * x * radius * math.cos(t), w * x * radius * math.sin(t)


def linear_t_ref(x, y, w, px, py):
    return (w * (1 if x >= 0 else -1) * max(abs(x), 1e-30) ** px,
            w * (1 if y >= 0 else -1) * max(abs(y), 1e-30) ** py)


def ovoid_ref(x, y, w, px, py):
    r = w / (x * x + y * y + 1e-6)
    return x * r * px, y * r * py


def unpolar_ref(x, y, w):
    v = w / PI * 0.5
    return v * math.exp(y) * math.sin(x), v * math.exp(y) * math.cos(x)


def target_ref(x, y, w, even, odd, size):
    a, r = math.atan2(y, x), math.hypot(x, y)
    t = math.log(max(r, 1e-30))
    if t < 0:
        t -= 0.5 * size
    ss = max(abs(size), 1e-30)
    t = abs(t) - math.floor(abs(t) / ss) * ss
    a += even if t < 0.5 * size else odd
    return r * math.cos(a) * w, r * math.sin(a) * w


def funnel_ref(x, y, w, effect):
    cx, cy = math.cos(x), math.cos(y)
    sx = 1 / (1e-30 if abs(cx) < 1e-30 else cx)
    sy = 1 / (1e-30 if abs(cy) < 1e-30 else cy)
    off = effect * PI
    return (w * math.tanh(min(max(x, -20.0), 20.0)) * (sx + off),
            w * math.tanh(min(max(y, -20.0), 20.0)) * (sy + off))


def holesq_ref(x, y, w):
    x, y = w * x, w * y
    fx, fy = abs(x), abs(y)
    if fx + fy > 1:
        return x, y
    if fx > fy:
        return ((x - fy + 1) * 0.5 if x >= 0 else (x + fy - 1) * 0.5), y
    return x, ((y - fx + 1) * 0.5 if y >= 0 else (y + fx - 1) * 0.5)


def fourth_ref(x, y, w, spin, space, twist, off_x, off_y):
    if x > 0 and y > 0:
        th, r = math.atan2(y, x), 1 / max(math.hypot(x, y), 1e-30)
        return w * r * math.cos(th), w * r * math.sin(th)
    if x > 0 and y < 0:
        r2 = x * x + y * y
        if r2 < w * w:
            r = w * math.sqrt(max(w * w / max(r2, 1e-30) - 1, 0.0))
            return r * x, r * y
        return w * x, w * y
    if x < 0 and y > 0:
        xx, yy = x - off_x, y + off_y
        r0 = math.hypot(xx, yy)
        if r0 < w:
            th = math.atan2(yy, xx) + spin + twist * (w - r0)
            return w * r0 * math.cos(th) + off_x, w * r0 * math.sin(th) - off_y
        r = w * (1 + space / max(r0, 1e-30))
        return r * xx + off_x, r * yy - off_y
    return w * x, w * y


def pulse_ref(x, y, w, fx, fy, sx, sy):
    return w * (x + sx * math.sin(x * fx)), w * (y + sy * math.sin(y * fy))


def rays1_ref(x, y, w):
    t = x * x + y * y
    tv = math.tan(math.sqrt(max(t, 1e-30)))
    u = 1 / (1e-30 if abs(tv) < 1e-30 else tv) + w * (2 / PI) ** 2
    return w * u * t / (x if x != 0 else 1e-30), w * u * t / (y if y != 0 else 1e-30)


def rays3_ref(x, y, w):
    t = x * x + y * y
    ts = max(t, 1e-30)
    inner = math.sin(t * t + 1e-6) * math.sin(1 / (ts * ts) + 1e-6)
    u = 1 / max(math.sqrt(max(math.cos(inner), 1e-30)), 1e-30)
    c = w / 10
    return (c * u * math.cos(t) * t / (x if x != 0 else 1e-30),
            c * u * math.tan(t) * t / (y if y != 0 else 1e-30))


def tancos_ref(x, y, w):
    d1 = 1e-6 + x * x + y * y
    d2 = w / d1
    return d2 * math.tanh(min(d1, 20.0)) * 2 * x, d2 * math.cos(d1) * 2 * y


def twoface_ref(x, y, w):
    v = w / max(x * x + y * y, 1e-30) if x > 0 else w
    return v * x, v * y


def _elliptic_xmax(x, y):
    tmp, tmp2 = y * y + x * x + 1, 2 * x
    xmax = (math.sqrt(max(tmp + tmp2, 0.0)) + math.sqrt(max(tmp - tmp2, 0.0))) * 0.5
    return max(xmax, 1.0)


def e_collide_ref(x, y, w, num_p, a):
    num = max(1.0, num_p)
    ecn_pi, pi_ecn, eca_ecn = num / PI, PI / num, PI * a / num
    xmax = _elliptic_xmax(x, y)
    nu = math.acos(min(max(x / xmax, -1.0), 1.0))
    alt = int(nu * ecn_pi)
    arg = nu + (eca_ecn if alt & 1 == 0 else -eca_ecn)
    nu = alt * pi_ecn + (arg - math.floor(arg / pi_ecn) * pi_ecn)
    if y <= 0:
        nu = -nu
    return (w * xmax * math.cos(nu),
            w * math.sqrt(max(xmax - 1, 0.0)) * math.sqrt(xmax + 1) * math.sin(nu))


def e_mod_ref(x, y, w, radius, distance):
    xmax = _elliptic_xmax(x, y)
    mu = math.acosh(xmax)
    nu = math.acos(min(max(x / xmax, -1.0), 1.0))
    if y < 0:
        nu = -nu
    if mu < radius and -mu < radius:
        r = max(radius, 1e-30)
        if nu > 0:
            mu = math.fmod(mu + r + distance * r, 2 * r) - r
        else:
            mu = math.fmod(mu - r - distance * r, 2 * r) + r
    return w * math.cosh(mu) * math.cos(nu), w * math.sinh(mu) * math.sin(nu)


def inv_squircular_ref(x, y, w):
    if w == 0:
        return 0.0, 0.0
    ri = x * x + y * y
    r2 = math.sqrt(max(ri * (w * w * ri - 4 * x * x * y * y) / w, 0.0))
    r = math.sqrt(max(ri - r2, 0.0)) / 1.41421356237
    return r / (x if x != 0 else 1e-30), r / (y if y != 0 else 1e-30)


def hypershift_ref(x, y, w, shift, stretch):
    rad1 = 1 / max(x * x + y * y, 1e-30)
    xx, yy = rad1 * x + shift, rad1 * y
    rad = w * (1 - shift * shift) / max(xx * xx + yy * yy, 1e-30)
    return rad * xx + shift, rad * yy * stretch


def hex_modulus_ref(x, y, w, size):
    s32, s3 = 0.8660254037844386, 1.7320508075688772
    hs = s32 / max(size, 1e-30)
    X, Y = x * hs, y * hs
    cx = 0.5773502691896258 * X - Y / 3
    cz = 2 * Y / 3
    cy = -cx - cz
    rx, ry, rz = round(cx), round(cy), round(cz)
    dx, dy, dz = abs(rx - cx), abs(ry - cy), abs(rz - cz)
    if dx > dy and dx > dz:
        rx = -ry - rz
    elif dy > dz:
        ry = -rx - rz
    else:
        rz = -rx - ry
    fx, fy = s3 * rx + s32 * rz, 1.5 * rz
    return (X - fx) * w / s32, (Y - fy) * w / s32


def b_mod_ref(x, y, w, radius, distance):
    xp1, xm1, y2 = x + 1, x - 1, y * y
    tau = 0.5 * (math.log(max(xp1 * xp1 + y2, 1e-30)) - math.log(max(xm1 * xm1 + y2, 1e-30)))
    sig = PI - math.atan2(y, xp1) - math.atan2(y, 1 - x)
    if tau < radius and -tau < radius:
        r = max(radius, 1e-30)
        arg = tau + r + distance * r
        tau = arg - math.floor(arg / (2 * r)) * 2 * r - r
    t = math.cosh(tau) - math.cos(sig)
    t = 1e-30 if abs(t) < 1e-30 else t
    return w * math.sinh(tau) / t, w * math.sin(sig) / t


def waves3_ref(x, y, w, sx, sy, fx, fy, sxf, syf):
    ssx = 0.5 * sx * (1 + math.sin(y * sxf))
    ssy = 0.5 * sy * (1 + math.sin(x * syf))
    return w * (x + math.sin(y * fx) * ssx), w * (y + math.sin(x * fy) * ssy)


def scry2_ref(x, y, w, sides_f, star, circle):
    sides = min(max(int(sides_f), 1), 50)
    a = 2 * PI / sides
    sina, cosa = math.sin(a), math.cos(a)
    sins, coss = math.sin(-star * PI * 0.5), math.cos(star * PI * 0.5)
    sinc, cosc = math.sin(circle * PI * 0.5), math.cos(circle * PI * 0.5)
    xr, yr = x, y
    r2 = xr * coss + abs(yr) * sins
    cr = math.hypot(xr, yr)
    iters = 0
    for i in range(sides - 1):
        xr, yr = xr * cosa - yr * sina, xr * sina + yr * cosa
        r2 = max(r2, xr * coss + abs(yr) * sins)
        iters = i + 1
    r2 = r2 * cosc + cr * sinc
    r1 = r2
    r2 = r2 * r2 if iters > 1 else abs(r2) * r2
    if w == 0:
        return 0.0, 0.0
    d = r1 * (r2 + 1 / w)
    return (0.0, 0.0) if d == 0 else (x / d, y / d)


def ennepers2_ref(x, y, w, ap, bp, cp):
    r2 = 1 / max(x * x + y * y, 1e-30)
    dxy = (ap * x) ** 2 - (bp * y) ** 2
    return (w * x * (ap * ap - dxy * r2 - cp * math.sqrt(abs(x))),
            w * y * (bp * bp - dxy * r2 - cp * math.sqrt(abs(y))))


def hole2_ref(x, y, w, a, b, c, d, inside_p, shape_p):
    rho = x * x + y * y
    th = math.atan2(y, x) * d
    de = max(th / PI + 1, 1e-30) ** a * c
    s = int(shape_p)
    f = {0: lambda: math.sqrt(max(rho, 0.0)) + de,
         1: lambda: math.sqrt(max(rho + de, 0.0)),
         2: lambda: math.sqrt(max(rho + math.sin(b * th) + de, 0.0)),
         3: lambda: math.sqrt(max(rho + math.sin(th) + de, 0.0)),
         4: lambda: math.sqrt(max(rho + th * th - de + 1, 0.0)),
         5: lambda: math.sqrt(max(rho + abs(math.tan(th)) + de, 0.0)),
         6: lambda: math.sqrt(max(rho * (1 + math.sin(b * th)) + de, 0.0)),
         7: lambda: math.sqrt(max(rho + abs(math.sin(0.5 * b * th)) + de, 0.0)),
         8: lambda: math.sqrt(max(rho + math.sin(PI * math.sin(b * th)) + de, 0.0)),
         9: lambda: math.sqrt(max(rho + (math.sin(b * th) + math.sin(2 * b * th + PI * 0.5)) * 0.5
                                   + de, 0.0))}
    r1 = f.get(s, lambda: 1.0)()
    rf = w / max(r1, 1e-30) if int(inside_p) != 0 else w * r1
    return rf * math.cos(th), rf * math.sin(th)


JWF = [   # (entry, oracle, the Lall row's parameters, more parameters)
    ("ennepers", ennepers_ref, (), None),
    ("circus", circus_ref, (0.92,), (1.7,)),
    ("clifford_js", clifford_js_ref, (-1.4, 1.6, 1.0, 0.7), (2.1, -0.3, -0.6, 1.2)),
    ("devil_warp", devil_warp_ref, (1.0, 1.0, 1.5, 0.07, 0.005, 0.005), (0.7, 1.3, 0.9, 0.5, -0.4, 0.6)),
    ("cardioid", cardioid_ref, (1.0,), (2.7,)),
    ("bcollide", bcollide_ref, (1.0, 1.0), (3.0, 0.35)),
    ("bsplit", bsplit_ref, (0.0, 0.0), (0.3, -0.2)),
    ("bulge", bulge_ref, (1.0,), (1.7,)),
    ("corners", corners_ref, (0.0, 0.0, 1.0, 1.0, 1.0, 1.0, 0.0, 0.0, 2.71828),
     (0.1, -0.2, 0.8, 1.1, 0.75, 1.25, 0.1, 1.0, 2.71828)),
    ("fibonacci2", fibonacci2_ref, (1.0, 1.0), (0.6, 0.4)),
    ("hypertile", hypertile_ref, (3.0, 7.0, 0.0), (4.0, 6.0, 2.0)),
    ("idisc", idisc_ref, (), None),
    ("hole", hole_ref, (1.0, 0.0), (0.6, 1.0)),
    ("kaleidoscope", kaleidoscope_ref, (1.0, 1.0, 1.0, 0.0, 0.0), (0.3, 0.8, 0.2, 0.1, -0.2)),
    ("layered_spiral", layered_spiral_ref, (1.0,), (2.5,)),
    ("linear_t", linear_t_ref, (1.0, 1.0), (1.4, 0.6)),
    ("ovoid", ovoid_ref, (1.0, 1.0), (0.7, 1.3)),
    ("unpolar", unpolar_ref, (), None),
    ("target", target_ref, (1.0, 2.0, 0.0), (0.3, -0.5, 0.8)),
    ("funnel", funnel_ref, (4.0,), (0.6,)),
    ("holesq", holesq_ref, (), None),
    ("fourth", fourth_ref, (1.0, 1.0, 1.0, 1.0, 1.0), (0.4, 0.3, 0.7, 0.2, -0.1)),
    ("pulse", pulse_ref, (1.0, 1.0, 0.1, 0.1), (2.0, 3.0, 0.3, -0.2)),
    ("rays1", rays1_ref, (), None),
    ("rays3", rays3_ref, (), None),
    ("tancos", tancos_ref, (), None),
    ("twoface", twoface_ref, (), None),
    ("e_collide", e_collide_ref, (1.0, 0.0), (3.0, 0.4)),
    ("e_mod", e_mod_ref, (1.0, 0.0), (0.8, 0.3)),
    ("inv_squircular", inv_squircular_ref, (), None),
    ("hypershift", hypershift_ref, (0.5, 1.0), (-0.3, 1.4)),
    ("hex_modulus", hex_modulus_ref, (0.5,), (0.8,)),
    ("b_mod", b_mod_ref, (1.0, 0.0), (0.6, 0.4)),
    ("waves3", waves3_ref, (0.05, 0.05, 7.0, 13.0, 0.0, 2.0), (0.2, 0.1, 3.0, 5.0, 1.0, 0.5)),
    ("scry2", scry2_ref, (4.0, 0.15, 0.25), (6.0, 0.4, 0.6)),
    ("ennepers2", ennepers2_ref, (1.0, 0.3333, 0.075), (0.8, 0.5, 0.2)),
    ("hole2", hole2_ref, (1.0, 0.0, 0.0, 0.0, 0.5, 0.5), None),
]


@pytest.mark.parametrize("case", JWF, ids=lambda c: c[0])
def test_jwildfire_ports(orc, case):
    name, ref, p0, p1 = case
    # the reference's smoke row: finite at its seven points (any draws)
    check_finite(orc, name, 1.0, p0, LALL_PTS)
    pts = LALL_PTS + P8 + [(0.3, -0.8), (-0.2, -0.45), (1.7, 0.9)]
    for prm in [p0] + ([p1] if p1 is not None else []):
        for w in (1.0, 0.7):
            check_pin(orc, name, ref, w, prm, pts, 3, f64_rel=1e-6, f32=False)


def test_hole2_shapes(orc):
    # every shape of hole2 (src/shaders/chaos.wgsl, var_hole2's switch)
    pts = LALL_PTS + P8
    for shape in range(11):
        for inside in (0.0, 1.0):
            check_pin(orc, "hole2", hole2_ref, 1.0, (0.8, 1.3, 0.6, 1.2, inside, float(shape)),
                      pts, 3, f64_rel=1e-6, f32=False)


# ------------------------- JWildfire ports with draws, V153 - V219 (WGSL restated) --
class Draws:
    """The draws a variation takes, in order: rand01 of each word as the
    kernels form it (low 28 bits times 1/(2^28 - 1), in f32)."""

    def __init__(self, words):
        self.words, self.k = list(words), 0

    def r01(self):
        w = self.words[self.k] if self.k < len(self.words) else 0
        self.k += 1
        return float(np.float32(np.float32(w & 0x0FFFFFFF) * np.float32(1.0 / 268435455.0)))


def erf_ref(x, y, w, D):   # A&S 7.1.26, as src/shaders/chaos.wgsl's erf_approx has it
    def e(v):
        t = 1 / (1 + 0.3275911 * abs(v))
        poly = ((((1.061405429 * t - 1.453152027) * t + 1.421413741) * t - 0.284496736) * t
                + 0.254829592) * t
        return math.copysign(1.0, v) * (1 - poly * math.exp(-v * v)) if v != 0 else 1 - poly
    return w * e(x), w * e(y)


def asteria_ref(x, y, w, alpha, D):
    sa, ca = math.sin(PI * alpha), math.cos(PI * alpha)
    x0, y0 = w * x, w * y
    r = x0 * x0 + y0 * y0
    r2 = math.sqrt((abs(y0) - 1) ** 2 + (abs(x0) - 1) ** 2)
    inside = (D.r01() > 0.35) if (r < 1 and r2 < 1) else not (r < 1)
    if inside:
        return x0, y0
    rx, ry = x0 * ca - y0 * sa, x0 * sa + y0 * ca
    nx = rx / math.sqrt(max(1 - ry * ry, 1e-30)) * (1 - math.sqrt(max(1 - (1 - abs(ry)) ** 2, 0.0)))
    return nx * ca + ry * sa, -nx * sa + ry * ca


def chrysanthemum_ref(x, y, w, D):
    u = 21 * PI * D.r01()
    p4, p8 = math.sin(17 * u / 3), math.sin(2 * math.cos(3 * u) - 28 * u)
    r = w * 0.1 * (5 * (1 + math.sin(11 * u / 5)) - 4 * p4 ** 4 * p8 ** 8)
    return r * math.cos(u), r * math.sin(u)


def checks_ref(x, y, w, cx, cy, size, rnd, D):
    cs = 1 / (size + 1e-6)
    is_xy = int(round(x * cs)) + int(round(y * cs))
    rnx, rny = rnd * D.r01(), rnd * D.r01()
    dx, dy = (-cx + rnx, -cy) if is_xy & 1 == 0 else (cx, cy + rny)
    return w * (x + dx), w * (y + dy)


def circleblur_ref(x, y, w, D):
    rad = math.sqrt(D.r01())
    a = D.r01() * 2 * PI
    return w * math.cos(a) * rad, w * math.sin(a) * rad


def hypertile1_ref(x, y, w, p_p, q, D):
    pa = 2 * PI / max(3.0, p_p)
    den = math.cos(pa) + math.cos(2 * PI / max(3.0, q))
    r2 = 1 - (math.cos(pa) - 1) / den if abs(den) > 1e-30 else 1.0
    r = 1 / math.sqrt(max(r2, 1e-30)) if r2 > 0 else 1.0
    rpa = math.floor(D.r01() * 10) * pa
    re, im = r * math.cos(rpa), r * math.sin(rpa)
    a, b = x + re, y - im
    c, d = re * x - im * y + 1, re * y + im * x
    vr = w / max(c * c + d * d, 1e-30)
    return vr * (a * c + b * d), vr * (b * c - a * d)


def line_ref(x, y, w, delta, phi, D):
    ux, uy = math.cos(delta * PI) * math.cos(phi * PI), math.sin(delta * PI) * math.cos(phi * PI)
    r = math.sqrt(max(ux * ux + uy * uy + math.sin(phi * PI) ** 2, 1e-30))
    rnd = D.r01() * w
    return ux / r * rnd, uy / r * rnd


def vogel_ref(x, y, w, n, scale, D):
    phi = 1.61803398874989
    i = math.floor(D.r01() * max(1.0, n)) + 1
    a = i * 2 * PI / (phi * phi)
    r = w * (math.hypot(x, y) + math.sqrt(i))
    return r * (math.cos(a) + scale * x), r * (math.sin(a) + scale * y)


def yin_yang_ref(x, y, w, radius, ang1, ang2, dual_t, outside, D):
    sa, ca = math.sin(PI * ang1), math.cos(PI * ang1)
    sb, cb = math.sin(PI * ang2), math.cos(PI * ang2)
    inv, RR = 1.0, radius
    if x * x + y * y < 1:
        nx, ny = x * ca - y * sa, x * sa + y * ca
        if int(dual_t) == 1 and D.r01() > 0.5:
            inv, RR = -1.0, 1 - radius
            nx, ny = x * cb - y * sb, x * sb + y * cb
        if ny > 0:
            t = math.sqrt(max(1 - ny * ny, 0.0))
            k = nx / max(abs(t), 1e-30) * (1.0 if t > 0 else 0.0)
            t1 = (t - 0.5) * 2
            alfa = (1 - k) * 0.5
            k1 = alfa * RR + (1 - alfa)
            return w * (t1 * k1 + alfa * (RR - 1)) * inv, w * math.sqrt(max(1 - t1 * t1, 0.0)) * k1 * inv
        return w * (nx * (1 - RR) + RR) * inv, w * ny * (1 - RR) * inv
    return (w * x, w * y) if int(outside) == 1 else (0.0, 0.0)


def squish_ref(x, y, w, power_p, D):
    pw = max(2.0, power_p)
    ax, ay = abs(x), abs(y)
    if ax > ay:
        s, pp = ax, (y if x > 0 else 4 * ax - y)
    else:
        s, pp = ay, (2 * ay - x if y > 0 else 6 * ay + x)
    pp = (pp + 8 * s * math.floor(pw * D.r01())) / pw
    if pp <= s:
        return w * s, w * pp
    if pp <= 3 * s:
        return w * (2 * s - pp), w * s
    if pp <= 5 * s:
        return -w * s, w * (4 * s - pp)
    if pp <= 7 * s:
        return -w * (6 * s - pp), -w * s
    return w * s, -w * (8 * s - pp)


def lace_js_ref(x, y, w, D):
    r0, wt, s32 = math.hypot(x, y), D.r01(), 0.8660254037844386
    if wt > 0.75:
        th = math.atan2(y, x - 1)
        return w * (-r0 * math.sin(th) / 2), w * (-r0 * math.cos(th) / 2 + 1)
    if wt > 0.5:
        th = math.atan2(y - s32, x + 0.5)
        return w * (-r0 * math.sin(th) / 2 + s32), w * (-r0 * math.cos(th) / 2 - 0.5)
    if wt > 0.25:
        th = math.atan2(y + s32, x + 0.5)
        return w * (-r0 * math.sin(th) / 2 - s32), w * (-r0 * math.cos(th) / 2 - 0.5)
    th = math.atan2(y, x)
    return w * (-r0 * math.sin(th) / 2), w * (-r0 * math.cos(th) / 2)


def _csqrt(z):   # WGSL complex_sqrt: the sign of the imaginary part from z.y < 0
    rad = abs(z)
    im = math.sqrt(max(0.5 * (rad - z.real), 0.0))
    return complex(math.sqrt(max(0.5 * (rad + z.real), 0.0)), -im if z.imag < 0 else im)


def julia_outside_ref(x, y, w, re_div, im_div, mode_p, D):
    m = int(mode_p)
    z, z2 = complex(x, y), complex(x, y)
    if m in (0, 2):
        z = _csqrt(z)
    z += 1
    if m in (0, 2):
        z, z2 = z * z, _csqrt(z2)
    z2 -= 1
    if m in (0, 2):
        z2 = z2 * z2
    z = z / z2
    if m in (0, 1):
        z = _csqrt(z)
    z = z / complex(re_div, im_div)
    if m in (0, 1):
        z = z * (1 if D.r01() < 0.5 else -1)
    return w * z.real, w * z.imag


def rays2_ref(x, y, w, D):
    t = x * x + y * y
    ts = max(t, 1e-30)
    ci = math.cos((ts + 1e-6) * math.tan(1 / ts + 1e-6))
    u = 1 / (1e-30 if abs(ci) < 1e-30 else ci)
    c = w / 10
    return c * u * t / (x if x != 0 else 1e-30), c * u * t / (y if y != 0 else 1e-30)


def e_julia_ref(x, y, w, power_p, D):
    ps = 1.0 if power_p == 0 else power_p
    r2 = x * x + y * y
    xx = x if ps > 0 else x / max(r2, 1e-30)
    if ps <= 0:
        r2 = 1 / max(r2, 1e-30)
    tmp, tmp2 = r2 + 1, 2 * xx
    xmax = max((math.sqrt(max(tmp + tmp2, 0.0)) + math.sqrt(max(tmp - tmp2, 0.0))) * 0.5, 1.0)
    mu = math.acosh(xmax)
    nu = math.acos(min(max(xx / xmax, -1.0), 1.0))
    if y < 0:
        nu = -nu
    nu = nu / ps + 2 * PI / ps * math.floor(D.r01() * abs(ps))
    mu = mu / ps
    return w * math.cosh(mu) * math.cos(nu), w * math.sinh(mu) * math.sin(nu)


def cannabis_ref(x, y, w, filled, D):
    a = math.atan2(y, x)
    r = ((1 + 0.9 * math.cos(8 * a)) * (1 + 0.1 * math.cos(24 * a)) * (0.9 + 0.1 * math.cos(200 * a))
         * (1 + math.sin(a)))
    a += PI * 0.5
    if filled > 0 and filled > D.r01():
        r *= D.r01()
    return w * math.sin(a) * r, w * math.cos(a) * r


def intersection_ref(x, y, w, xw, xt, xm1, xm2, xh, yh, yt, ym1, ym2, yw, D):
    if w == 0:
        return 0.0, 0.0
    if D.r01() < 0.5:
        s = xw if D.r01() < 0.5 else -xw
        ox = xt * (x + round(s * math.log(max(D.r01(), 1e-30))))
        r1 = max(xm2 * xm1, 1e-30)
        if y > xm1:
            oy = xh * (-xm1 + math.fmod(y + xm1, r1))
        elif y < -xm1:
            oy = xh * (xm1 - math.fmod(xm1 - y, r1))
        else:
            oy = xh * y
        return ox, oy
    s = yh if D.r01() < 0.5 else -yh
    oy = yt * (y + round(s * math.log(max(D.r01(), 1e-30))))
    r1 = max(ym2 * ym1, 1e-30)
    if x > ym1:
        ox = yw * (-ym1 + math.fmod(x + ym1, r1))
    elif x < -ym1:
        ox = yw * (ym1 - math.fmod(ym1 - x, r1))
    else:
        ox = yw * x
    return ox, oy


def boarders2_ref(x, y, w, c, left, right, D):
    ac, al, ar = max(abs(c), 1e-6), max(abs(left), 1e-6), max(abs(right), 1e-6)
    cl, cr = ac * al, ac + ac * ar
    rx, ry = round(x), round(y)
    ox, oy = x - rx, y - ry
    if D.r01() >= cr:
        return w * (ox * ac + rx), w * (oy * ac + ry)
    nz = lambda v: 1e-30 if v == 0 else v
    if abs(ox) >= abs(oy):
        sg = 1 if ox >= 0 else -1
        return w * (ox * ac + rx + sg * cl), w * (oy * ac + ry + sg * cl * oy / nz(ox))
    sg = 1 if oy >= 0 else -1
    return w * (ox * ac + rx + sg * ox / nz(oy) * cl), w * (oy * ac + ry + sg * cl)


def b_transform_ref(x, y, w, rotate, power_p, move_v, split, D):
    pw = max(1.0, power_p)
    xp1, xm1, y2 = x + 1, x - 1, y * y
    tau = 0.5 * (math.log(max(xp1 * xp1 + y2, 1e-30)) - math.log(max(xm1 * xm1 + y2, 1e-30))) / pw + move_v
    sig = PI - math.atan2(y, xp1) - math.atan2(y, 1 - x) + rotate
    sig = sig / pw + 2 * PI / pw * math.floor(D.r01() * pw)
    tau += split if x >= 0 else -split
    t = math.cosh(tau) - math.cos(sig)
    t = 1e-30 if abs(t) < 1e-30 else t
    return w * math.sinh(tau) / t, w * math.sin(sig) / t


def apollony_ref(x, y, w, D):
    s3 = 1.7320508075688772
    dx = 1 + s3 - x
    den = max(dx * dx + y * y, 1e-30)
    a0, b0 = 3 * dx / den - (1 + s3) / (2 + s3), 3 * y / den
    ab = max(a0 * a0 + b0 * b0, 1e-30)
    f1x, f1y = a0 / ab, -b0 / ab
    br = int(math.floor(4 * D.r01())) % 3
    if br == 0:
        return w * a0, w * b0
    if br == 1:
        return w * (-f1x * 0.5 - f1y * s3 * 0.5), w * (f1x * s3 * 0.5 - f1y * 0.5)
    return w * (-f1x * 0.5 + f1y * s3 * 0.5), w * (-f1x * s3 * 0.5 - f1y * 0.5)


def circlecrop_ref(x, y, w, radius, cx, cy, scatter, zero_p, D):
    ca = min(max(scatter, -1.0), 1.0)
    dx, dy = x - cx, y - cy
    rad, ang = math.hypot(dx, dy), math.atan2(dy, dx)
    rdc = radius + D.r01() * 0.5 * ca
    esc = rad > radius
    if zero_p > 0.5 and esc:
        return 0.0, 0.0
    if esc:
        return w * rdc * math.cos(ang) + cx, w * rdc * math.sin(ang) + cy
    return w * dx + cx, w * dy + cy


def henon_ref(x, y, w, a, b, c, D):
    return w * (c - a * x * x + y), w * b * x


def lozi_ref(x, y, w, a, b, c, D):
    return w * (c - a * abs(x) + y), w * b * x


def _i32(v):
    v &= 0xFFFFFFFF
    return v - (1 << 32) if v >= 1 << 31 else v


def _voron_noise(x):   # JWildfire's DiscretNoise in wrapping 32-bit integers
    s = _i32((x << 13) ^ x)
    r = _i32(s * _i32(_i32(s * s) * 15731 + 789221) + 1376312589) & 0x7FFFFFFF
    return r / 2147483647.0


def voron_ref(x, y, w, k, step, num_f, xseed_f, yseed_f, D):
    xs, ys = int(xseed_f), int(yseed_f)
    st = 1e-30 if step == 0 else step
    M, N = math.floor(x / st), math.floor(y / st)
    best = (20.0, 0.0, 0.0)
    for i in (-1, 0, 1):
        for j in (-1, 0, 1):
            m1, n1 = M + i, N + j
            cnt = min(max(1 + math.floor(num_f * _voron_noise(19 * m1 + 257 * n1 + xs)), 1), 26)
            for l in range(cnt):
                cx = (_voron_noise(l + 64 * m1 + 15 * n1 + xs) + m1) * step
                cy = (_voron_noise(l + 21 * m1 + 33 * n1 + ys) + n1) * step
                r = math.hypot(x - cx, y - cy)
                if r < best[0]:
                    best = (r, cx, cy)
    _, x0, y0 = best
    return w * (k * (x - x0) + x0), w * (k * (y - y0) + y0)


JWF_DRAWS = [   # (entry, oracle, the reference's smoke-row parameters, more parameters)
    ("voron", voron_ref, (0.99, 0.25, 1.0, 3.0, 7.0), (0.5, 0.4, 3.0, -11.0, 5.0)),
    ("erf", erf_ref, (), None),
    ("asteria", asteria_ref, (0.5,), (0.2,)),
    ("henon", henon_ref, (1.4, 0.3, 0.6), (1.1, -0.4, 1.0)),
    ("chrysanthemum", chrysanthemum_ref, (), None),
    ("checks", checks_ref, (0.3, 0.3, 0.0, 0.0), (0.2, -0.1, 0.5, 0.7)),
    ("circleblur", circleblur_ref, (), None),
    ("hypertile1", hypertile1_ref, (3.0, 7.0), (5.0, 4.0)),
    ("line", line_ref, (0.5, 0.5), (0.2, -0.3)),
    ("vogel", vogel_ref, (8.0, 0.3), (13.0, -0.2)),
    ("yin_yang", yin_yang_ref, (1.0, 1.0, 0.5, 0.5, 0.5), (0.4, 0.3, -0.6, 1.0, 1.0)),
    ("squish", squish_ref, (2.0,), (5.0,)),
    ("lace_js", lace_js_ref, (), None),
    ("julia_outside", julia_outside_ref, (2.0, 0.0, 0.0), (1.5, 0.5, 1.0)),
    ("rays2", rays2_ref, (), None),
    ("e_julia", e_julia_ref, (2.0,), (-3.0,)),
    ("cannabis_curve_wf", cannabis_ref, (1.0,), (0.4,)),
    ("intersection", intersection_ref, (1.0, 0.0, 0.5, 1.0, 0.5, 1.0, 0.5, 0.5, 0.0, 0.5),
     (0.7, 0.4, 0.6, 2.0, 0.9, 0.3, 0.8, 0.4, 1.5, 0.6)),
    ("lozi", lozi_ref, (1.7, 0.5, 0.5), (1.2, -0.3, 1.0)),
    ("boarders2", boarders2_ref, (0.4, 0.65, 0.35), (0.6, 0.3, 0.5)),
    ("b_transform", b_transform_ref, (1.0, 0.0, 0.0, 0.0), (0.3, 3.0, 0.2, 0.1)),
    ("apollony", apollony_ref, (), None),
    ("circlecrop", circlecrop_ref, (1.0, 0.0, 0.0, 0.0, 1.0), (0.8, 0.1, -0.2, 0.6, 0.0)),
]


# flam3's draw-taking variations (flam3 variations.c, as src/shaders/chaos.wgsl:
# 1036-1170 and 1540-1620 state them). The reference's fixture rows for these
# only check that the output stays finite, since flam3's own RNG stream is not
# reproduced; with the draws scripted, the values are fixed.
def noise_ref(x, y, w, D):
    a = D.r01() * 2 * PI
    r = w * D.r01()
    return x * r * math.cos(a), y * r * math.sin(a)


def blur_ref(x, y, w, D):
    a = D.r01() * 2 * PI
    r = w * D.r01()
    return r * math.cos(a), r * math.sin(a)


def gaussian_blur_ref(x, y, w, D):
    a = D.r01() * 2 * PI
    r = w * (sum(D.r01() for _ in range(4)) - 2)
    return r * math.cos(a), r * math.sin(a)


def arch_ref(x, y, w, D):
    a = D.r01() * w * PI
    return w * math.sin(a), w * math.sin(a) ** 2 / math.cos(a)


def radial_blur_ref(x, y, w, angle, D):
    spin, zoom = math.sin(angle * PI / 2), math.cos(angle * PI / 2)
    g = w * (sum(D.r01() for _ in range(4)) - 2)
    ra, a = math.hypot(x, y), math.atan2(y, x) + spin * g
    rz = zoom * g - 1
    return ra * math.cos(a) + rz * x, ra * math.sin(a) + rz * y


def square_ref(x, y, w, D):
    r0 = D.r01()
    return w * (r0 - 0.5), w * (D.r01() - 0.5)


def rays_ref(x, y, w, D):
    a = w * D.r01() * PI
    t = w * math.tan(a) * w / (x * x + y * y + 1e-10)
    return t * math.cos(x), t * math.sin(y)


def blade_ref(x, y, w, D):
    r = D.r01() * w * math.hypot(x, y)
    return w * x * (math.cos(r) + math.sin(r)), w * x * (math.cos(r) - math.sin(r))


def twintrian_ref(x, y, w, D):
    r = D.r01() * w * math.hypot(x, y)
    s = math.sin(r)
    diff = math.log10(s * s) + math.cos(r) if s != 0 else -30.0
    if not -1e10 <= diff <= 1e10:
        diff = -30.0
    return w * x * diff, w * x * (diff - s * PI)


def super_shape_ref(x, y, w, rnd, m, n1, n2, n3, holes, D):
    r0 = math.hypot(x, y)
    th = m / 4 * math.atan2(y, x) + PI / 4
    t = abs(math.cos(th)) ** n2 + abs(math.sin(th)) ** n3
    r = w * ((rnd * D.r01() + (1 - rnd) * r0) - holes) * t ** (-1 / n1) / r0
    return r * x, r * y


def flower_ref(x, y, w, petals, holes, D):
    r0 = math.hypot(x, y)
    r = w * (D.r01() - holes) * math.cos(petals * math.atan2(y, x)) / r0
    return r * x, r * y


def conic_ref(x, y, w, ecc, holes, D):
    r0 = math.hypot(x, y)
    r = w * (D.r01() - holes) * ecc / (1 + ecc * x / r0) / r0
    return r * x, r * y


def parabola_ref(x, y, w, height, width, D):
    r = math.hypot(x, y)
    r0 = D.r01()
    return height * w * math.sin(r) ** 2 * r0, width * w * math.cos(r) * D.r01()


def pie_ref(x, y, w, slices, rotation, thickness, D):
    sl = math.trunc(D.r01() * slices + 0.5)
    a = rotation + 2 * PI * (sl + D.r01() * thickness) / slices
    r = w * D.r01()
    return r * math.cos(a), r * math.sin(a)


def boarders_ref(x, y, w, D):
    r0 = D.r01()
    rx, ry = math.floor(x + 0.5), math.floor(y + 0.5)
    ox, oy = x - rx, y - ry
    if r0 >= 0.75:
        return w * (ox * 0.5 + rx), w * (oy * 0.5 + ry)
    if abs(ox) >= abs(oy):
        s = 0.25 if ox >= 0 else -0.25
        return w * (ox * 0.5 + rx + s), w * (oy * 0.5 + ry + s * oy / ox)
    s = 0.25 if oy >= 0 else -0.25
    return w * (ox * 0.5 + rx + s * ox / oy), w * (oy * 0.5 + ry + s)


# The spatial-hash entries (src/shaders/chaos.wgsl:4046-4077, 4200-4222,
# 4331-4352, 5219-5233): the hash multiplies a sine by 43758.5453, so an f32
# result moves with the last bit of its input; the f64 instantiation, pinned
# here, does not.
def _frac_trunc(v):
    return v - math.trunc(v)


def circular_ref(x, y, w, angle_deg, seed, D, kx=12.9898, ky=78.233):
    ca = angle_deg * PI / 180
    aux = _frac_trunc(math.sin(x * kx + y * ky + seed) * 43758.5453)
    rnd = (2 * (D.r01() + aux) - 2) * ca
    rad, ang = math.hypot(x, y), math.atan2(y, x)
    return w * math.cos(ang + rnd) * rad, w * math.sin(ang + rnd) * rad


def circular2_ref(x, y, w, angle_deg, seed, xx, yy, D):
    return circular_ref(x, y, w, angle_deg, seed, D, kx=xx, ky=yy)


def hypertile2_ref(x, y, w, p_p, q, D):
    pa = 2 * PI / max(3.0, p_p)
    den = math.cos(pa) + math.cos(2 * PI / max(3.0, q))
    r2 = 1 - (math.cos(pa) - 1) / den if abs(den) > 1e-30 else 1.0
    r = 1 / math.sqrt(max(r2, 1e-30)) if r2 > 0 else 1.0
    a, b, c, d = x + r, y, r * x + 1, r * y
    xx, yy = a * c + b * d, b * c - a * d
    vr = w / max(c * c + d * d, 1e-30)
    rpa = math.floor(D.r01() * 32767) * pa
    return vr * (xx * math.cos(rpa) + yy * math.sin(rpa)), vr * (yy * math.cos(rpa) - xx * math.sin(rpa))


def phoenix_julia_ref(x, y, w, power, dist, xd, yd, D):
    pw = 1.0 if power == 0 else power
    k = math.floor(D.r01() * 32768)
    a = math.atan2(y * (yd + 1), x * (xd + 1)) * dist / pw + k * 2 * PI / pw
    r = w * max(x * x + y * y, 1e-30) ** (dist / (2 * pw))
    return r * math.cos(a), r * math.sin(a)


def waves4_ref(x, y, w, sx, sy, fx, fy, cont, yfact, D):
    cell = math.floor(y * fx / (2 * PI))
    ax = _frac_trunc(math.sin(cell * 12.9898 + cell * 78.233 + 1 + y * 0.001 * yfact) * 43758.5453)
    if cont > 0.5:
        ax = 1.0 if ax > 0.5 else 0.0
    return w * (x + math.sin(y * fx) * ax * ax * sx), w * (y + math.sin(x * fy) * sy)


# Entries the reference's own tests check only for finite output
# (src/post-flam3-batch2b-a.gpu.test.ts:186-317, issue243-bwraps-crackle.gpu.
# test.ts:53-110, issue120-b5-glynnsim.gpu.test.ts:122-144), restated from
# src/shaders/chaos.wgsl:1988-2140 (bwraps, crackle), 2149-2344 (juliaq,
# glynnia, falloff, falloff2, falloff3) and 3453-3545 (glynnSim1-3); and the
# position part of the direct-colour entries (:7849-7856).
def bwraps_ref(x, y, w, cellsize, space, gain, inner, outer, D):
    if abs(cellsize) < 1e-30:
        return w * x, w * y
    radius = 0.5 * cellsize / (1 + space * space)
    g2 = gain * gain / cellsize + 1e-6
    mb = g2 * radius
    mb = 1.0 if mb > 2 else mb / (mb * mb / 4 + 1)
    r2, rfac = radius * radius, radius / max(mb, 1e-30)
    cx, cy = (math.floor(x / cellsize) + 0.5) * cellsize, (math.floor(y / cellsize) + 0.5) * cellsize
    lx, ly = x - cx, y - cy
    if lx * lx + ly * ly > r2:
        return w * x, w * y
    lx, ly = lx * g2, ly * g2
    rd = rfac / ((lx * lx + ly * ly) / 4 + 1)
    lx, ly = lx * rd, ly * rd
    f = (lx * lx + ly * ly) / max(r2, 1e-30)
    th = inner * (1 - f) + outer * f
    return w * (cx + math.cos(th) * lx + math.sin(th) * ly), w * (cy - math.sin(th) * lx + math.cos(th) * ly)


def crackle_ref(x, y, w, cellsize, power, distort, scale, D):
    if abs(cellsize) < 1e-30:
        return w * x, w * y
    s = cellsize / 2
    r = [D.r01() for _ in range(4)]
    blurr = (r[0] + r[1]) / 2 + (r[2] - 0.5) / 4
    ux, uy = blurr * math.sin(2 * PI * r[3]), blurr * math.cos(2 * PI * r[3])

    def centre(cx, cy):   # a cell's centre moved by two noise samples
        return ((cx + distort * perlin2d_ref(cx * 2.5, cy * 2.5)) * s,
                (cy + distort * perlin2d_ref(cy * 2.5 + 30.2, cx * 2.5 - 12.1)) * s)

    def grid(gx, gy):
        return [centre(gx + i // 3 - 1, gy + i % 3 - 1) for i in range(9)]

    gx, gy = math.floor(ux / s), math.floor(uy / s)
    P = grid(gx, gy)
    d2 = [(px - ux) ** 2 + (py - uy) ** 2 for px, py in P]
    q = d2.index(min(d2))
    P = grid(gx + q // 3 - 1, gy + q % 3 - 1)
    (qx, qy), L = P[4], -1e30
    for i in range(9):
        if i != 4:
            ex, ey = P[i][0] - qx, P[i][1] - qy
            den = ex * ex + ey * ey
            L = max(L, 1.0 if den < 1e-30 else 2 * ((ux - qx) * ex + (uy - qy) * ey) / den)
    ls = L + 1e-30
    R = abs(ls) ** power * scale / ls
    return w * (qx + (ux - qx) * R), w * (qy + (uy - qy) * R)


def juliaq_ref(x, y, w, power, divisor, D):
    pw = 1.0 if abs(power) < 1e-30 else power
    n = math.floor(D.r01() * max(1.0, abs(pw)))
    a = math.atan2(y, x) * divisor / pw + n * 2 * PI / pw
    r = w * (x * x + y * y) ** (0.5 * divisor / pw)
    return r * math.cos(a), r * math.sin(a)


def glynnia_ref(x, y, w, D):
    v2 = w * math.sqrt(0.5)
    r, coin = math.hypot(x, y), D.r01()
    sg = 1.0 if r >= 1 else -1.0
    if coin > 0.5:
        if r + x <= 0:
            return 0.0, 0.0
        d = math.sqrt(r + x)
        return sg * v2 * d, -(v2 / d) * y
    d = r + x
    rad = r * (y * y + d * d)
    if rad <= 0:
        return 0.0, 0.0
    rr = w / math.sqrt(rad)
    return sg * rr * d, rr * y


def _falloff_dist(x, y, scatter, mindist, x0, y0):
    return max((max(math.hypot(x - x0, y - y0), 0.0) - mindist) * 0.04 * scatter, 0.0)


def falloff_ref(x, y, w, scatter, mindist, mx, my, x0, y0, D):
    d = _falloff_dist(x, y, scatter, mindist, x0, y0)
    r0 = D.r01()
    return w * (x + mx * r0 * d), w * (y + my * D.r01() * d)


def falloff2_ref(x, y, w, scatter, typ, mx, my, x0, y0, mindist, D):
    d = _falloff_dist(x, y, scatter, mindist, x0, y0)
    r0, r1, r2 = D.r01(), D.r01(), D.r01()
    t = int(min(max(typ, 0.0), 2.0))
    if t == 1:
        phi = math.atan2(y, x) + my * d * r1
        rr = math.hypot(x, y) + mx * r0 * d
        return w * rr * math.cos(phi), w * rr * math.sin(phi)
    if t == 2:
        sc, phi, rad = math.cos(d * r1 * 2 * PI), d * r2 * PI, d * r0
        return w * (x + mx * rad * sc * math.cos(phi)), w * (y + my * rad * sc * math.sin(phi))
    return w * (x + mx * r0 * d), w * (y + my * r1 * d)


def falloff3_ref(x, y, w, scatter, mx, my, x0, y0, mindist, inv, D):
    rad0 = math.hypot(x - x0, y - y0)
    base = max(1 - rad0, 0.0) if inv > 0.5 else max(rad0, 0.0)
    d = max((base - mindist) * 0.04 * scatter, 0.0)
    r0, r1, r2 = D.r01() - 0.5, D.r01() - 0.5, D.r01() - 0.5
    sc, phi, rad = math.cos(d * r1 * 2 * PI), d * r2 * PI, d * r0
    return w * (x + mx * rad * sc * math.cos(phi)), w * (y + my * rad * sc * math.sin(phi))


def _glynn_inside(w, radius1, thickness, x1, y1, D):
    ri = radius1 * (thickness + (1 - thickness) * D.r01())
    phi = 2 * PI * D.r01()
    return w * (ri * math.cos(phi) + x1), w * (ri * math.sin(phi) + y1)


def glynnsim1_ref(x, y, w, radius, radius1, phi1, thickness, pw, contrast, D):
    a = phi1 * PI / 180
    x1, y1 = radius * math.cos(a), radius * math.sin(a)
    r = math.hypot(x, y)
    if r < radius:
        return _glynn_inside(w, radius1, thickness, x1, y1, D)
    alpha = radius / max(r, 1e-30)
    if D.r01() > contrast * abs(alpha) ** abs(pw):
        px, py = x, y
    else:
        px, py = alpha * alpha * x, alpha * alpha * y
    if (px - x1) ** 2 + (py - y1) ** 2 < radius1 * radius1:
        return _glynn_inside(w, radius1, thickness, x1, y1, D)
    return w * px, w * py


def glynnsim2_ref(x, y, w, radius, thickness, contrast, pw, phi1, phi2, D):
    p10, p20 = phi1 * PI / 180, phi2 * PI / 180
    rpt = radius + thickness
    gamma = thickness * (2 * radius + thickness) / max(rpt, 1e-30)
    r = math.hypot(x, y)
    if r < radius:
        ri = rpt - gamma * D.r01()
        ph = p10 + (p20 - p10) * D.r01()
        return w * ri * math.cos(ph), w * ri * math.sin(ph)
    alpha = radius / max(r, 1e-30)
    if D.r01() > contrast * abs(alpha) ** abs(pw):
        return w * x, w * y
    return w * alpha * alpha * x, w * alpha * alpha * y


def glynnsim3_ref(x, y, w, radius, thickness, contrast, pw, D):
    r1 = radius + thickness
    r2 = radius * radius / max(r1, 1e-30)
    gamma = r1 / max(r1 + r2, 1e-30)
    r = math.hypot(x, y)
    if r < r1:
        phi = 2 * PI * D.r01()
        ri = r1 if D.r01() < gamma else r2
        return w * ri * math.cos(phi), w * ri * math.sin(phi)
    alpha = radius / max(r, 1e-30)
    if D.r01() > contrast * abs(alpha) ** abs(pw):
        return w * x, w * y
    return w * alpha * alpha * x, w * alpha * alpha * y


def dc_zero_ref(x, y, w, *rest):   # colour-only entries add nothing to the position
    return 0.0, 0.0


def dc_cylinder_ref(x, y, w, D):
    return w * math.sin(x), w * y


SMOKE_DRAWS = [
    ("bwraps", bwraps_ref, (1.0, 0.0, 1.0, -1.04, 0.71), (0.7, 0.3, 1.6, 0.5, -0.9)),
    ("crackle", crackle_ref, (1.0, 0.2, 0.0, 1.0), (1.0, 0.2, 1.0, 1.0)),
    ("juliaq", juliaq_ref, (3.0, 2.0), (-4.0, 1.5)),
    ("glynnia", glynnia_ref, (), None),
    ("falloff", falloff_ref, (1.0, 0.5, 1.0, 1.0, 0.0, 0.0), (3.0, 0.2, 0.7, -1.3, 0.4, -0.2)),
    ("falloff2", falloff2_ref, (1.0, 0.0, 1.0, 1.0, 0.0, 0.0, 0.5), (2.0, 1.0, 0.8, 1.2, 0.3, 0.1, 0.2)),
    ("falloff2", falloff2_ref, (1.0, 2.0, 1.0, 1.0, 0.0, 0.0, 0.5), (2.5, 2.7, 0.6, -0.9, -0.3, 0.2, 0.1)),
    ("falloff3", falloff3_ref, (1.0, 1.0, 1.0, 0.0, 0.0, 0.5, 0.0), (2.0, 0.8, 1.1, 0.2, -0.1, 0.1, 1.0)),
    ("glynnSim1", glynnsim1_ref, (1.0, 0.1, 110.0, 0.1, 1.5, 0.5), (0.45, 0.43, 110.0, 0.25, 1.5, 0.5)),
    ("glynnSim2", glynnsim2_ref, (1.0, 0.1, 0.5, 1.5, 110.0, 150.0), (0.25, 0.1, 0.8, 1.5, 30.0, 200.0)),
    ("glynnSim3", glynnsim3_ref, (1.0, 0.1, 0.5, 1.5), (0.5, 0.2, 0.9, 1.3)),
    ("dc_linear", dc_zero_ref, (), None),
    ("dc_perlin", dc_zero_ref, (1.0, 3.0, 0.0), None),
    ("dc_gridout", dc_zero_ref, (4.0,), None),
    ("dc_cylinder", dc_cylinder_ref, (), None),
]


HASH_DRAWS = [
    ("circular", circular_ref, (90.0, 0.0), (35.0, 2.5)),
    ("circular2", circular2_ref, (90.0, 0.0, 12.9898, 78.233), (40.0, 1.0, 7.3, -3.1)),
    ("hypertile2", hypertile2_ref, (3.0, 7.0), (5.0, 4.0)),
    ("phoenix_julia", phoenix_julia_ref, (2.0, 1.0, 0.0, 0.0), (-3.0, 0.7, 0.4, -0.2)),
    ("waves4", waves4_ref, (0.05, 0.05, 7.0, 13.0, 0.0, 1.0), (0.3, 0.2, 3.0, 5.0, 1.0, 2.0)),
]


FLAM3_DRAWS = [   # (entry, oracle, the fixture row's parameters, more parameters)
    ("noise", noise_ref, (), None),
    ("blur", blur_ref, (), None),
    ("gaussian_blur", gaussian_blur_ref, (), None),
    ("arch", arch_ref, (), None),
    ("radial_blur", radial_blur_ref, (0.5,), (-1.3,)),
    ("square", square_ref, (), None),
    ("rays", rays_ref, (), None),
    ("blade", blade_ref, (), None),
    ("twintrian", twintrian_ref, (), None),
    ("super_shape", super_shape_ref, (0.5, 6.0, 2.0, 2.0, 2.0, 0.0), (0.2, 5.0, 3.0, 1.5, 0.7, 0.3)),
    ("flower", flower_ref, (6.0, 0.3), (3.0, -0.6)),
    ("conic", conic_ref, (0.7, 0.2), (0.4, -0.1)),
    ("parabola", parabola_ref, (0.5, 0.5), (0.55, -0.8)),
    ("pie", pie_ref, (6.0, 0.5, 0.5), (5.0, -0.3, 0.9)),
    ("boarders", boarders_ref, (), None),
]


@pytest.mark.parametrize("case", JWF_DRAWS + SMOKE_DRAWS + HASH_DRAWS + FLAM3_DRAWS, ids=lambda c: c[0])
def test_values_with_scripted_draws(orc, case):
    """Entries restated with scripted RNG words (those that draw nothing take
    none): the value for every point, weight, parameter set and six word
    sequences, f64 within 1e-6 relative, draws consumed in the order the
    restatement takes them. Where the restatement divides by zero (a lattice
    point of boarders, the origin), the oracle must give a non-finite value
    there."""
    name, ref, p0, p1 = case
    if case in JWF_DRAWS:   # the reference's smoke rows for these ports
        check_finite(orc, name, 1.0, p0, LALL_PTS)
    pts = LALL_PTS + P8 + [(0.3, -0.8), (-0.2, -0.45), (1.7, 0.9)]
    n = 0
    for prm in [p0] + ([p1] if p1 is not None else []):
        for s in range(6):
            words = WORDS[s * 16:(s + 1) * 16]
            for w in (1.0, 0.7):
                for x, y in pts:
                    got = run(orc, name, x, y, w, prm, words=words)
                    try:
                        want = ref(x, y, w, *prm, Draws(words))
                    except ZeroDivisionError:
                        assert not all(math.isfinite(v) for v in got), (name, prm, x, y, got)
                        continue
                    for k in range(2):
                        assert abs(got[k] - want[k]) <= 1e-6 * max(1.0, abs(want[k])), \
                            (name, prm, s, w, x, y, k, got[k], want[k])
                    n += 1
    assert n > 0


# -- end of pins --
