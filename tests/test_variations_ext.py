"""Oracle restatement of catalog entries V99..V325 (CPU only).

No flam3 fixture covers these entries (the reference pins them with small
in-file checks in its src/issue*.gpu.test.ts and post-flam3-batch*.gpu.test.ts
files). The oracle is pinned here by:
  * closed forms the mathematics fixes (Gamma(n+1) = n!, Lambert W e^W = z,
    the Collatz interpolation at integers, Chebyshev T_n(cos t) = cos(n t),
    Klein j(i) = 1728, modular lambda(i) = 1/2, Blaschke |B| = 1 on the unit
    circle, the A&S / Hastings / Winitzki approximations against scipy within
    their published error bounds, ...), chosen so that a dropped term, a
    wrong sign or a swapped argument fails;
  * the reference tests' literal expectations where they state numbers
    (file:line cited per case);
  * the RNG draw contract of every randomised entry (how many draws, always
    taken or conditional), which the GPU parity test then checks per walker;
  * an f32 / f64 consistency sweep over every entry at the reference's
    default parameters (src/serialize.ts:487-860).
"""
import math

import numpy as np
import pytest
from scipy import special

from pyr3_amd import catalog as cat
from pyr3_amd.variations_supported import SUPPORTED

V = cat.VARIATION_INDEX
EXT = list(range(99, 326))
PI = math.pi


def var(orc, name, x, y, params=None, w=1.0, words=(), dtype=np.float64):
    p = cat.VARIATION_DEFAULTS.get(name, []) if params is None else params
    return orc.variation(V[name], x, y, w, p, rng_words=words, dtype=dtype)


def test_every_entry_is_restated(orc):
    assert all(orc.variation_supported(i) for i in EXT)
    assert set(EXT) <= SUPPORTED and max(SUPPORTED) == 325


# -- f32 / f64 consistency ----------------------------------------------------

# Entries whose f32 value is a discontinuous function of the f64 one at
# ordinary inputs: a sin-hash fract (circular, circular2, waves4), a random
# integer times a rounded constant (hypertile2, phoenix_julia), reflections at
# a boundary hit at t ~ 0 (billiards), 1/t near an integer (rays2).
ILL = {"circular", "circular2", "waves4", "hypertile2", "phoenix_julia",
       "billiard_circle", "billiard_stadium", "rays2"}


@pytest.mark.parametrize("idx", EXT)
def test_f32_tracks_f64_at_defaults(orc, idx):
    name = cat.VARIATION_NAMES[idx]
    rng = np.random.default_rng(1000 + idx)
    words = [int(v) for v in rng.integers(0, 2 ** 32, 64, dtype=np.uint64)]
    pts = rng.uniform(-1.5, 1.5, (48, 2)).astype(np.float32).astype(np.float64)
    ok = 0
    for x, y in pts:
        a = var(orc, name, x, y, w=0.8, words=words)
        b = var(orc, name, x, y, w=0.8, words=words, dtype=np.float32)
        assert a[2] == b[2], (name, x, y, "draw count differs")
        fa, fb = np.isfinite(a[:2]).all(), np.isfinite(b[:2]).all()
        if fa and fb:
            rel = max(abs(a[0] - b[0]), abs(a[1] - b[1])) / max(1.0, abs(a[0]), abs(a[1]))
            ok += rel <= 1e-3
        else:
            ok += fa == fb
    if name not in ILL:   # ill-conditioned entries: draw counts only
        assert ok >= 0.95 * len(pts), (name, ok)


# -- RNG draw contract (chaos.wgsl per entry) ------------------------------------

FIXED_DRAWS = {
    "cpow2": 3, "cpow3": 4, "epispiral": 1, "crackle": 4, "juliaq": 1, "glynnia": 1,
    "falloff": 2, "falloff2": 3, "falloff3": 3, "xtrb": 2, "blur_circle": 2, "acosh": 1,
    "acosech": 1, "lissajous": 2, "spirograph": 2, "chrysanthemum": 1, "checks": 2,
    "circular": 1, "circular2": 1, "circleblur": 2, "hypertile1": 1, "hypertile2": 1,
    "line": 1, "phoenix_julia": 1, "vogel": 1, "squish": 1, "lace_js": 1, "e_julia": 1,
    "boarders2": 1, "b_transform": 1, "apollony": 1, "circlecrop": 1,
}


@pytest.mark.parametrize("name", sorted(FIXED_DRAWS))
def test_fixed_draw_counts(orc, name):
    for x, y in ((0.3, -0.2), (1.7, 0.9), (-0.6, 1.1)):
        assert var(orc, name, x, y)[2] == FIXED_DRAWS[name], (name, x, y)


def _w(u):
    """An RNG word whose rand01 is ~u."""
    return int(round(u * 0x0FFFFFFF)) & 0x0FFFFFFF


def test_conditional_draws(orc):
    # bcircle draws only outside the unit circle (scale 1)
    assert var(orc, "bcircle", 0.3, 0.2, [1.0, 0.5])[2] == 0
    assert var(orc, "bcircle", 1.3, 0.2, [1.0, 0.5])[2] == 1
    # cannabis_curve_wf: a second draw only when filled > first draw
    assert var(orc, "cannabis_curve_wf", 0.3, 0.2, [0.0])[2] == 0
    assert var(orc, "cannabis_curve_wf", 0.3, 0.2, [0.5], words=[_w(0.9)])[2] == 1
    assert var(orc, "cannabis_curve_wf", 0.3, 0.2, [0.5], words=[_w(0.1), 0])[2] == 2
    # yin_yang: the dual draw only inside the unit disc with dual_t = 1
    assert var(orc, "yin_yang", 0.3, 0.2, [0.5, 0, 0, 1, 0])[2] == 1
    assert var(orc, "yin_yang", 0.3, 0.2, [0.5, 0, 0, 0, 0])[2] == 0
    assert var(orc, "yin_yang", 1.3, 0.2, [0.5, 0, 0, 1, 0])[2] == 0
    # julia_outside: the root-sign draw for modes 0 and 1 only
    assert [var(orc, "julia_outside", 0.3, 0.2, [1, 0, m])[2] for m in (0, 1, 2)] == [1, 1, 0]
    # intersection / parallel: branch draw + 2; waffle: 1 + per-mode count
    assert var(orc, "intersection", 0.3, 0.2)[2] == 3
    assert var(orc, "parallel", 0.3, 0.2)[2] == 3
    per_mode = {0: 5, 1: 4, 2: 4, 3: 4, 4: 4}
    for mode, n in per_mode.items():
        words = [_w((mode + 0.5) / 5)] + [_w(0.3)] * 8
        assert var(orc, "waffle", 0.1, 0.1, words=words)[2] == n, mode
    # glynnSim1: two draws inside the disc; one contrast draw outside (+2 if
    # the point lands in the small circle)
    assert var(orc, "glynnSim1", 0.1, 0.1, [1.0, 0.1, 110, 0.1, 1.5, 0.5])[2] == 2
    assert var(orc, "glynnSim1", 3.0, 0.1, [1.0, 0.1, 110, 0.1, 1.5, 0.5])[2] == 1
    # asteria: a draw only where the disc and the corner region overlap
    assert var(orc, "asteria", 0.6, 0.6, [0.0])[2] == 1
    assert var(orc, "asteria", 0.05, 0.02, [0.0])[2] == 0


# -- closed forms and reference literal checks --------------------------------------

def test_box_and_sphere_fold(orc):
    # src/issue129-folds.gpu.test.ts:100-116
    assert var(orc, "box_fold", 0.5, 0.5, [1.0])[:2] == pytest.approx((0.5, 0.5))
    assert var(orc, "box_fold", 1.5, -1.5, [1.0])[:2] == pytest.approx((0.5, -0.5))
    assert var(orc, "sphere_fold", 0.0, 0.0, [0.5, 1.0])[:2] == pytest.approx((0.0, 0.0))
    assert var(orc, "sphere_fold", 0.75, 0.0, [0.5, 1.0])[0] == pytest.approx(0.75 / 0.5625)
    assert var(orc, "sphere_fold", 0.2, 0.1, [0.5, 1.0])[:2] == pytest.approx((0.8, 0.4))
    assert var(orc, "sphere_fold", 2.0, 0.0, [0.5, 1.0])[0] == pytest.approx(2.0)


def test_complex_gamma_values(orc):
    for n in range(1, 7):   # Gamma(n) = (n-1)!
        g = var(orc, "complex_gamma", float(n), 0.0, [1.0])
        assert g[0] == pytest.approx(math.factorial(n - 1), rel=2e-6) and abs(g[1]) < 1e-6
    assert var(orc, "complex_gamma", 0.5, 0.0, [1.0])[0] == pytest.approx(math.sqrt(PI), rel=1e-6)
    # reflection branch: Gamma(-1/2) = -2 sqrt(pi); Gamma(1 + i) against scipy
    assert var(orc, "complex_gamma", -0.5, 0.0, [1.0])[0] == pytest.approx(-2 * math.sqrt(PI), rel=1e-6)
    g = var(orc, "complex_gamma", 1.0, 1.0, [2.0])
    ref = 2.0 * special.gamma(1 + 1j)
    assert complex(*g[:2]) == pytest.approx(ref, rel=1e-5)


def test_lambert_w_inverts(orc):
    for z in (0.5 + 0.2j, 1.5 - 0.7j, 3.0 + 4.0j, -0.2 + 0.1j):
        wv = complex(*var(orc, "lambert_w", z.real, z.imag, [4.0])[:2])
        assert abs(wv * np.exp(wv) - z) < 1e-6 * max(1, abs(z))
        assert wv == pytest.approx(special.lambertw(z), abs=1e-6)


def test_collatz_at_integers(orc):
    # even n -> n / 2, odd n -> (3n + 1) / 2
    for n in range(0, 9):
        out = var(orc, "collatz", float(n), 0.0, [1.0, 0.0])
        want = n / 2 if n % 2 == 0 else (3 * n + 1) / 2
        assert out[0] == pytest.approx(want, abs=1e-9) and abs(out[1]) < 1e-9


def test_orthogonal_polynomials(orc):
    t = np.linspace(0.1, 3.0, 9)
    for n in range(0, 9):
        for tt in t:
            x = math.cos(tt)
            out = var(orc, "chebyshev", x, 0.5, [float(n), 2.0])
            assert out[0] == pytest.approx(math.cos(n * tt), abs=1e-9)
            assert out[1] == pytest.approx(-0.5, abs=1e-9)          # T_2(1/2)
            lp = var(orc, "legendre", x, 1.0, [float(n), float(n)])
            assert lp[0] == pytest.approx(special.eval_legendre(n, x), abs=1e-9)
            assert lp[1] == pytest.approx(1.0)                      # P_n(1) = 1


def test_inverse_hyperbolics_match_numpy(orc):
    k = 2.0 / PI
    for z in (0.3 + 0.4j, -1.2 + 0.5j, 0.7 - 1.1j, 2.0 + 0.1j):
        assert complex(*var(orc, "arcsinh", z.real, z.imag)[:2]) == pytest.approx(k * np.arcsinh(z), abs=1e-9)
        # log((1 + z) / (1 - z)) = 2 atanh(z), scaled by 2/pi (chaos.wgsl:3172)
        assert complex(*var(orc, "arctanh", z.real, z.imag)[:2]) == pytest.approx(2 * k * np.arctanh(z), abs=1e-9)
        a = complex(*var(orc, "acosh", z.real, z.imag, words=[0])[:2])     # draw < 0.5: +
        assert np.cosh(a / k) == pytest.approx(z, abs=1e-9)
        b = complex(*var(orc, "acosh", z.real, z.imag, words=[0x0FFFFFFF])[:2])
        assert b == pytest.approx(-a)
        c = complex(*var(orc, "acoth", z.real, z.imag)[:2])   # (Im, Re) of atanh(1/z)
        ref = k * np.arctanh(1 / z)
        assert c == pytest.approx(complex(ref.imag, ref.real), abs=1e-9)


def test_conformal_maps(orc):
    # Blaschke factor z (z - a) / (1 - conj(a) z): modulus 1 on |z| = 1
    for t in np.linspace(0, 6, 7):
        out = var(orc, "blaschke", math.cos(t), math.sin(t), [0.3, -0.2])
        assert math.hypot(*out[:2]) == pytest.approx(1.0, abs=1e-9)
    # Cayley (z - is) / (z + is) sends the upper half-plane into the disc
    for z in (0.3 + 0.2j, -2 + 1j, 5 + 0.01j):
        out = var(orc, "cayley", z.real, z.imag, [1.0])
        assert math.hypot(*out[:2]) < 1.0
        assert complex(*out[:2]) == pytest.approx((z - 1j) / (z + 1j), abs=1e-12)
    # Newton / Halley / nova fix the cube roots of unity (c = 0)
    for k in range(3):
        r = complex(math.cos(2 * PI * k / 3), math.sin(2 * PI * k / 3))
        assert complex(*var(orc, "newton", r.real, r.imag, [3.0])[:2]) == pytest.approx(r, abs=1e-12)
        assert complex(*var(orc, "halley", r.real, r.imag, [0, 0])[:2]) == pytest.approx(r, abs=1e-12)
        assert complex(*var(orc, "nova", r.real, r.imag, [0, 0, 1])[:2]) == pytest.approx(r, abs=1e-12)
    # one Newton step for z^2 = 1 from z: (z^2 + 1) / (2z)
    z = 0.7 + 0.4j
    assert complex(*var(orc, "newton", z.real, z.imag, [2.0])[:2]) == pytest.approx((z * z + 1) / (2 * z))
    # magnet1: ((z^2 + c - 1) / (2z + c - 2))^2; burning ship (|x| + i|y|)^2 + c
    c = 0.2 - 0.3j
    want = ((z * z + c - 1) / (2 * z + c - 2)) ** 2
    assert complex(*var(orc, "magnet1", z.real, z.imag, [c.real, c.imag])[:2]) == pytest.approx(want)
    want = complex(abs(z.real), abs(z.imag)) ** 2 + c
    assert complex(*var(orc, "burning_ship", -z.real, -z.imag, [c.real, c.imag])[:2]) == pytest.approx(want)


def test_modular_forms_special_values(orc):
    # tau = i: lambda = 1/2, j = 1728 (log-compressed: log(1 + |j|) j/|j|)
    lam = var(orc, "modular_lambda", 0.0, 0.98, [0.0])
    assert lam[0] == pytest.approx(0.5, abs=1e-9) and abs(lam[1]) < 1e-9
    j = var(orc, "klein_j", 0.0, 0.98, [0.0])
    assert j[0] == pytest.approx(math.log(1729.0), rel=1e-6) and abs(j[1]) < 1e-6
    # theta3(q) at q = e^-pi: pi^(1/4) / Gamma(3/4)
    th = var(orc, "jacobi_theta", 0.0, 0.98, [0.0])
    assert th[0] == pytest.approx(PI ** 0.25 / special.gamma(0.75), rel=1e-9)
    # Weierstrass P is even: P(-z) = P(z)
    a = var(orc, "weierstrass_p", 0.31, 0.17)
    b = var(orc, "weierstrass_p", -0.31, -0.17)
    assert a[:2] == pytest.approx(b[:2], rel=1e-9)


def test_special_function_profiles(orc):
    # J0 below 3 is A&S 9.4.1 (|err| < 5e-8). Above 3 the reference evaluates
    # the 9.4.3 modulus/phase polynomials in (3/x)^2 rather than 3/x
    # (chaos.wgsl:6606-6614); the restatement keeps that.
    r = 0.9
    for freq in (0.5, 2.0, 3.2):
        out = var(orc, "bessel_j0", r, 0.0, [freq])
        assert out[0] == pytest.approx(special.j0(freq * r) * r, abs=1e-7)
    for freq in (1.0, 5.0, 10.0):
        out = var(orc, "struve_h1", r, 0.0, [freq])
        assert out[0] == pytest.approx(0.6 * special.struve(1, freq * r) * r, abs=1e-6)
    # Airy Ai: series region and both asymptotic tails (about 1% in the tails)
    for x in (-6.0, -2.0, 0.0, 1.5, 5.0):
        out = var(orc, "airy_radial", 1.0, 0.0, [x, 0.0])     # scale * (r - 0) = x
        ai = special.airy(x)[0]
        assert out[0] / 3.0 == pytest.approx(ai, abs=2e-3 if abs(x) > 4 else 1e-7)
    for m in (0.1, 0.5, 0.9):
        r2 = m / (1 - m)
        rr = math.sqrt(r2)
        assert var(orc, "elliptic_E", rr, 0.0, [1.0])[0] == pytest.approx(special.ellipe(m), abs=5e-5)
        assert var(orc, "elliptic_K", rr, 0.0, [1.0, 100.0])[0] == pytest.approx(special.ellipk(m), abs=5e-5)
    for x in (-2.0, -0.5, 0.1, 1.3):
        assert var(orc, "erf", x, x)[0] == pytest.approx(math.erf(x), abs=1.5e-7)
    for z in (1.5 + 0.0j, 3.2 + 0.0j, 0.7 + 0.0j):
        out = var(orc, "digamma", z.real, 0.0, [1.0, 0.0])
        assert out[0] == pytest.approx(special.digamma(z.real), abs=1e-6)


def test_distribution_warps(orc):
    # u = r^2 / (1 + r^2); Gaussian CDF inverse via Winitzki (|err| < 4e-3)
    for r in (0.5, 1.0, 2.0):
        u = r * r / (1 + r * r)
        out = var(orc, "gaussian_cdf", r, 0.0, [0.0, 1.0])
        assert out[0] == pytest.approx(max(special.ndtri(u), 0.0), abs=4e-3 * 1.5)
        lg = var(orc, "logistic_cdf", r, 0.0, [0.5, 1.0])
        assert lg[0] == pytest.approx(max(0.5 + math.log(u / (1 - u)), 0.0), rel=1e-9)
        wb = var(orc, "weibull_cdf", r, 0.0, [1.0, 1.0])
        assert wb[0] == pytest.approx(-math.log(1 - u), rel=1e-9)


def test_digit_codes(orc):
    ext = 1.0
    # 4-bit grid over [-1, 1): cell i -> centre -1 + (i + 0.5) / 8
    cells = np.arange(16)
    xs = -1 + (cells + 0.5) / 8.0
    def cell_of(v):
        return int(round((v + 1) * 8 - 0.5))
    gray = [cell_of(var(orc, "gray_code", x, 0.0, [ext, 4])[0]) for x in xs]
    assert gray == [i ^ (i >> 1) for i in cells]
    rad = [cell_of(var(orc, "radical_inverse", x, 0.0, [ext, 4])[0]) for x in xs]
    assert rad == [int(f"{i:04b}"[::-1], 2) for i in cells]
    # Morton: x bits to even positions, y bits to odd, split back
    out = var(orc, "morton_zorder", xs[5], xs[9], [ext, 4])
    code = 0
    for b in range(4):
        code |= ((5 >> b) & 1) << (2 * b) | ((9 >> b) & 1) << (2 * b + 1)
    assert (cell_of(out[0]), cell_of(out[1])) == (code & 15, (code >> 4) & 15)
    # Peano scramble is a bijection on 3^t cells
    t = 3
    n = 3 ** t
    cs = -1 + 2 * (np.arange(n) + 0.5) / n
    got = sorted(int(round((var(orc, "peano", v, 0.0, [ext, t])[0] + 1) * n / 2 - 0.5)) for v in cs)
    assert got == list(range(n))


def test_orbits_and_maps(orc):
    # Kepler with e = 0: E = M, output a (cos M, sin M)
    out = var(orc, "kepler_orbit", 0.6, 0.8, [0.0, 1.0])
    a = 1.0 + 1e-6
    assert out[:2] == pytest.approx((a * 0.6, a * 0.8))
    # e > 0: two Newton steps leave a small Kepler residual
    e = 0.3
    M = math.atan2(0.8, 0.6)
    out = var(orc, "kepler_orbit", 0.6, 0.8, [e, 1.0])
    E = math.atan2(out[1] / math.sqrt(1 - e * e), out[0] + e * a)
    assert abs(E - e * math.sin(E) - M) < 1e-3
    # Henon, Lozi, Tinkerbell, de Jong at a point (direct maps)
    assert var(orc, "henon", 0.5, 0.2, [1.4, 0.3, 1.0])[:2] == pytest.approx((1.0 - 1.4 * 0.25 + 0.2, 0.15))
    assert var(orc, "lozi", -0.5, 0.2, [1.7, 0.5, 1.0])[:2] == pytest.approx((1.0 - 0.85 + 0.2, -0.25))
    # Arnold cat map is area-preserving mod 1: integer points map to the origin cell
    assert var(orc, "arnold_cat", 1.0, 2.0)[:2] == pytest.approx((0.0, 0.0), abs=1e-12)
    # Mollweide: auxiliary angle satisfies 2t + sin 2t = pi sin(phi)
    phi = 0.7
    out = var(orc, "mollweide", 0.0, phi)
    t = math.asin(out[1] / 1.4142135)
    assert abs(2 * t + math.sin(2 * t) - 3.14159265 * math.sin(phi)) < 1e-6
    # Billiard in a circle of radius 1: stays inside, speed preserved
    for x, y, ang in ((0.2, 0.1, 0.3), (-0.5, 0.4, 2.0), (0.9, 0.0, 1.0)):
        out = var(orc, "billiard_circle", x, y, [1.0, 0.5, ang])
        assert math.hypot(*out[:2]) <= 1.0 + 1e-9
    # Gauss map: fract(1/x)
    assert var(orc, "gauss_map", 0.3, -0.7)[:2] == pytest.approx((1 / 0.3 % 1, (-1 / 0.7) % 1))


def test_noise_and_colour_helpers(orc):
    # Perlin noise vanishes on the integer lattice (dc_perlin at scale 1, one
    # octave: hue = fract(0.5 + 0.5 * 0 + seed))
    rgb = orc.dc_color(V["dc_perlin"], 3.0, -2.0, (1.0, 1.0, 0.0))
    assert rgb == pytest.approx((0.1, 1.0, 1.0), abs=1e-6)      # hue 1/2
    rgb = orc.dc_color(V["dc_perlin"], 3.0, -2.0, (1.0, 1.0, 0.5))  # seed rotates hue to 0
    assert rgb == pytest.approx((1.0, 0.1, 0.1), abs=1e-6)
    assert orc.dc_color(V["dc_linear"], 0.0, 0.0) == pytest.approx((0.5, 0.5, 0.5))
    assert orc.dc_color(V["dc_linear"], 5.0, -5.0) == pytest.approx((1.0, 0.0, 0.5))
    g = orc.dc_color(V["dc_gridout"], 0.31, 0.77, (4.0,))
    assert all(0 <= c < 1 for c in g) and g == orc.dc_color(V["dc_gridout"], 0.26, 0.99, (4.0,))
    # Newton basin colour at the k-th root of z^3 = 1: hue k / 3
    hues = []
    for k in range(3):
        r = complex(math.cos(2 * PI * k / 3), math.sin(2 * PI * k / 3))
        hues.append(orc.dc_color(V["newton"], r.real, r.imag, (3.0,)))
    assert hues[0] == pytest.approx((1.0, 0.1, 0.1), abs=1e-6)
    assert hues[1] == pytest.approx((0.1, 1.0, 0.1), abs=1e-6)
    assert hues[2] == pytest.approx((0.1, 0.1, 1.0), abs=1e-6)
    # entries without a colour side
    assert orc.dc_color(V["linear"], 0.1, 0.1) is None and orc.dc_color(V["box_fold"], 0, 0) is None
    # escape colours are valid RGB in [0, 1]
    for name in ("burning_ship", "magnet1", "nova", "halley"):
        c = orc.dc_color(V[name], 0.3, 0.2, (0.1, 0.2, 1.0))
        assert c is not None and all(-1e-6 <= v <= 1 + 1e-6 for v in c)


# -- the reference tests' literal expectations -------------------------------------

def _known_cases():
    import json
    from conftest import GOLDEN
    return json.loads((GOLDEN / "variation_known_answers.json").read_text())["cases"]


@pytest.mark.parametrize("case", _known_cases(), ids=lambda c: c["name"])
def test_reference_literal_expectations(orc, case):
    for i, k, want, digits in case["checks"]:
        x, y = case["points"][i]
        got = var(orc, case["name"], x, y, case["params"], dtype=np.float32)[k]
        assert abs(got - want) < 10.0 ** -digits / 2, (case["src"], i, k, got, want)


def test_fractal_series_match_reference_oracles(orc):
    """src/issue136-fractals.gpu.test.ts:107-185: the series summed in f64
    (with Math.PI) agree with the kernel to 4 decimals."""
    pts = [(0.1, 0.2), (-0.35, 0.6), (0.8, -0.45)]
    for px, py in pts:
        a, b, terms, amp = 0.5, 3.0, 4, 0.5
        wx = wy = 0.0
        ap = bp = 1.0
        for _ in range(terms):
            wx += ap * math.cos(bp * PI * px)
            wy += ap * math.cos(bp * PI * py)
            ap *= a
            bp *= b
        out = var(orc, "weierstrass", px, py, [a, b, terms, amp], dtype=np.float32)
        assert out[:2] == pytest.approx((px + amp * wx, py + amp * wy), abs=5e-5)
        tx = ty = 0.0
        p2 = 1.0
        for _ in range(4):
            tx += abs(p2 * px - math.floor(p2 * px + 0.5)) / p2
            ty += abs(p2 * py - math.floor(p2 * py + 0.5)) / p2
            p2 *= 2
        out = var(orc, "takagi", px, py, [4.0, 0.5], dtype=np.float32)
        assert out[:2] == pytest.approx((px + 0.5 * tx, py + 0.5 * ty), abs=5e-5)
        cx, cy = px, py
        for _ in range(4):
            cx = (cx + math.sin(cx * 2 * PI)) * 0.5
            cy = (cy + math.sin(cy * 2 * PI)) * 0.5
        out = var(orc, "cantor_stairs", px, py, [4.0, 0.5], dtype=np.float32)
        assert out[:2] == pytest.approx((px + 0.5 * cx, py + 0.5 * cy), abs=5e-5)
