"""Keyframe animation host logic (pyr3_amd.animate): interpolation, motion,
easing, temporal filter, channel curves.

The cases restate the reference's own unit tests with their expected
numbers (src/interpolate.test.ts, src/motion.test.ts, src/easing.test.ts,
src/temporal-filter.test.ts, src/channel-curves.test.ts), plus closed forms
that a dropped term or a swapped operand would break: a pure rotation keeps
unit magnitude, Catmull-Rom weights reproduce quadratics exactly, permutations
gather in the documented direction.
"""
import math

import pytest

from pyr3_amd import animate as A
from pyr3_amd import curves as cv
from pyr3_amd.catalog import VARIATION_INDEX as V
from pyr3_amd.genome import (Affine, ColorStop, Density, Genome, Palette, Tonemap, Variation,
                             Xform, genome_from_json, genome_to_json)


def vr(name, w, *params):
    v = Variation(V[name], float(w))
    for i, p in enumerate(params):
        v.params[i] = float(p)
    return v


def xf(a=1, b=0, c=0, d=0, e=1, f=0, weight=1.0, color=0.0, speed=0.5, variations=None, **kw):
    x = Xform(weight=weight, color=color, color_speed=speed, affine=Affine(a, b, c, d, e, f),
              variations=variations if variations is not None else [vr("linear", 1)])
    for k, v in kw.items():
        setattr(x, k, v)
    return x


def pal(name="p", rgb0=(0.0, 0.0, 0.0), rgb1=(1.0, 0.5, 0.25)):
    return Palette(name, [ColorStop(0.0, *rgb0), ColorStop(1.0, *rgb1)])


def gen(time=0.0, **kw):
    g = Genome(name="k", xforms=[xf()], scale=100.0, cx=0.0, cy=0.0, palette=pal(), time=time)
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def anim(k0, k1, **kw):
    return A.Animation(keyframes=[k0, k1], **kw)


def rot(deg, animate=None):
    th = math.radians(deg)
    return xf(math.cos(th), -math.sin(th), 0, math.sin(th), math.cos(th), 0, animate=animate)


# ------------------------------------------------------------ keyframes --

def test_pick_keyframes():
    ks = [gen(0), gen(1), gen(3)]
    with pytest.raises(ValueError):
        A.pick_keyframes(ks[:1], 0)
    assert A.pick_keyframes(ks, -1)[:2] == (0, 1)
    i1, i2, c0, c1 = A.pick_keyframes(ks, -1)
    assert c0 == pytest.approx(2.0) and c1 == pytest.approx(-1.0)     # extrapolates
    assert A.pick_keyframes(ks, 5)[:2] == (1, 2)
    assert A.pick_keyframes(ks, 0.5)[2:] == (0.5, 0.5)
    assert A.pick_keyframes(ks, 0)[2] == 1.0
    assert A.pick_keyframes(ks, 2)[:2] == (1, 2)
    assert A.pick_keyframes(ks, 2)[2] == pytest.approx(0.5)
    # undefined time reads as 0; coincident keyframes give all weight to i2
    assert A.pick_keyframes([gen(None), gen(None)], 0)[2:] == (0.0, 1.0)


def test_scalar_fields_and_time():
    with pytest.raises(ValueError):
        A.interpolate(A.Animation([gen(0)]), 0)
    r = A.interpolate(anim(gen(0, cx=0.0, cy=2.0, scale=100.0, rotate=10.0),
                           gen(1, cx=4.0, cy=6.0, scale=300.0)), 0.5)
    assert (r.cx, r.cy, r.scale, r.time) == (2.0, 4.0, 200.0, 0.5)
    assert r.rotate == pytest.approx(5.0)
    assert A.interpolate(anim(gen(0), gen(1)), 0.5).rotate == 0.0


# ------------------------------------------------------------ alignment --

def test_alignment_pads_zero_weight_identity_and_final():
    r = A.interpolate(anim(gen(0, xforms=[xf()]), gen(1, xforms=[xf(), xf()])), 0.5)
    assert len(r.xforms) == 2
    assert r.xforms[1].weight == pytest.approx(0.5)      # 1 -> pad weight 0
    r0 = A.interpolate(anim(gen(0, xforms=[xf()]), gen(1, xforms=[xf(), xf()])), 0)
    assert r0.xforms[1].weight == 0.0
    r = A.interpolate(anim(gen(0, finalxform=xf(c=2.0)), gen(1)), 0.5)
    assert r.finalxform is not None and r.finalxform.affine.c == pytest.approx(1.0)


@pytest.mark.parametrize("name", ["spherical", "polar", "julian", "ngon", "juliascope",
                                  "wedge_sph", "wedge_julia"])
def test_flipped_identity_pad_for_the_spherical_family(name):
    assert V[name] in (2, 5, 14, 39, 48, 89, 96)      # the reference's index list
    k1 = gen(1, xforms=[xf(), xf(variations=[vr(name, 1)])])
    r = A.interpolate(anim(gen(0), k1, interpolation_type="log"), 0)
    assert r.xforms[1].affine.a == pytest.approx(-1) and r.xforms[1].affine.e == pytest.approx(-1)


def test_flipped_identity_scans_every_slot_and_is_log_only():
    k1 = gen(1, xforms=[xf(), xf(variations=[vr("linear", 0.5), vr("spherical", 0.5)])])
    assert A.interpolate(anim(gen(0), k1, interpolation_type="log"), 0).xforms[1].affine.a \
        == pytest.approx(-1)
    assert A.interpolate(anim(gen(0), k1, interpolation_type="linear"), 0).xforms[1].affine.a \
        == pytest.approx(1)
    plain = gen(1, xforms=[xf(), xf()])
    assert A.interpolate(anim(gen(0), plain, interpolation_type="log"), 0).xforms[1].affine.a \
        == pytest.approx(1)


# ------------------------------------------------------------- affines --

def test_affine_linear_midpoint():
    k0 = gen(0, xforms=[xf(1, 0, 0, 0, 1, 0)])
    k1 = gen(1, xforms=[xf(3, 2, 4, -1, 5, 6)])
    a = A.interpolate(anim(k0, k1, interpolation_type="linear"), 0.5).xforms[0].affine
    assert (a.a, a.b, a.c, a.d, a.e, a.f) == pytest.approx((2, 1, 2, -0.5, 3, 3))


def test_affine_log_polar():
    r = A.interpolate(anim(gen(0, xforms=[rot(0)]), gen(1, xforms=[rot(90)])), 0.5)
    a = r.xforms[0].affine
    assert a.a == pytest.approx(math.cos(math.pi / 4), abs=1e-6)
    assert a.d == pytest.approx(math.sin(math.pi / 4), abs=1e-6)
    # a rotation stays a rotation: unit columns, orthogonal
    assert math.hypot(a.a, a.d) == pytest.approx(1.0) and math.hypot(a.b, a.e) == pytest.approx(1)
    assert a.a * a.b + a.d * a.e == pytest.approx(0.0, abs=1e-12)
    # short arc 350 -> 10 passes through 0
    assert A.interpolate(anim(gen(0, xforms=[rot(350)]), gen(1, xforms=[rot(10)])),
                         0.5).xforms[0].affine.a > 0.9
    # magnitude preserved (2 -> 2 through a 90 degree turn)
    r = A.interpolate(anim(gen(0, xforms=[xf(2, 0, 0, 0, 2, 0)]),
                           gen(1, xforms=[xf(0, -2, 0, 2, 0, 0)])), 0.5).xforms[0].affine
    assert math.hypot(r.a, r.d) == pytest.approx(2.0, abs=1e-6)
    # log magnitude: 1 -> 4 at the midpoint is 2 (geometric), not 2.5
    r = A.interpolate(anim(gen(0, xforms=[xf(1, 0, 0, 0, 1, 0)]),
                           gen(1, xforms=[xf(4, 0, 0, 0, 4, 0)])), 0.5).xforms[0].affine
    assert r.a == pytest.approx(2.0, rel=1e-8)


def test_zero_column_inherits_sibling_angle():
    k0 = gen(0, xforms=[xf(0, 0, 0, 1, 0, 0)])
    k1 = gen(1, xforms=[xf(0, 1, 0, 1, 0, 0)])
    a = A.interpolate(anim(k0, k1, interpolation_type="log"), 0.5).xforms[0].affine
    assert a.b == pytest.approx(0.35355, abs=1e-4) and a.e == pytest.approx(0.35355, abs=1e-4)


def test_asymmetric_wind():
    sym = A.interpolate(anim(gen(0, xforms=[rot(10, 1)]), gen(1, xforms=[rot(-10, 1)])), 0.5)
    assert sym.xforms[0].affine.a == pytest.approx(1, abs=1e-3)
    assert sym.xforms[0].affine.d == pytest.approx(0, abs=1e-3)
    asym = A.interpolate(anim(gen(0, xforms=[rot(10, 1)]), gen(1, xforms=[rot(-10, 0)])), 0.5)
    assert asym.xforms[0].affine.a == pytest.approx(-1, abs=1e-3)
    assert asym.xforms[0].affine.d == pytest.approx(0, abs=1e-3)
    fin = A.interpolate(anim(gen(0, finalxform=rot(10, 1)), gen(1, finalxform=rot(-10, 0))), 0.5)
    assert fin.finalxform.affine.a == pytest.approx(1, abs=1e-3)


def test_post_affine():
    post = lambda *v: Affine(*v)  # noqa: E731
    k0 = gen(0, xforms=[xf(post=post(2, 0, 1, 0, 2, 1))])
    k1 = gen(1, xforms=[xf(post=post(4, 0, 3, 0, 4, 3))])
    p = A.interpolate(anim(k0, k1, interpolation_type="linear"), 0.3).xforms[0].post
    assert (p.a, p.c, p.e, p.f) == pytest.approx((2.6, 1.6, 2.6, 1.6))
    ident = gen(0, xforms=[xf(post=Affine())])
    assert A.interpolate(anim(ident, gen(1, xforms=[xf(post=Affine())]),
                              interpolation_type="linear"), 0.3).xforms[0].post is None
    k0 = gen(0, xforms=[xf(post=post(2, 0, 4, 0, 2, 0))])
    p = A.interpolate(anim(k0, gen(1), interpolation_type="linear"), 0.3).xforms[0].post
    assert (p.a, p.c, p.e) == pytest.approx((1.7, 2.8, 1.7))
    assert A.interpolate(anim(k0, gen(1), interpolation_type="linear"), 1).xforms[0].post is None


# ------------------------------------------------------ xform contents --

def test_variations_union_params_and_fallback():
    r = A.interpolate(anim(gen(0, xforms=[xf(variations=[vr("linear", 1)])]),
                           gen(1, xforms=[xf(variations=[vr("linear", 3)])])), 0.5)
    v = r.xforms[0].variations[0]
    assert v.index == V["linear"] and v.weight == pytest.approx(2)
    r = A.interpolate(anim(gen(0, xforms=[xf(variations=[vr("linear", 1), vr("julian", 2, 3, 1)])]),
                           gen(1, xforms=[xf(variations=[vr("linear", 1)])])), 0.5)
    vs = r.xforms[0].variations
    assert [v.index for v in vs] == [V["linear"], V["julian"]]
    assert vs[1].weight == pytest.approx(1.0)
    r = A.interpolate(anim(gen(0, xforms=[xf(variations=[vr("julian", 1, 2, 1)])]),
                           gen(1, xforms=[xf(variations=[vr("julian", 1, 4, 3)])])), 0.5)
    v = r.xforms[0].variations[0]
    assert v.params[0] == pytest.approx(3) and v.params[1] == pytest.approx(2)
    r = A.interpolate(anim(gen(0, xforms=[xf(variations=[])]),
                           gen(1, xforms=[xf(variations=[])])), 0.5)
    assert [(v.index, v.weight) for v in r.xforms[0].variations] == [(V["linear"], 1.0)]
    # at most 8 in the union
    many0 = [vr(n, 1) for n in ("linear", "sinusoidal", "spherical", "swirl", "horseshoe")]
    many1 = [vr(n, 1) for n in ("polar", "handkerchief", "heart", "disc", "spiral")]
    r = A.interpolate(anim(gen(0, xforms=[xf(variations=many0)]),
                           gen(1, xforms=[xf(variations=many1)])), 0.5)
    assert len(r.xforms[0].variations) == 8


def test_active_flags():
    def off_dc():
        return xf(variations=[vr("linear", 1), Variation(V["dc_linear"], 0.8, active=False)])
    r = A.interpolate(anim(gen(0, xforms=[off_dc()]), gen(1, xforms=[off_dc()])), 0.5)
    dc = [v for v in r.xforms[0].variations if v.index == V["dc_linear"]][0]
    assert dc.active is False and dc.weight == 0.0
    off = xf(weight=0.7, active=False)
    r = A.interpolate(anim(gen(0, xforms=[xf(), off]), gen(1, xforms=[xf(), xf(weight=0.7,
                                                                              active=False)])), 0.5)
    assert r.xforms[1].active is False and r.xforms[1].weight == 0.0
    on = xf(variations=[vr("linear", 1), vr("dc_linear", 0.8)])
    r = A.interpolate(anim(gen(0, xforms=[off_dc()]), gen(1, xforms=[on])), 0.5)
    dc = [v for v in r.xforms[0].variations if v.index == V["dc_linear"]][0]
    assert dc.weight == pytest.approx(0.4) and dc.active is True


def test_xaos_and_opacity():
    r = A.interpolate(anim(gen(0, xforms=[xf(xaos=[0, 2, 4])]), gen(1, xforms=[xf(xaos=[10, 20])]),
                           interpolation_type="linear"), 0.3)
    assert r.xforms[0].xaos == pytest.approx([3.0, 7.4, 3.1])
    r = A.interpolate(anim(gen(0, xforms=[xf(xaos=[1, 1])]), gen(1, xforms=[xf(xaos=[1, 1])])), 0.3)
    assert r.xforms[0].xaos is None
    assert A.interpolate(anim(gen(0), gen(1)), 0.3).xforms[0].xaos is None
    r = A.interpolate(anim(gen(0, xforms=[xf(opacity=0.2)]), gen(1, xforms=[xf(opacity=0.7)]),
                           interpolation_type="linear"), 0.3)
    assert r.xforms[0].opacity == pytest.approx(0.35)
    assert A.interpolate(anim(gen(0), gen(1)), 0.3).xforms[0].opacity == 1.0


def test_colour_fields_clamped():
    r = A.interpolate(anim(gen(0, xforms=[xf(color=0.2, speed=0.4)]),
                           gen(1, xforms=[xf(color=1.0, speed=0.0)])), 2.0)   # extrapolated
    assert r.xforms[0].color == 1.0 and r.xforms[0].color_speed == 0.0


# --------------------------------------------------------------- palette --

def _solid(r, g, b):
    return Palette("s", [ColorStop(0, r, g, b), ColorStop(1, r, g, b)])


def test_palette_modes():
    r = A.interpolate(anim(gen(0), gen(1)), 0.5)
    assert len(r.palette.stops) == 256
    assert all(0 <= s.r <= 1 and 0 <= s.g <= 1 and 0 <= s.b <= 1 for s in r.palette.stops)
    assert r.palette.stops[0].r == pytest.approx(0.0, abs=1e-6)
    r = A.interpolate(anim(gen(0, palette=_solid(0, 0, 0)), gen(1, palette=_solid(1, 1, 1)),
                           palette_interpolation="rgb"), 0.5)
    assert all(s.r == pytest.approx(0.5) and s.b == pytest.approx(0.5) for s in r.palette.stops)

    def sweep(t):
        return A.interpolate(anim(gen(0, palette=_solid(1, 0, 0)), gen(1, palette=_solid(0, 0, 1)),
                                  palette_interpolation="sweep"), t).palette.stops
    s = sweep(0.3)     # c0 = 0.7: entries i < 179.2 from the first keyframe
    assert [s[i].r for i in (0, 178, 179, 180, 255)] == [1, 1, 1, 0, 0]
    s = sweep(0)       # c0 = 1: every entry from the first keyframe
    assert [s[i].r for i in (0, 255)] == [1, 1] and s[255].b == 0
    s = sweep(0.5)
    assert (s[127].r, s[128].r) == (1, 0)
    # hsv_circular takes the short hue arc red -> blue (through magenta)
    r = A.interpolate(anim(gen(0, palette=_solid(1, 0, 0)), gen(1, palette=_solid(0, 0, 1))), 0.3)
    assert r.palette.stops[0].r > 0.5
    r = A.interpolate(anim(gen(0, palette=_solid(1, 0, 0)), gen(1, palette=_solid(1, 0, 0))), 0.3)
    assert (r.palette.stops[0].r, r.palette.stops[0].g, r.palette.stops[0].b) == \
        pytest.approx((1, 0, 0), abs=1e-6)
    # hsv (no shortcut): red (h 0) -> blue (h 4) at 0.5 is h 2 = green
    r = A.interpolate(anim(gen(0, palette=_solid(1, 0, 0)), gen(1, palette=_solid(0, 0, 1)),
                           palette_interpolation="hsv"), 0.5)
    assert (r.palette.stops[0].r, r.palette.stops[0].g, r.palette.stops[0].b) == \
        pytest.approx((0, 1, 0), abs=1e-6)


def test_flam3_hsv_round_trip():
    for rgb in [(1, 0, 0), (0.2, 0.5, 0.9), (0.9, 0.9, 0.1), (0.3, 0.3, 0.3), (0, 0, 0)]:
        assert A._hsv2rgb(*A._rgb2hsv(*rgb)) == pytest.approx(rgb)


# --------------------------------------------------------------- grading --

def test_tonemap():
    k0 = gen(0, tonemap=Tonemap(gamma=2, vibrancy=0, highlight_power=-1, brightness=4,
                                gamma_threshold=0.01))
    k1 = gen(1, tonemap=Tonemap(gamma=4, vibrancy=1, highlight_power=1, brightness=8,
                                gamma_threshold=0.05))
    t = A.interpolate(anim(k0, k1), 0.5).tonemap
    assert (t.gamma, t.brightness, t.vibrancy, t.highlight_power) == pytest.approx((3, 6, 0.5, 0))
    assert A.interpolate(anim(gen(0), gen(1)), 0.5).tonemap is None
    # one side missing: the flam3 partial-fill values stand in
    k0 = gen(0, tonemap=Tonemap(gamma=2, vibrancy=0, highlight_power=1, brightness=8,
                                gamma_threshold=0.02))
    t = A.interpolate(anim(k0, gen(1)), 0.3).tonemap
    assert t.gamma == pytest.approx(0.7 * 2 + 0.3 * 4)
    assert t.brightness == pytest.approx(0.7 * 8 + 0.3 * 4)
    assert t.vibrancy == pytest.approx(0.3) and t.highlight_power == pytest.approx(0.4)


def test_hsl_adjust():
    h = lambda hue, sat=100, light=0: {"hue": hue, "sat": sat, "light": light}  # noqa: E731
    r = A.interpolate(anim(gen(0, hsl_adjust=h(0)), gen(1, hsl_adjust=h(60, 200, 50))), 0.5)
    assert (r.hsl_adjust["hue"], r.hsl_adjust["sat"], r.hsl_adjust["light"]) == \
        pytest.approx((30, 150, 25))
    r = A.interpolate(anim(gen(0), gen(1, hsl_adjust=h(60, 200, 50))), 0.5)
    assert (r.hsl_adjust["hue"], r.hsl_adjust["sat"]) == pytest.approx((30, 150))
    r = A.interpolate(anim(gen(0, hsl_adjust=h(170)), gen(1, hsl_adjust=h(-170))), 0.5)
    assert abs(r.hsl_adjust["hue"]) == pytest.approx(180)
    a, b = h(-40, 80, -30), h(50, 160, 40)
    ab = A.interpolate(anim(gen(0, hsl_adjust=a), gen(1, hsl_adjust=b)), 0.3).hsl_adjust
    ba = A.interpolate(anim(gen(0, hsl_adjust=b), gen(1, hsl_adjust=a)), 0.7).hsl_adjust
    assert [ab[k] for k in ("hue", "sat", "light")] == pytest.approx([ba[k] for k in
                                                                      ("hue", "sat", "light")])
    assert A.interpolate(anim(gen(0), gen(1)), 0.5).hsl_adjust is None


def test_channel_curves_blend():
    lifted = [(0.0, 0.0), (0.5, 0.8), (1.0, 1.0)]
    c1 = cv.identity_curves()
    c1["composite"] = lifted
    r = A.interpolate(anim(gen(0, channel_curves=cv.identity_curves()),
                           gen(1, channel_curves=c1)), 0.5)
    comp = dict(r.channel_curves["composite"])
    assert comp[0.5] == pytest.approx(0.65) and comp[0.0] == 0 and comp[1.0] == 1
    r = A.interpolate(anim(gen(0), gen(1, channel_curves=c1)), 0.5)
    assert dict(r.channel_curves["composite"])[0.5] == pytest.approx(0.65)
    ca, cb = cv.identity_curves(), cv.identity_curves()
    ca["composite"] = [(0, 0.1), (0.5, 0.3), (1, 0.9)]
    cb["composite"] = [(0, 0.0), (0.5, 0.7), (1, 1.0)]
    ab = dict(A.interpolate(anim(gen(0, channel_curves=ca), gen(1, channel_curves=cb)),
                            0.3).channel_curves["composite"])
    ba = dict(A.interpolate(anim(gen(0, channel_curves=cb), gen(1, channel_curves=ca)),
                            0.7).channel_curves["composite"])
    for x in (0, 0.5, 1):
        assert ab[x] == pytest.approx(ba[x])
    assert A.interpolate(anim(gen(0), gen(1)), 0.5).channel_curves is None


def test_render_fields():
    k0 = gen(0, density=Density(9, 0, 0.4), background=[0.1, 0.2, 0.3])
    r = A.interpolate(anim(k0, gen(1)), 0.5)
    assert r.density == k0.density and r.background == k0.background
    k0 = gen(0, quality=10.0, oversample=1, background=[0, 0, 0], size={"width": 100, "height": 100},
             filter_radius=1.0, density=Density(9, 0, 0.4))
    k1 = gen(1, quality=30.0, oversample=3, background=[1, 1, 1], size={"width": 200, "height": 300},
             filter_radius=3.0, density=Density(5, 0, 0.6))
    r = A.interpolate(anim(k0, k1), 0.5)
    assert r.quality == pytest.approx(20) and r.oversample == 2
    assert r.background == [0.5, 0.5, 0.5] and r.size == {"width": 150, "height": 200}
    assert r.filter_radius == pytest.approx(2)
    assert r.density.max_rad == pytest.approx(7) and r.density.curve == pytest.approx(0.5)


# ---------------------------------------------------- easing, permutation --

def test_segment_easing():
    k0, k1 = gen(0, cx=0.0), gen(1, cx=10.0)
    assert A.interpolate(anim(k0, k1), 0.5).cx == pytest.approx(5)
    lin = {"kind": "preset", "name": "linear"}
    ein = {"kind": "preset", "name": "easeIn"}
    assert A.interpolate(anim(k0, k1, segment_easing=[lin]), 0.5).cx == pytest.approx(5)
    assert A.interpolate(anim(k0, k1, segment_easing=[ein]), 0.5).cx == pytest.approx(2.5)
    assert A.interpolate(anim(k0, k1, segment_easing=[ein]), 1.5).cx == pytest.approx(15)
    assert A.interpolate(anim(k0, k1, segment_easing=[]), 0.5).cx == pytest.approx(5)


def _tagged(tags, time, **kw):
    return gen(time, xforms=[xf(c=float(t)) for t in tags], **kw)


def _cs(g):
    return [x.affine.c for x in g.xforms]


def test_segment_permutation():
    k0, k1 = _tagged([10, 20, 30], 0), _tagged([11, 21, 31], 1)
    assert _cs(A.interpolate(anim(k0, k1), 0.5)) == [10.5, 20.5, 30.5]
    assert _cs(A.interpolate(anim(k0, k1, segment_permutation=[[0, 1, 2]]), 0.5)) == \
        [10.5, 20.5, 30.5]
    assert _cs(A.interpolate(anim(k0, k1, segment_permutation=[[2, 1, 0]]), 0.5)) == \
        [20.5, 20.5, 20.5]
    z = _tagged([0, 0, 0], 0)
    assert _cs(A.interpolate(anim(z, _tagged([10, 20, 30], 1), segment_permutation=[[1, 2, 0]]),
                             0.5)) == [10, 15, 5]
    for bad in ([], [[0, 1]], [[0, 0, 1]], [[0, 1, 5]]):
        assert _cs(A.interpolate(anim(k0, k1, segment_permutation=bad), 0.5)) == \
            [10.5, 20.5, 30.5]
    # a real xform mapped onto a padded slot fades out
    r = A.interpolate(anim(_tagged([10, 20], 0), _tagged([99], 1), segment_permutation=[[1, 0]]),
                      0.5)
    assert r.xforms[0].weight == pytest.approx(0.5) and r.xforms[1].affine.c == pytest.approx(59.5)
    # the final xform is not permuted
    r = A.interpolate(anim(_tagged([10, 20], 0, finalxform=xf(c=5.0)),
                           _tagged([11, 21], 1, finalxform=xf(c=7.0)),
                           segment_permutation=[[1, 0]]), 0.5)
    assert r.finalxform.affine.c == pytest.approx(6)
    # a short permutation on a symmetry-baked flame gets an identity tail
    sym = {"kind": "rotational", "n": 4}
    pos = A.interpolate(anim(_tagged([10, 20], 0, symmetry=sym), _tagged([11, 21], 1, symmetry=sym)),
                        0.5)
    sw = A.interpolate(anim(_tagged([10, 20], 0, symmetry=sym), _tagged([11, 21], 1, symmetry=sym),
                            segment_permutation=[[1, 0]]), 0.5)
    assert len(pos.xforms) > 2
    assert pos.xforms[0].affine.c == pytest.approx(10.5)
    assert (sw.xforms[0].affine.c, sw.xforms[1].affine.c) == pytest.approx((15.5, 15.5))
    assert A.resolve_segment_permutation([1, 0], 4) == [1, 0, 2, 3]
    assert A.resolve_segment_permutation([1, 1], 4) is None


def test_symmetry_baked_both_directions():
    plain = lambda t: gen(t)  # noqa: E731
    symm = lambda t: gen(t, symmetry={"kind": "rotational", "n": 4})  # noqa: E731
    ab = A.interpolate(anim(plain(0), symm(1)), 0.5)
    ba = A.interpolate(anim(symm(0), plain(1)), 0.5)
    assert len(ab.xforms) == len(ba.xforms) == 4 and ab.symmetry is None and ba.symmetry is None


# -------------------------------------------------------- Catmull-Rom --

def test_catmull_rom():
    for t in (0, 0.25, 0.5, 0.75, 1):
        assert sum(A.catmull_rom_weights(t)) == pytest.approx(1)
    assert A.catmull_rom_weights(0) == (0, 1, 0, 0)
    assert A.catmull_rom_weights(1) == (0, 0, 1, 0)
    # the basis reproduces quadratics exactly: keyframes at k^2 give (1 + t)^2
    for t in (0.1, 0.5, 0.8):
        w = A.catmull_rom_weights(t)
        assert sum(wk * (k * k) for k, wk in enumerate(w)) == pytest.approx((1 + t) ** 2)
    ks = [gen(0, cx=0.0), gen(1, cx=10.0), gen(2, cx=40.0), gen(3, cx=90.0)]
    sm = A.Animation(keyframes=ks, interpolation="smooth")
    assert A.interpolate(sm, 1).cx == pytest.approx(10)
    assert A.interpolate(sm, 2).cx == pytest.approx(40)
    assert A.interpolate(sm, 0.5).cx == pytest.approx(5)       # first segment: linear
    assert A.interpolate(sm, 2.5).cx == pytest.approx(65)      # last segment: linear
    assert A.interpolate(A.Animation(keyframes=ks), 1.5).cx == pytest.approx(25)
    assert A.interpolate(sm, 1.5).cx == pytest.approx(22.5)


# ------------------------------------------------------------- motion --

def _mel(func, freq=1.0, a=0.0, **kw):
    m = xf(a, 0, 0, 0, 0, 0, weight=0.0, color=0.0, speed=0.0, variations=[],
           motion_func=func, motion_freq=freq)
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def test_motion_funcs():
    for t, v in ((0, 0), (0.25, 1), (0.5, 0), (0.75, -1), (1, 0)):
        assert A.motion_funcs(1, t) == pytest.approx(v, abs=1e-12)
        assert A.motion_funcs(2, t) == pytest.approx(v, abs=1e-12)
    assert A.motion_funcs(2, -0.75) == pytest.approx(A.motion_funcs(2, 0.25))
    assert A.motion_funcs(3, 0) == 0 and A.motion_funcs(3, 0.5) == pytest.approx(1)
    assert all(0 <= A.motion_funcs(3, t / 20) <= 1 for t in range(20))


def test_apply_motion_parameters():
    x = xf()
    assert A.apply_motion_parameters(x, 0.5) is x
    x = xf(motion=[_mel(1, a=10.0)])
    snap = genome_to_json(gen(xforms=[x]))
    assert A.apply_motion_parameters(x, 0).affine.a == pytest.approx(1)
    assert A.apply_motion_parameters(x, 0.25).affine.a == pytest.approx(11)
    assert genome_to_json(gen(xforms=[x])) == snap                  # input untouched
    assert A.apply_motion_parameters(xf(motion=[_mel(3, a=5.0)]), 0.5).affine.a == pytest.approx(6)
    assert A.apply_motion_parameters(xf(motion=[_mel(1, 2.0, a=10.0)]), 0.25).affine.a == \
        pytest.approx(1)
    assert A.apply_motion_parameters(xf(motion=[_mel(1, a=5.0), _mel(1, a=3.0)]),
                                     0.25).affine.a == pytest.approx(9)
    assert A.apply_motion_parameters(xf(motion=[_mel(0, a=10.0)]), 0.25).affine.a == 1
    assert A.apply_motion_parameters(xf(color=0.5, motion=[_mel(1, color=100.0)]), 0.25).color == 1
    assert A.apply_motion_parameters(xf(weight=0.5, motion=[_mel(1, weight=-100.0)]),
                                     0.25).weight == 0
    r = A.apply_motion_parameters(xf(motion=[_mel(1, variations=[vr("linear", 2)])]), 0.25)
    assert r.variations[0].weight == pytest.approx(3)
    r = A.apply_motion_parameters(xf(motion=[_mel(1, variations=[vr("julian", 2, 3)])]), 0.25)
    assert [v.index for v in r.variations] == [V["linear"], V["julian"]]
    assert r.variations[1].weight == pytest.approx(2) and r.variations[1].params[0] == \
        pytest.approx(3)
    full = xf(variations=[vr(n, 1) for n in ("linear", "sinusoidal", "spherical", "swirl",
                                             "horseshoe", "polar", "handkerchief", "heart")],
              motion=[_mel(1, variations=[vr("julian", 2)])])
    assert len(A.apply_motion_parameters(full, 0.25).variations) == 8
    # opacity moves only when the motion element gives one
    m = _mel(1)
    m.opacity = 0.5
    assert A.apply_motion_parameters(xf(opacity=0.8, motion=[m]), 0.25).opacity == 0.8
    m.opacity_given = True
    assert A.apply_motion_parameters(xf(opacity=0.8, motion=[m]), 0.25).opacity == \
        pytest.approx(1.3)
    # post: added to the identity when the xform has none
    m = _mel(1, post=Affine(1, 0, 2, 0, 1, 0))
    r = A.apply_motion_parameters(xf(motion=[m]), 0.25)
    assert (r.post.a, r.post.c) == pytest.approx((2, 2))


def test_motion_runs_on_the_raw_segment_clock():
    x0 = xf(motion=[_mel(1, a=10.0)])
    k0, k1 = gen(0, xforms=[x0]), gen(1, xforms=[xf()])
    ein = {"kind": "preset", "name": "easeIn"}
    # t = 0.25: the motion clock is 0.25 (peak) on k0 whatever the easing;
    # k0's a = 11, k1's a = 1 -> linear blend with c1 = 0.25 or eased 0.0625
    r = A.interpolate(anim(k0, k1, interpolation_type="linear"), 0.25)
    assert r.xforms[0].affine.a == pytest.approx(0.75 * 11 + 0.25 * 1)
    r = A.interpolate(anim(k0, k1, interpolation_type="linear", segment_easing=[ein]), 0.25)
    assert r.xforms[0].affine.a == pytest.approx(0.9375 * 11 + 0.0625 * 1)


# ------------------------------------------------------------- easing --

def test_eval_easing():
    p = lambda n: {"kind": "preset", "name": n}  # noqa: E731
    assert [A.eval_easing(p("linear"), t) for t in (0, 0.5, 1)] == [0, 0.5, 1]
    assert A.eval_easing(p("easeIn"), 0.5) == pytest.approx(0.25)
    assert A.eval_easing(p("easeOut"), 0.5) == pytest.approx(0.75)
    assert A.eval_easing(p("easeInOut"), 0.25) == pytest.approx(0.15625)
    assert A.eval_easing(p("easeInOut"), 0.25) + A.eval_easing(p("easeInOut"), 0.75) == \
        pytest.approx(1)
    assert (A.eval_easing(p("hold"), 0), A.eval_easing(p("hold"), 0.99),
            A.eval_easing(p("hold"), 1)) == (0, 0, 1)
    assert (A.eval_easing(p("linear"), -0.5), A.eval_easing(p("linear"), 1.5)) == (0, 1)
    bz = lambda *c: {"kind": "cubicBezier", "x1": c[0], "y1": c[1], "x2": c[2], "y2": c[3]}  # noqa
    assert A.eval_easing(bz(0, 0, 1, 1), 0.3) == pytest.approx(0.3, abs=1e-3)
    assert A.eval_easing(bz(0.42, 0, 1, 1), 0.5) < 0.5
    assert A.eval_easing(bz(0.42, 0, 0.58, 1), 0.5) == pytest.approx(0.5, abs=1e-3)
    ys = [A.eval_easing(bz(0.25, 0.1, 0.25, 1), t / 20) for t in range(21)]
    assert all(b >= a for a, b in zip(ys, ys[1:]))
    assert all(0 <= A.eval_easing(bz(0.5, -1, 0.5, 2), t / 10) <= 1 for t in range(11))
    assert A.eval_easing({"kind": "spring"}, 0.3) == 0.3
    # the solved bezier point lies on the curve: x(s) = t within the solver's tolerance
    c = bz(0.3, 0.2, 0.7, 0.9)
    for t in (0.1, 0.4, 0.9):
        s = A._solve_bezier(0.3, 0.7, t)
        assert A._bezier(0.3, 0.7, s) == pytest.approx(t, abs=2e-6)
        assert A.eval_easing(c, t) == pytest.approx(A._bezier(0.2, 0.9, s))


# ---------------------------------------------------- temporal filter --

def test_temporal_filter():
    f = A.temporal_filter(1)
    assert (f.deltas, f.filter, f.sumfilt) == ([0.0], [1.0], 1.0)
    f = A.temporal_filter(4, "box", 1.0)
    assert f.deltas[0] == pytest.approx(-0.5) and f.deltas[3] == pytest.approx(0.5)
    assert f.deltas[1] - f.deltas[0] == pytest.approx(1 / 3)
    assert f.filter == [1, 1, 1, 1] and f.sumfilt == pytest.approx(1)
    f = A.temporal_filter(4, "box", 2.0)
    assert (f.deltas[0], f.deltas[3]) == pytest.approx((-1, 1))
    g = A.temporal_filter(11, "gaussian")
    assert g.filter[5] > g.filter[0] and max(g.filter) == pytest.approx(1)
    assert 0 < g.sumfilt < 1
    # closed form: weight i is exp(-2 x_i^2 + 2 x_max^2), x_i = 1.5 |i - N/2| / (N/2)
    xs = [1.5 * abs(i - 5.5) / 5.5 for i in range(11)]
    xmin = min(xs)
    assert g.filter == pytest.approx([math.exp(-2 * (x * x - xmin * xmin)) for x in xs])
    e = A.temporal_filter(10, "exp", 1.0, 2.0)
    assert e.filter[9] == pytest.approx(1) and e.filter[0] < e.filter[9]
    assert e.filter == pytest.approx([((i + 1) / 10) ** 2 for i in range(10)])
    e = A.temporal_filter(10, "exp", 1.0, -2.0)
    assert e.filter[0] == pytest.approx(1) and e.filter[9] < e.filter[0]
    assert A.temporal_filter(10, "exp", 1.0, 0.0).filter == pytest.approx([1] * 10)
    for n, kind in ((7, "box"), (11, "gaussian"), (9, "exp")):
        f = A.temporal_filter(n, kind, 1.0, 1.5)
        assert sum(w / (n * f.sumfilt) for w in f.filter) == pytest.approx(1, abs=1e-9)


def test_apportion_largest_remainder():
    from pyr3_amd.batch import apportion
    assert apportion(10, [1, 1, 1]) == [4, 3, 3]
    assert apportion(7, [0.1, 0.0, 1.0]) == [1, 0, 6]
    assert apportion(5, [1, -1, 1]) == [3, 0, 2]
    assert sum(apportion(131072, A.temporal_filter(1000, "gaussian").filter)) == 131072
    assert apportion(3, [0, 0]) == [0, 0] and apportion(0, [1]) == [0]


# ------------------------------------------------------- channel curves --

def test_channel_curves_eval_and_bake():
    ident = cv.IDENTITY_POINTS
    assert cv.is_identity(ident) and cv.active_mask(None) == 0
    assert cv.active_mask(cv.identity_curves()) == 0 and cv.bake_curves(cv.identity_curves()) is None
    c = cv.identity_curves()
    c["g"] = [(0, 0), (0.5, 0.7), (1, 1)]
    c["luma"] = [(0, 0.2), (1, 0.8)]
    assert cv.active_mask(c) == (1 << 2) | (1 << 4)
    lut = cv.bake_curves(c)
    assert lut.shape == (5, 256)
    assert lut[0, 128] == pytest.approx(128 / 255) and lut[4, 0] == pytest.approx(0.2)
    # spline through its control points, monotone between them, clamped at the ends
    pts = [(0.1, 0.2), (0.4, 0.6), (0.7, 0.65), (0.9, 0.95)]
    for x, y in pts:
        assert cv.eval_curve(pts, x) == pytest.approx(y)
    assert cv.eval_curve(pts, 0.0) == 0.2 and cv.eval_curve(pts, 1.0) == 0.95
    # tension 0.5 Catmull-Rom on equal spacing: the middle of segment 1 is
    # (-y0 + 9 y1 + 9 y2 - y3) / 16
    eq = [(0, 0.1), (1 / 3, 0.5), (2 / 3, 0.4), (1, 0.9)]
    assert cv.eval_curve(eq, 0.5) == pytest.approx((-0.1 + 9 * 0.5 + 9 * 0.4 - 0.9) / 16)
    assert cv.eval_curve([(0, 0.2), (1, 0.8)], 0.25) == pytest.approx(0.35)
    for bad in ([(0, 0)], [(0, 0), (0, 1)], [(0, 0), (1, 1.5)], [(i / 9, 0) for i in range(9)]):
        with pytest.raises(ValueError):
            cv.validate(bad)


# ------------------------------------------------- model and import --

def test_genome_json_round_trips_animation_and_grading_fields():
    m = _mel(2, 3.0, a=0.5)
    m.animate = 0.0
    x = xf(motion=[m], motion_freq=2.0, motion_func=1, animate=0.0)
    c = cv.identity_curves()
    c["r"] = [(0.0, 0.1), (1.0, 0.9)]
    g = gen(2.5, xforms=[x], channel_curves=c, hsl_adjust={"hue": 10.0, "sat": 90.0, "light": 5.0},
            nick="someone")
    back = genome_from_json(genome_to_json(g))
    assert back.time == 2.5 and back.nick == "someone"
    assert back.hsl_adjust == g.hsl_adjust and back.channel_curves["r"] == [(0.0, 0.1), (1.0, 0.9)]
    bx = back.xforms[0]
    assert (bx.motion_freq, bx.motion_func, bx.animate) == (2.0, 1, 0.0)
    assert bx.motion[0].motion_func == 2 and bx.motion[0].affine.a == 0.5
    j = genome_to_json(g)
    j["hslAdjust"] = {"hue": 350, "sat": 500, "light": -400}
    h = genome_from_json(j).hsl_adjust
    assert h == {"hue": -10.0, "sat": 200.0, "light": -100.0}
    j["channelCurves"]["g"] = [{"x": 0.5, "y": 0}, {"x": 0.4, "y": 1}]
    with pytest.raises(ValueError):
        genome_from_json(j)


MOTION_FLAME = """
<flames>
<flame name="a" time="0" size="64 64" center="0 0" scale="20" temporal_samples="6"
       temporal_filter_type="gaussian" interpolation_type="linear" palette_interpolation="rgb">
  <xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1" animate="0"
         motion_frequency="2" motion_function="triangle">
    <motion motion_frequency="1" motion_function="sin" julian="0.5" julian_power="3"
            coefs="0.1 0 0 0 0 0"/>
    <motion motion_frequency="1" motion_function="hill" opacity="0.25"/>
    <motion motion_function="hill" weight="5"/>
  </xform>
  <color index="0" rgb="255 0 0"/><color index="255" rgb="0 0 255"/>
</flame>
<flame name="b" time="10" size="64 64" center="1 0" scale="40">
  <xform weight="1" color="1" coefs="0 1 -1 0 0 0" linear="1"/>
  <color index="0" rgb="0 255 0"/><color index="255" rgb="0 0 255"/>
</flame>
</flames>
"""


def test_flam3_motion_import_and_animation():
    from pyr3_amd.flam3 import parse_flame
    imp = parse_flame(MOTION_FLAME)
    x = imp.genome.xforms[0]
    assert (x.animate, x.motion_freq, x.motion_func) == (0.0, 2.0, 2)
    m0, m1, m2 = x.motion
    assert m0.motion_func == 1 and m0.motion_freq == 1.0
    assert [(v.index, v.weight) for v in m0.variations] == [(V["julian"], 0.5)]
    assert m0.variations[0].params[0] == 3.0
    assert (m0.affine.a, m0.affine.e) == (0.1, 0.0)          # coefs given: 0.1 0 0 0 0 0
    assert m1.variations == [] and m1.affine == Affine(0, 0, 0, 0, 0, 0)   # no coefs: zero deltas
    assert m1.opacity_given and m1.opacity == 0.25 and not m0.opacity_given
    assert m1.weight == 1.0 and m1.color == 0.0 and m1.color_speed == 0.5
    assert not any("motion" in f["field"] or "animate" in f["field"]
                   for f in imp.report.ignored_fields)
    an = A.Animation.from_flame(imp)
    assert [k.time for k in an.keyframes] == [None, 10.0]
    assert (an.ntemporal_samples, an.temporal_filter_type, an.interpolation_type,
            an.palette_interpolation) == (6.0, "gaussian", "linear", "rgb")
    mid = A.interpolate(an, 5.0)
    assert mid.cx == pytest.approx(0.5) and mid.scale == pytest.approx(30)
    # motion on k0 at raw clock 0.5: sin(2 pi 0.5) = 0 and hill(0.5) = 1 ->
    # weight 1 + 1 (the hill element's default weight 1), opacity 1 + 0.25;
    # the element without a frequency runs at frequency 0: hill(0) = 0
    assert m2.motion_freq is None and m2.weight == 5.0
    assert mid.xforms[0].weight == pytest.approx(0.5 * 2 + 0.5 * 1)
    assert mid.xforms[0].opacity == pytest.approx(0.5 * 1.25 + 0.5 * 1)


# --------------------------------------------------- frame loop plumbing --

class _Recorder:
    """Stands in for pyr3_amd.renderer.Renderer: records the calls the frame
    loop makes (no device)."""

    def __init__(self, w, h):
        self.width, self.height = w, h
        self.calls = []

    def reset(self, g):
        self.calls.append(("reset", g))

    def iterate(self, g, seed, walkers, iters, jitter=None, fuse=None, packed=None):
        self.calls.append(("iterate", [g], [seed], [walkers], iters))

    def iterate_multi(self, gs, seeds, walkers, iters, jitter=None, fuse=None, packed=None):
        assert len({A.packed_shape(p) for p in packed}) == 1 and len(gs) <= 64
        self.calls.append(("iterate", list(gs), list(seeds), list(walkers), iters))

    def present(self, g, samples, force_de_off=False, transparent=False):
        self.calls.append(("present", g, samples))


def test_frame_loop_subframes_seeds_walkers_and_groups():
    k0 = gen(0.0, xforms=[xf(), xf(c=0.5)], quality=20.0)
    k1 = gen(1.0, xforms=[xf(c=0.1), xf(c=0.4)], quality=20.0)
    k2 = gen(2.0, xforms=[xf(), xf(variations=[vr("spherical", 1)]), xf()], quality=20.0)
    an = A.Animation([k0, k1, k2], ntemporal_samples=200, temporal_filter_type="gaussian",
                     temporal_filter_width=1.0)
    rec = _Recorder(64, 48)
    res = A.render_animation_frame(rec, an, 1.0, seed=0xFFFFFFF0, walkers=50000)
    kinds = [c[0] for c in rec.calls]
    assert kinds[0] == "reset" and kinds[-1] == "present" and set(kinds[1:-1]) == {"iterate"}
    launched = [(g, s, w) for c in rec.calls if c[0] == "iterate" for g, s, w in zip(*c[1:4])]
    assert [s for _, s, _ in launched] == res.seeds and [w for _, _, w in launched] == res.walkers
    tf = A.temporal_filter(200, "gaussian", 1.0)
    from pyr3_amd.batch import apportion
    counts = apportion(50000, tf.filter)
    live = [i for i in range(200) if counts[i]]
    assert res.walkers == [counts[i] for i in live] and sum(res.walkers) == 50000
    assert res.seeds == [(0xFFFFFFF0 + i * 0x9E3779B1) & 0xFFFFFFFF for i in live]
    assert [g.time for g in res.genomes] == pytest.approx([1.0 + tf.deltas[i] for i in live])
    # two shapes (either side of the middle keyframe), each launched in runs of <= 64
    n_launch = sum(1 for k in kinds if k == "iterate")
    assert len({A.packed_shape(_pg(g)) for g in res.genomes}) == 2
    assert n_launch >= 4 and n_launch <= len(live)
    _, center, samples = rec.calls[-1]
    iters = rec.calls[1][4]
    assert samples == 50000 * iters == res.total_samples and center.time == 1.0
    assert iters == -(-round(20.0 * 64 * 48) // 50000)
    # one temporal sample: the centre genome, the base seed, every walker
    rec = _Recorder(64, 48)
    res = A.render_animation_frame(rec, A.Animation([k0, k1], ntemporal_samples=1), 0.5, seed=9,
                                   walkers=4096)
    assert [c[0] for c in rec.calls] == ["reset", "iterate", "present"]
    assert rec.calls[1][2:4] == ([9], [4096]) and res.genomes[0].time == 0.5


def _pg(g):
    from pyr3_amd.renderer import pack_genome
    return pack_genome(g)
