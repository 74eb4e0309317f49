"""`.flam3` import (pyr3_amd/flam3.py) against the reference importer's own
expectations (src/flame-import.test.ts), restated: coefs layout, variation
attributes and their parameters, colour / colour-speed defaults, palettes in
the three inline formats, tone-map partial fill, density, symmetry, the
malformed-value recoveries and the errors; plus every golden fixture's
source flame parses with nothing dropped."""
import json
import math

import pytest

from conftest import GOLDEN
from pyr3_amd import catalog as cat
from pyr3_amd.flam3 import coefs_to_affine, load_genome, parse_flame
from pyr3_amd.genome import MAX_XFORMS, Affine, GenomeError

MIN_PALETTE = '<color index="0" rgb="0 0 0"/><color index="255" rgb="255 255 255"/>'
ONE_XFORM = '<xform weight="1" color="0" coefs="1 0 0 1 0 0" linear="1"/>'
MIN_XFORM = '<xform weight="1" color="0" color_speed="0.5" coefs="1 0 0 1 0 0" linear="1"/>'
V = cat.VARIATION_INDEX


def wrap(children, attrs="", palette=MIN_PALETTE):
    extra = " " + attrs if attrs else ""
    return (f'<flame name="t" size="1024 1024" center="0 0" scale="100"{extra}>'
            f'{palette}{children}</flame>')


def test_malformed_and_missing_flame():
    with pytest.raises(GenomeError, match="malformed"):
        parse_flame("<flame oops")
    with pytest.raises(GenomeError, match="no <flame> element"):
        parse_flame("<root><other /></root>")


def test_name_is_inert_text():
    xml = ('<flame name="&lt;img src=x onerror=alert(1)&gt;" size="1024 1024" center="0 0" '
           f'scale="100">{MIN_PALETTE}{ONE_XFORM}</flame>')
    assert parse_flame(xml).genome.name == "<img src=x onerror=alert(1)>"


def test_xform_cap():
    r = parse_flame(wrap(ONE_XFORM * (MAX_XFORMS + 5)))
    assert len(r.genome.xforms) == MAX_XFORMS
    assert r.report.clamped_xforms == {"had": MAX_XFORMS + 5, "cap": MAX_XFORMS}
    r = parse_flame(wrap(ONE_XFORM * MAX_XFORMS))
    assert len(r.genome.xforms) == MAX_XFORMS and r.report.clamped_xforms is None


def test_coefs_are_column_major():
    assert coefs_to_affine("1 2 3 4 5 6") == Affine(a=1, b=3, c=5, d=2, e=4, f=6)
    assert coefs_to_affine("  1\t2\n3 4   5 6 ") == Affine(a=1, b=3, c=5, d=2, e=4, f=6)
    with pytest.raises(GenomeError):
        coefs_to_affine("1 2 3 4 5")
    with pytest.raises(GenomeError):
        coefs_to_affine("1 2 3 4 5 nan")


def test_variation_params_and_defaults():
    x = parse_flame(wrap('<xform weight="0.5" color="0.25" coefs="1 0 0 1 0 0" '
                         'julian="0.7" julian_power="5" julian_dist="-1.5"/>')).genome.xforms[0]
    assert x.weight == 0.5 and x.color == 0.25 and x.color_speed == 0.5
    (v,) = x.variations
    assert v.index == V["julian"] and v.weight == 0.7 and v.params[:2] == [5.0, -1.5]
    # underscore-named variations are variations, with their own parameters
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" radial_blur="0.3" '
                         'radial_blur_angle="0.25" gaussian_blur="0.1"/>')).genome.xforms[0]
    assert [(v.index, v.weight) for v in x.variations] == [(V["radial_blur"], 0.3),
                                                           (V["gaussian_blur"], 0.1)]
    assert x.variations[0].params[0] == 0.25
    # a missing parameter takes the variation's default
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" blob="1" blob_high="3"/>')
                    ).genome.xforms[0]
    low, _, waves = cat.VARIATION_DEFAULTS["blob"][:3]
    assert x.variations[0].params[:3] == [float(low), 3.0, float(waves)]


def test_aliases():
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" mobius="1" Re_A="2" Im_D="-3" '
                         'oscilloscope="0.5" oscope_frequency="7"/>')).genome.xforms[0]
    mob, osc = x.variations
    assert mob.index == V["mobius"] and osc.index == V["oscilloscope"]
    names = cat.VARIATION_PARAMS["mobius"]
    assert mob.params[names.index("re_a")] == 2.0 and mob.params[names.index("im_d")] == -3.0
    assert osc.params[cat.VARIATION_PARAMS["oscilloscope"].index("frequency")] == 7.0


def test_dropped_and_fallback_linear():
    r = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" linear="0.5" '
                         'frobnicate="0.25" mystery_thing="x"/>'))
    assert [v.index for v in r.genome.xforms[0].variations] == [V["linear"]]
    dropped = {d["name"]: d["weight"] for d in r.report.dropped_variations}
    assert dropped == {"frobnicate": 0.25, "mystery_thing": 0.0}
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" frobnicate="1"/>')).genome.xforms[0]
    assert [(v.index, v.weight) for v in x.variations] == [(V["linear"], 1.0)]
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0"/>')).genome.xforms[0]
    assert [(v.index, v.weight) for v in x.variations] == [(V["linear"], 1.0)]
    # explicit zero weights are kept
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" linear="0" julia="0"/>')
                    ).genome.xforms[0]
    assert [(v.index, v.weight) for v in x.variations] == [(V["linear"], 0.0), (V["julia"], 0.0)]
    r = parse_flame(wrap(ONE_XFORM + '<finalxform coefs="1 0 0 1 0 0" frobnicate="1"/>'))
    assert r.report.dropped_variations[0].get("final") is True


def test_malformed_scalars_default():
    r = parse_flame(wrap('<xform weight="oops" color="0 1" color_speed="nan" '
                         'coefs="1 0 0 1 0 0" linear="1"/>'))
    x = r.genome.xforms[0]
    assert (x.weight, x.color, x.color_speed) == (1.0, 0.0, 0.0)
    fields = {f["field"]: f["value"] for f in r.report.ignored_fields}
    assert fields["color@xform[0]"] == "0 1" and "weight@xform[0]" in fields
    r = parse_flame(f'<flame name="t" scale="nan" center="nan nan">{MIN_PALETTE}{ONE_XFORM}</flame>')
    assert r.genome.scale == 50 and (r.genome.cx, r.genome.cy) == (0.0, 0.0)
    assert {d["field"] for d in r.report.defaulted_fields} == {"scale", "center"}


def test_colour_defaults():
    xs = '<xform weight="1" coefs="1 0 0 1 0 0" linear="1"/>'
    g = parse_flame(wrap(xs * 3 + '<finalxform coefs="1 0 0 1 0 0" linear="1"/>')).genome
    assert [x.color for x in g.xforms] == [0, 1, 0] and g.finalxform.color == 1
    assert g.finalxform.weight == 0
    g = parse_flame(wrap(xs * 2 + '<finalxform coefs="1 0 0 1 0 0" linear="1"/>')).genome
    assert g.finalxform.color == 0
    g = parse_flame(wrap('<xform weight="1" color="0.75" symmetry="-1" coefs="1 0 0 1 0 0"/>'
                         '<xform weight="1" symmetry="0.5" color_speed="0.2" coefs="1 0 0 1 0 0"/>')
                    ).genome
    assert g.xforms[0].color == 0.75 and g.xforms[0].color_speed == 1.0   # (1 - s) / 2
    assert g.xforms[1].color_speed == 0.2                                # explicit wins


def test_palettes():
    pal = "".join(f'<color index="{i}" rgb="{i} {i} {i}"/>' for i in range(256))
    s = parse_flame(wrap(MIN_XFORM, palette=pal)).genome.palette.stops
    assert len(s) == 256 and (s[0].t, s[0].r) == (0.0, 0.0) and (s[255].t, s[255].b) == (1.0, 1.0)
    assert math.isclose(s[128].r, 128 / 255)
    s = parse_flame(wrap(MIN_XFORM, palette='<color index="0" rgb="255 0 0"/>'
                                            '<color index="255" rgb="0 0 255"/>')).genome.palette.stops
    assert (s[0].r, s[0].b) == (1.0, 0.0) and (s[255].r, s[255].b) == (0.0, 1.0)
    assert math.isclose(s[100].r, 155 / 255) and math.isclose(s[100].b, 100 / 255)
    s = parse_flame(wrap(MIN_XFORM, palette='<color index="0" rgba="255 128 64 128"/>'
                                            '<color index="1" rgb="255 128 64" a="128"/>')
                    ).genome.palette.stops
    for k in (0, 1):
        assert math.isclose(s[k].r, 128 / 255) and math.isclose(s[k].g, (128 / 255) ** 2)
    hexd = "".join("00" + f"{i:02x}" * 3 for i in range(256))
    s = parse_flame(wrap(MIN_XFORM, palette=f'<colors count="256" data="{hexd}"/>')).genome.palette.stops
    assert math.isclose(s[128].g, 128 / 255)
    s = parse_flame(wrap(MIN_XFORM, palette='<colors count="256" data="' + "0080a040" * 256 + '"/>')
                    ).genome.palette.stops
    assert (s[0].r, s[0].g, s[0].b) == (128 / 255, 160 / 255, 64 / 255)
    hexp = "".join(f"{i:02x}" * 3 for i in range(256))
    s = parse_flame(wrap(MIN_XFORM, palette=f'<palette count="256" format="RGB">{hexp}</palette>')
                    ).genome.palette.stops
    assert len(s) == 256 and math.isclose(s[7].g, 7 / 255)
    with pytest.raises(GenomeError, match="old palette format"):
        parse_flame(wrap(MIN_XFORM, palette='<palette index0="10" index1="20" blend="0.5"/>'))
    r = parse_flame(wrap(MIN_XFORM, palette=""))
    assert r.report.palette_fallback["kind"] == "default" and r.genome.palette.name == "pyre"
    assert parse_flame(wrap(MIN_XFORM)).report.palette_fallback is None


def test_flame_fields():
    g = parse_flame(wrap(MIN_XFORM, 'rotate="30" quality="500" supersample="3" filter="0.8" '
                                    'estimator_radius="11" estimator_minimum="1" '
                                    'estimator_curve="0.6" background="0.1 0.2 0.3" '
                                    'palette_mode="linear"')).genome
    assert (g.rotate, g.quality, g.oversample, g.filter_radius) == (30.0, 500.0, 3, 0.8)
    assert (g.density.max_rad, g.density.min_rad, g.density.curve) == (11.0, 1.0, 0.6)
    assert g.background == [0.1, 0.2, 0.3] and g.palette_mode == "linear"
    assert g.size == {"width": 1024, "height": 1024} and g.scale == 100
    g = parse_flame(f'<flame name="t" center="1 2">{MIN_PALETTE}{MIN_XFORM}</flame>').genome
    assert g.scale == 50 and (g.cx, g.cy) == (1.0, 2.0) and g.density is None and g.tonemap is None
    g = parse_flame(wrap(MIN_XFORM, 'background="0 0 0" supersample="1"')).genome
    assert g.background is None and g.oversample is None
    r = parse_flame(wrap(MIN_XFORM, 'palette_mode="smooth"'))
    assert {"field": "palette_mode", "value": "smooth"} in r.report.ignored_fields


def test_tonemap_partial_fill():
    t = parse_flame(wrap(MIN_XFORM, 'gamma="2.5"')).genome.tonemap
    assert (t.gamma, t.vibrancy, t.highlight_power, t.brightness) == (2.5, 1.0, -1.0, 4.0)
    t = parse_flame(wrap(MIN_XFORM, 'brightness="2"')).genome.tonemap
    assert (t.gamma, t.brightness, t.vibrancy, t.highlight_power) == (4.0, 2.0, 1.0, -1.0)
    t = parse_flame(wrap(MIN_XFORM, 'gamma="2.2" vibrancy="0.5" highlight_power="1" '
                                    'brightness="3" gamma_threshold="0.05"')).genome.tonemap
    assert (t.gamma, t.vibrancy, t.highlight_power, t.brightness, t.gamma_threshold) == \
        (2.2, 0.5, 1.0, 3.0, 0.05)


def test_symmetry_and_density_errors():
    g = parse_flame(wrap(MIN_XFORM + '<symmetry kind="3"/>')).genome
    assert g.symmetry == {"kind": "rotational", "n": 3}
    g = parse_flame(wrap(MIN_XFORM + '<symmetry kind="-2"/>')).genome
    assert g.symmetry == {"kind": "dihedral", "n": 2}
    with pytest.raises(GenomeError):
        parse_flame(wrap(MIN_XFORM + '<symmetry kind="0"/>'))
    with pytest.raises(GenomeError):
        parse_flame(wrap(MIN_XFORM, 'estimator_radius="31"'))
    with pytest.raises(GenomeError):
        parse_flame(wrap(MIN_XFORM, 'estimator_curve="5"'))


def test_post_opacity_chaos():
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" linear="1" post="1 2 3 4 5 6" '
                         'opacity="0.5" chaos="1 0 2.5"/>')).genome.xforms[0]
    assert x.post == Affine(a=1, b=3, c=5, d=2, e=4, f=6)
    assert x.opacity == 0.5 and x.xaos == [1.0, 0.0, 2.5]
    x = parse_flame(wrap('<xform weight="1" coefs="1 0 0 1 0 0" linear="1" post="1 0 0 1 0 0"/>')
                    ).genome.xforms[0]
    assert x.post is None and x.opacity == 1.0 and x.xaos is None
    g = parse_flame(wrap(ONE_XFORM + '<finalxform coefs="1 0 0 1 0 0" linear="1" '
                                     'post="2 0 0 2 0 0"/>')).genome
    assert g.finalxform.post == Affine(a=2, e=2)


def test_structural_errors():
    with pytest.raises(GenomeError, match="missing coefs"):
        parse_flame(wrap('<xform weight="1" linear="1"/>'))
    with pytest.raises(GenomeError, match="no <xform>"):
        parse_flame(wrap(""))
    with pytest.raises(GenomeError, match="no <xform>"):
        parse_flame(wrap('<finalxform coefs="1 0 0 1 0 0" linear="1"/>'))


def test_multiple_flames_are_keyframes():
    a = wrap(ONE_XFORM, 'time="10" interpolation="smooth" temporal_samples="8"')
    b = wrap(ONE_XFORM.replace('color="0"', 'color="0.5"'), 'time="0"')
    r = parse_flame(f"<?xml version='1.0'?>\n{a}\n{b}")
    assert r.report.flame_count == 2
    assert r.genome.xforms[0].color == 0.0                 # document order first
    assert r.times == [0.0, 10.0] and r.keyframes[0].xforms[0].color == 0.5
    assert r.animation == {"interpolation": "smooth", "temporal_samples": 8.0}
    r = parse_flame(f"<flames>{wrap(ONE_XFORM)}</flames>")
    assert r.report.flame_count == 1 and not r.keyframes


def test_load_genome_sniffs_kind(tmp_path):
    p = tmp_path / "x.flam3"
    p.write_text(wrap(ONE_XFORM))
    g, kind, rep = load_genome(str(p))
    assert kind == "flame" and rep is not None and g.name == "t"
    q = tmp_path / "noext"
    q.write_text("   " + wrap(ONE_XFORM))
    assert load_genome(str(q))[1] == "flame"


def test_golden_fixture_flames_parse_cleanly():
    dirs = sorted(d for d in (GOLDEN / "flam3").iterdir() if d.is_dir())
    assert dirs
    for d in dirs:
        meta = json.loads((d / "meta.json").read_text())
        g, kind, rep = load_genome(str(d / f"{d.name}.flam3"))
        assert kind == "flame" and not rep.dropped_variations, d.name
        assert rep.palette_fallback is None, d.name
        assert g.size == {"width": meta["width"], "height": meta["height"]}, d.name
        assert all(v.index in cat.VARIATION_NAMES for x in g.xforms for v in x.variations)
