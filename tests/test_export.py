"""Output side: 16-bit PNG, OpenEXR, scene-linear export, auto-exposure.

Vectors from the reference's own tests (src/export-linear.test.ts,
src/half-float.test.ts, src/srgb.test.ts, src/auto-exposure.test.ts,
src/exr-encode.test.ts), an independent minimal EXR reader and a 16-bit PNG
decoder written here, and the linear-export oracle against a pure-Python
brute force.
"""
import math
import struct
import zlib

import numpy as np
import pytest

from pyr3_amd import export as X
from pyr3_amd.genome import Density, Genome
from pyr3_amd.png import encode_png


# ----------------------------------------------------------- linear export --

def test_linear_export_reference_vectors(orc):
    h = np.array([400, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1],
                 dtype=np.uint32).reshape(2, 2, 4)
    out = orc.histogram_to_linear_rgba(h, 1, 1, 2, 2.0, 0.001)
    assert out[0, 0, 0] == pytest.approx(0.2, abs=1e-6)
    assert out[0, 0, 3] == pytest.approx(0.002, abs=1e-6)
    h = np.array([255000, 0, 0, 1000], dtype=np.uint32).reshape(1, 1, 4)
    out = orc.histogram_to_linear_rgba(h, 1, 1, 1, 1.0, 0.01)
    assert out[0, 0, 0] == pytest.approx(2550, abs=0.1) and out[0, 0, 3] == 1.0


def test_linear_export_oracle_vs_brute_force(orc):
    rng = np.random.default_rng(3)
    W, H, os_ = 5, 3, 3
    h = rng.integers(0, 2 ** 32, (H * os_, W * os_, 4), dtype=np.uint64).astype(np.uint32)
    k1, k2 = 279.84375, 3.7e-9
    out = orc.histogram_to_linear_rgba(h, W, H, os_, k1, k2)
    scale = (k1 * k2) / (os_ * os_)
    for y in range(H):
        for x in range(W):
            s = [0, 0, 0, 0]
            for sy in range(os_):
                for sx in range(os_):
                    for c in range(4):
                        s[c] += int(h[y * os_ + sy, x * os_ + sx, c])
            want = [np.float32(s[c] * scale) for c in range(3)]
            want.append(np.float32(max(0.0, min(1.0, s[3] * scale))))
            assert list(out[y, x]) == want


# ---------------------------------------------------------- conversions --

def test_half_float_and_srgb():
    assert X.half_to_float(0x0000) == 0 and X.half_to_float(0x3C00) == 1
    assert X.half_to_float(0x4000) == 2 and X.half_to_float(0xC000) == -2
    assert X.half_to_float(0x0001) == pytest.approx(5.9604645e-8, abs=1e-12)
    assert X.half_to_float(0x7BFF) == 65504
    assert X.half_to_float(0x7C00) == math.inf and math.isnan(X.half_to_float(0x7E00))
    # numpy's binary16 decode agrees with the restatement on every pattern
    bits = np.arange(65536, dtype=np.uint32).astype(np.uint16)
    f = bits.view(np.float16).astype(np.float64)
    mine = np.array([X.half_to_float(int(b)) for b in bits[::97]])
    assert np.array_equal(np.nan_to_num(mine, nan=-7), np.nan_to_num(f[::97], nan=-7))
    assert X.srgb_to_linear(0) == 0 and X.srgb_to_linear(1) == pytest.approx(1, abs=1e-6)
    assert X.srgb_to_linear(0.5) == pytest.approx(0.21404, abs=1e-4)
    assert X.srgb_to_linear(0.04045) == pytest.approx(0.04045 / 12.92, abs=1e-7)
    assert X.srgb_to_linear(0.02) == pytest.approx(0.02 / 12.92, abs=1e-7)


def test_display_half_conversions():
    f = np.array([[[0.0, 0.5, 1.0, 1.0], [2.0, -1.0, np.nan, 0.25]]], dtype=np.float32)
    bits = X.display_to_half(f)
    assert bits.dtype == np.uint16 and bits[0, 0, 2] == 0x3C00
    p16 = X.display_half_to_png16(bits)
    assert p16[0, 0].tolist() == [0, 32768, 65535, 65535]     # 0.5 * 65535 rounds half up
    assert p16[0, 1].tolist() == [65535, 0, 0, 16384]
    lin = X.display_half_to_linear_exr(bits)
    assert lin[0, 0, 1] == pytest.approx(0.21404, abs=1e-4)
    assert lin[0, 1, 2] == 0 and lin[0, 1, 3] == 0.25        # NaN -> 0; alpha not curved
    # rounding to binary16 is to nearest even
    assert X.display_to_half(np.array([1 + 2 ** -11], dtype=np.float32))[0] == 0x3C00
    assert X.display_to_half(np.array([1 + 3 * 2 ** -11], dtype=np.float32))[0] == 0x3C02


# ------------------------------------------------------------------ EXR --

def _read_exr(b: bytes):
    """Minimal independent reader for uncompressed scanline EXR."""
    assert struct.unpack_from("<i", b, 0)[0] == 0x01312F76 and b[4] == 2
    off, attrs = 8, {}
    while b[off] != 0:
        e = b.index(b"\x00", off)
        name = b[off:e].decode()
        e2 = b.index(b"\x00", e + 1)
        kind = b[e + 1:e2].decode()
        n = struct.unpack_from("<i", b, e2 + 1)[0]
        attrs[name] = (kind, b[e2 + 5:e2 + 5 + n])
        off = e2 + 5 + n
    off += 1
    x0, y0, x1, y1 = struct.unpack("<iiii", attrs["dataWindow"][1])
    w, h = x1 - x0 + 1, y1 - y0 + 1
    chl, names, p = attrs["channels"][1], [], 0
    while chl[p] != 0:
        e = chl.index(b"\x00", p)
        names.append(chl[p:e].decode())
        assert struct.unpack_from("<i", chl, e + 1)[0] == 2      # FLOAT
        p = e + 1 + 16
    offs = struct.unpack_from(f"<{h}Q", b, off)
    img = np.zeros((h, w, 4), dtype=np.float32)
    for y in range(h):
        yy, size = struct.unpack_from("<ii", b, offs[y])
        assert yy == y and size == 4 * w * len(names)
        data = np.frombuffer(b, "<f4", w * len(names), offs[y] + 8).reshape(len(names), w)
        for ci, nm in enumerate(names):
            img[y, :, "RGBA".index(nm)] = data[ci]
    return attrs, names, img


def test_exr_structure_and_round_trip():
    out = X.encode_exr(np.full((2, 2, 4), 0.5, dtype=np.float32))
    assert struct.unpack_from("<I", out, 0)[0] == 0x01312F76 and out[4] == 2
    attrs, names, img = _read_exr(out)
    assert names == ["A", "B", "G", "R"]
    assert attrs["compression"] == ("compression", b"\x00")
    assert attrs["lineOrder"] == ("lineOrder", b"\x00")
    assert struct.unpack("<f", attrs["screenWindowWidth"][1])[0] == 1.0
    assert np.all(img == 0.5)
    one = X.encode_exr(np.array([[[3.25, 0, 0, 1]]], dtype=np.float32))
    assert struct.unpack_from("<f", one, len(one) - 4)[0] == 3.25   # R is the last channel
    rng = np.random.default_rng(8)
    src = (rng.standard_normal((7, 5, 4)) * 100).astype(np.float32)
    _, _, back = _read_exr(X.encode_exr(src))
    assert np.array_equal(back, src)
    hdr = len(one) - 8 - 16 - 8
    assert len(X.encode_exr(src)) == hdr + 7 * 8 + 7 * (8 + 4 * 5 * 4)


# ---------------------------------------------------------------- PNG16 --

def _decode_png16(png: bytes) -> np.ndarray:
    assert png[:8] == b"\x89PNG\r\n\x1a\n"
    off, idat, ihdr = 8, b"", None
    while off < len(png):
        n, = struct.unpack_from(">I", png, off)
        kind, body = png[off + 4:off + 8], png[off + 8:off + 8 + n]
        assert struct.unpack_from(">I", png, off + 8 + n)[0] == zlib.crc32(kind + body)
        if kind == b"IHDR":
            ihdr = struct.unpack(">IIBBBBB", body)
        elif kind == b"IDAT":
            idat += body
        off += 12 + n
    w, h, depth, ctype = ihdr[:4]
    assert depth == 16 and ctype == 6
    raw = zlib.decompress(idat)
    stride = w * 8
    rows = np.zeros((h, stride), dtype=np.int64)
    for y in range(h):
        f = raw[y * (stride + 1)]
        line = np.frombuffer(raw, np.uint8, stride, y * (stride + 1) + 1).astype(np.int64)
        if f == 0:
            rows[y] = line
        elif f == 1:
            for i in range(stride):
                rows[y, i] = (line[i] + (rows[y, i - 8] if i >= 8 else 0)) & 255
        elif f == 2:
            rows[y] = (line + (rows[y - 1] if y else 0)) & 255
        else:
            raise AssertionError(f"filter {f}")
    b = rows.astype(np.uint16).reshape(h, w, 4, 2)
    return (b[..., 0] << 8) | b[..., 1]


def test_png16_round_trip():
    rng = np.random.default_rng(2)
    img = rng.integers(0, 65536, (9, 13, 4), dtype=np.uint32).astype(np.uint16)
    img[3] = img[2]                  # an Up-friendly row
    img[5, :] = 40000                # a Sub-friendly row
    assert np.array_equal(_decode_png16(encode_png(img)), img)


# -------------------------------------------------------- auto-exposure --

def test_mean_luminance():
    assert X.compute_mean_luminance(np.zeros(0, dtype=np.uint8)) == 0
    assert X.compute_mean_luminance(np.array([0, 0, 0, 255, 0, 0, 0, 255], np.uint8)) == 0
    assert X.compute_mean_luminance(np.array([255, 255, 255, 255], np.uint8)) == 255
    assert X.compute_mean_luminance(np.array([60, 90, 120, 255, 0, 0, 0, 255], np.uint8)) == 45
    assert X.compute_mean_luminance_half(np.zeros(0, dtype=np.uint16)) == 0
    white = X.display_to_half(np.ones((2, 4), np.float32))
    assert X.compute_mean_luminance_half(white) == pytest.approx(255, abs=1e-4)
    over = X.display_to_half(np.full((2, 4), 3.0, np.float32))
    assert X.compute_mean_luminance_half(over) == pytest.approx(255, abs=1e-4)
    grey = X.display_to_half(np.full((3, 4), 0.5, np.float32))
    assert X.compute_mean_luminance_half(grey) == pytest.approx(127.5, abs=0.1)
    nan = X.display_to_half(np.array([[np.nan, np.nan, np.nan, 1]], np.float32))
    assert X.compute_mean_luminance_half(nan) == 0


def test_fit_brightness():
    calls = []

    def never(b):
        calls.append(b)
        return 0.0
    for args in ((3.0, 100.0, 95.0), (3.0, 100.0, 0.0), (3.0, 0.0, 50.0), (3.0, math.nan, 50.0),
                 (3.0, 50.0, math.inf)):
        fit = X.fit_brightness_to_target(*args, never)
        assert not fit.corrected and fit.brightness == 3.0 and fit.iters == 0
    assert calls == []
    fit = X.fit_brightness_to_target(1.0, 80.0, 10.0, lambda b: b * 10)
    assert fit.corrected and fit.brightness == pytest.approx(8.0) and fit.iters == 1
    assert fit.final_mean == pytest.approx(80.0)
    fit = X.fit_brightness_to_target(1.0, 60.0, 30 * math.log(2), lambda b: 30 * math.log(1 + b))
    assert fit.corrected and fit.iters <= X.AE_MAX_ITERS
    assert 1 - X.AE_DEADBAND < fit.final_mean / 60.0 < 1 + X.AE_DEADBAND
    fit = X.fit_brightness_to_target(1.0, 100.0, 5 * (1 - math.exp(-1 / 1e6)),
                                     lambda b: 5 * (1 - math.exp(-b / 1e6)))
    assert fit.brightness <= X.AE_BRIGHTNESS_MAX and fit.iters <= X.AE_MAX_ITERS


def test_probe_genome_geometry():
    from helpers import genome, sierpinski
    g = genome(sierpinski(), scale=100.0)
    g.size = {"width": 1920, "height": 1080}
    g.quality = 500.0
    g.density = Density(9.0, 1.0, 0.4)
    g.filter_radius = 0.8
    p = X.probe_genome(g)
    assert p.size == {"width": 1024, "height": 576} and p.quality == 16 and p.oversample == 1
    assert p.scale == pytest.approx(100.0 * 1024 / 1920)
    assert p.density.max_rad == pytest.approx(9 * 1024 / 1920)
    assert p.filter_radius == pytest.approx(0.8 * 1024 / 1920)
    assert g.size == {"width": 1920, "height": 1080} and g.density.max_rad == 9.0   # untouched
    g.size = {"width": 700, "height": 1000}
    p = X.probe_genome(g)
    assert p.size == {"width": 716, "height": 1024}        # floor(1024 * 0.7)
    g.size = None
    g.quality = 4.0
    p = X.probe_genome(g)
    assert p.size is None and p.quality == 4.0 and p.scale == g.scale
    b = X.with_brightness(g, 7.5)
    assert b.tonemap.brightness == 7.5 and g.tonemap.brightness != 7.5
    assert isinstance(b, Genome)
