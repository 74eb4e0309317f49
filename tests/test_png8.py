"""The test-side PNG reader (tests/golden/png8.py) on PNGs written here with
every row filter of the PNG specification (None, Sub, Up, Average, Paeth),
RGB and RGBA: decoding must give back the exact pixels."""
import struct
import sys
import zlib

import numpy as np

from conftest import GOLDEN

sys.path.insert(0, str(GOLDEN))
from png8 import read_png  # noqa: E402


def _paeth(a, b, c):
    p = a + b - c
    pa, pb, pc = abs(p - a), abs(p - b), abs(p - c)
    if pa <= pb and pa <= pc:
        return a
    return b if pb <= pc else c


def _filter_row(ft, row, prev, bpp):
    out = bytearray(len(row))
    for i, x in enumerate(row):
        a = row[i - bpp] if i >= bpp else 0
        b = prev[i]
        c = prev[i - bpp] if i >= bpp else 0
        pred = (0, a, b, (a + b) // 2, _paeth(a, b, c))[ft]
        out[i] = (x - pred) & 255
    return bytes([ft]) + bytes(out)


def _write_png(path, img, filters):
    h, w, bpp = img.shape
    raw, prev = b"", bytes(w * bpp)
    for y in range(h):
        row = img[y].tobytes()
        raw += _filter_row(filters[y % len(filters)], row, prev, bpp)
        prev = row

    def chunk(kind, body):
        return struct.pack(">I", len(body)) + kind + body + struct.pack(
            ">I", zlib.crc32(kind + body) & 0xffffffff)

    ihdr = struct.pack(">IIBBBBB", w, h, 8, 6 if bpp == 4 else 2, 0, 0, 0)
    path.write_bytes(b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", ihdr) + chunk(b"IDAT", zlib.compress(raw))
                     + chunk(b"IEND", b""))


def test_round_trip_every_filter(tmp_path):
    rng = np.random.default_rng(11)
    for bpp in (3, 4):
        img = rng.integers(0, 256, size=(23, 17, bpp), dtype=np.uint8)
        img[5:9] = img[4]           # runs where Up / Paeth predict well
        p = tmp_path / f"t{bpp}.png"
        _write_png(p, img, filters=[0, 1, 2, 3, 4, 4, 3, 2, 1])
        got = read_png(p)
        want = img if bpp == 4 else np.concatenate(
            [img, np.full(img.shape[:2] + (1,), 255, np.uint8)], axis=2)
        assert np.array_equal(got, want), bpp
