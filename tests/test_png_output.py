"""PNG output (pyr3_amd/png.py): pixels survive encode -> decode (checked with
the independent reader tests/golden/png8.py, and chunk CRCs with zlib), tEXt
chunks as src/png-text-chunk.ts writes them (before IEND, non-ASCII escaped
as \\uXXXX, key length 1..79), and the genome round trip through a saved
render (src/loader.ts:45-56)."""
import json
import struct
import sys
import zlib

import numpy as np
import pytest

from conftest import GOLDEN
from helpers import genome, sierpinski
from pyr3_amd import png
from pyr3_amd.flam3 import load_genome
from pyr3_amd.genome import GenomeError, genome_to_json

sys.path.insert(0, str(GOLDEN))
from png8 import read_png  # noqa: E402


def _chunk_list(data):
    off, out = 8, []
    while off < len(data):
        n, = struct.unpack(">I", data[off:off + 4])
        kind, body = data[off + 4:off + 8], data[off + 8:off + 8 + n]
        crc, = struct.unpack(">I", data[off + 8 + n:off + 12 + n])
        assert crc == zlib.crc32(kind + body) & 0xFFFFFFFF
        out.append((kind, body))
        off += 12 + n
    return out


def test_pixels_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    img = rng.integers(0, 256, size=(37, 53, 4), dtype=np.uint8)
    img[10:20] = img[9]                      # rows where Up wins
    img[:, 30:] = img[:, 29:30]              # runs where Sub wins
    data = png.encode_png(img)
    kinds = [k for k, _ in _chunk_list(data)]
    assert kinds == [b"IHDR", b"IDAT", b"IEND"]
    p = tmp_path / "a.png"
    p.write_bytes(data)
    assert np.array_equal(read_png(p), img)


def test_text_chunks():
    base = png.encode_png(np.zeros((2, 3, 4), np.uint8))
    data = png.inject_text_chunk(base, "pyr3", "café \U0001F525 ok")
    kinds = [k for k, _ in _chunk_list(data)]
    assert kinds == [b"IHDR", b"IDAT", b"tEXt", b"IEND"]
    got = png.read_text_chunks(data)
    assert got == {"pyr3": "caf\\u00e9 \\ud83d\\udd25 ok"}
    assert json.loads('"' + got["pyr3"] + '"') == "café \U0001F525 ok"
    with pytest.raises(ValueError):
        png.inject_text_chunk(base, "", "x")
    with pytest.raises(ValueError):
        png.inject_text_chunk(base, "k" * 80, "x")
    with pytest.raises(ValueError):
        png.inject_text_chunk(b"not a png at all, too", "k", "x")
    assert png.read_text_chunks(b"garbage" * 10) == {}
    # IEND not last (trailing bytes): the chunk still goes before IEND
    two = png.inject_text_chunk(base + b"\x00" * 5, "a", "1")
    assert png.read_text_chunks(two) == {"a": "1"}


def test_genome_round_trip_through_png(tmp_path):
    g = genome(sierpinski(), scale=12.5, cx=0.25, cy=-0.5)
    g.name = "sierpiñski"
    rgba = np.full((4, 5, 4), 200, np.uint8)
    p = tmp_path / "render.png"
    p.write_bytes(png.encode_render_png(rgba, g))
    g2, kind, rep = load_genome(str(p))
    assert kind == "pyr3-png" and rep is None
    assert genome_to_json(g2) == genome_to_json(g)
    assert np.array_equal(read_png(p), rgba)
    q = tmp_path / "plain.png"
    q.write_bytes(png.encode_png(rgba))
    with pytest.raises(GenomeError, match="no pyr3 metadata"):
        load_genome(str(q))
