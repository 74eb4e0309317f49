"""PNG output with the genome embedded, and the tEXt reader behind `.png` loads.

* `encode_png(rgba)`: 8- or 16-bit RGBA, non-interlaced; each row takes the filter
  (None, Sub or Up) with the smallest sum of absolute filtered bytes, then
  zlib.
* `inject_text_chunk` / `read_text_chunks`: tEXt chunks as
  src/png-text-chunk.ts:64-190 writes and reads them — key (1..79 Latin-1
  bytes), NUL, value; characters above U+007F escaped as `\\uXXXX` (UTF-16
  units, like JSON) so the value stays ASCII; the chunk goes before IEND.
* `encode_render_png(rgba, genome)`: the image with the genome's
  `.pyr3.json` (compact) under key `pyr3`, as bin/pyr3-render.ts:517-559
  saves it; `pyr3_amd.flam3.load_genome` reads it back from a `.png`.
"""
from __future__ import annotations

import json
import struct
import zlib
from typing import Dict

import numpy as np

from .genome import Genome, GenomeError, genome_to_json

SIGNATURE = b"\x89PNG\r\n\x1a\n"


def _chunk(kind: bytes, body: bytes) -> bytes:
    return (struct.pack(">I", len(body)) + kind + body
            + struct.pack(">I", zlib.crc32(kind + body) & 0xFFFFFFFF))


def _filtered_rows(img: np.ndarray) -> bytes:
    h, w, bpp = img.shape
    rows = img.reshape(h, w * bpp).astype(np.int16)
    prev = np.zeros_like(rows)
    prev[1:] = rows[:-1]
    left = np.zeros_like(rows)
    left[:, bpp:] = rows[:, :-bpp]
    cands = [rows & 255, (rows - left) & 255, (rows - prev) & 255]   # None, Sub, Up
    cost = np.stack([np.minimum(c, 256 - c).sum(axis=1) for c in cands])  # |signed byte|
    choice = cost.argmin(axis=0)
    out = np.empty((h, w * bpp + 1), dtype=np.uint8)
    out[:, 0] = choice
    for k in range(3):
        sel = choice == k
        out[sel, 1:] = cands[k][sel]
    return out.tobytes()


def encode_png(rgba: np.ndarray, level: int = 6) -> bytes:
    """(H, W, 4) RGBA -> PNG bytes: uint8 gives 8-bit samples, uint16 16-bit
    samples stored big-endian (src/png16-encode.ts:39-82)."""
    img = np.asarray(rgba)
    if img.ndim != 3 or img.shape[2] != 4:
        raise ValueError(f"expected (H, W, 4) RGBA, got {img.shape}")
    h, w, _ = img.shape
    if img.dtype == np.uint16:
        depth = 16
        raw = np.ascontiguousarray(img, dtype=">u2").view(np.uint8).reshape(h, w, 8)
    else:
        depth = 8
        raw = np.ascontiguousarray(img, dtype=np.uint8)
    ihdr = struct.pack(">IIBBBBB", w, h, depth, 6, 0, 0, 0)
    return (SIGNATURE + _chunk(b"IHDR", ihdr)
            + _chunk(b"IDAT", zlib.compress(_filtered_rows(raw), level)) + _chunk(b"IEND", b""))


def _ascii_safe(value: str) -> str:
    out = []
    for ch in value:
        c = ord(ch)
        if c < 0x80:
            out.append(ch)
        elif c <= 0xFFFF:
            out.append(f"\\u{c:04x}")
        else:                      # UTF-16 surrogate pair, as a JS string holds it
            c -= 0x10000
            out.append(f"\\u{0xD800 + (c >> 10):04x}\\u{0xDC00 + (c & 0x3FF):04x}")
    return "".join(out)


def text_chunk(key: str, value: str) -> bytes:
    if not 1 <= len(key) <= 79:
        raise ValueError(f"pyr3 PNG tEXt: key length must be 1..79, got {len(key)}")
    return _chunk(b"tEXt", key.encode("latin-1") + b"\x00" + _ascii_safe(value).encode("ascii"))


def _chunks(png: bytes):
    off = len(SIGNATURE)
    while off + 12 <= len(png):
        n, = struct.unpack(">I", png[off:off + 4])
        yield off, png[off + 4:off + 8], png[off + 8:off + 8 + n]
        off += 12 + n


def inject_text_chunk(png: bytes, key: str, value: str) -> bytes:
    """Insert a tEXt chunk right before IEND."""
    if len(png) < len(SIGNATURE) + 12:
        raise ValueError("pyr3 PNG tEXt: input too short to be a PNG")
    if png[:8] != SIGNATURE:
        raise ValueError("pyr3 PNG tEXt: bad PNG signature")
    for off, kind, _ in _chunks(png):
        if kind == b"IEND":
            return png[:off] + text_chunk(key, value) + png[off:]
    raise ValueError("pyr3 PNG tEXt: IEND not found")


def read_text_chunks(png: bytes) -> Dict[str, str]:
    """Every tEXt chunk (key -> raw value); {} for input that is not a PNG."""
    out: Dict[str, str] = {}
    if len(png) < len(SIGNATURE) + 12 or png[:4] != SIGNATURE[:4]:
        return out
    for _, kind, body in _chunks(png):
        if kind == b"tEXt":
            sep = body.find(b"\x00")
            if sep > 0:
                out[body[:sep].decode("latin-1")] = body[sep + 1:].decode("latin-1")
        if kind == b"IEND":
            break
    return out


def encode_render_png(rgba: np.ndarray, genome: Genome) -> bytes:
    """The rendered image (RGBA8, or RGBA16 as uint16) with the genome under
    tEXt key `pyr3`."""
    text = json.dumps(genome_to_json(genome), separators=(",", ":"), ensure_ascii=False)
    return inject_text_chunk(encode_png(rgba), "pyr3", text)


def genome_from_png(png: bytes) -> Genome:
    from .genome import genome_from_text
    text = read_text_chunks(png).get("pyr3")
    if not text:
        raise GenomeError("PNG has no pyr3 metadata — only PNGs saved by pyr3 (Save Render) "
                          "carry the genome.")
    return genome_from_text(text)
