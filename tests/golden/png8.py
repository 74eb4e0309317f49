"""Minimal PNG reader for 8-bit RGB/RGBA, non-interlaced images (the flam3-C
golden renders): zlib inflate, then the five per-row filters of the PNG
specification (None, Sub, Up, Average, Paeth) undone row by row."""
from __future__ import annotations

import struct
import zlib

import numpy as np


def _chunks(data: bytes):
    if data[:8] != b"\x89PNG\r\n\x1a\n":
        raise ValueError("not a PNG file")
    pos = 8
    while pos < len(data):
        n, = struct.unpack(">I", data[pos:pos + 4])
        kind = data[pos + 4:pos + 8]
        yield kind, data[pos + 8:pos + 8 + n]
        pos += 12 + n


def read_png(path) -> np.ndarray:
    """(H, W, 4) uint8 RGBA (alpha 255 for RGB files)."""
    data = open(path, "rb").read()
    idat = []
    w = h = ctype = 0
    for kind, body in _chunks(data):
        if kind == b"IHDR":
            w, h, depth, ctype, _, _, interlace = struct.unpack(">IIBBBBB", body)
            if depth != 8 or ctype not in (2, 6) or interlace != 0:
                raise ValueError(f"unsupported PNG: depth {depth}, colour type {ctype}, "
                                 f"interlace {interlace}")
        elif kind == b"IDAT":
            idat.append(body)
        elif kind == b"IEND":
            break
    bpp = 4 if ctype == 6 else 3
    raw = np.frombuffer(zlib.decompress(b"".join(idat)), dtype=np.uint8)
    stride = w * bpp
    raw = raw.reshape(h, stride + 1)
    out = np.zeros((h, stride), dtype=np.uint8)
    prev = np.zeros(stride, dtype=np.int32)
    for y in range(h):
        ft = int(raw[y, 0])
        f = raw[y, 1:].astype(np.int32)
        if ft == 0:
            cur = f
        elif ft == 1:    # Sub: running sum along the row, per channel
            cur = (np.cumsum(f.reshape(w, bpp), axis=0) & 255).reshape(-1)
        elif ft == 2:    # Up
            cur = (f + prev) & 255
        elif ft in (3, 4):
            cur = np.zeros(stride, dtype=np.int32)
            left = np.zeros(bpp, dtype=np.int32)
            upleft = np.zeros(bpp, dtype=np.int32)
            for x in range(0, stride, bpp):
                up = prev[x:x + bpp]
                if ft == 3:  # Average
                    v = (f[x:x + bpp] + ((left + up) >> 1)) & 255
                else:        # Paeth
                    p = left + up - upleft
                    pa, pb, pc = np.abs(p - left), np.abs(p - up), np.abs(p - upleft)
                    pred = np.where((pa <= pb) & (pa <= pc), left, np.where(pb <= pc, up, upleft))
                    v = (f[x:x + bpp] + pred) & 255
                cur[x:x + bpp] = v
                left, upleft = v, up
        else:
            raise ValueError(f"bad PNG filter type {ft} in row {y}")
        out[y] = cur
        prev = cur
    img = out.reshape(h, w, bpp)
    if bpp == 3:
        img = np.concatenate([img, np.full((h, w, 1), 255, np.uint8)], axis=2)
    return img
