"""ctypes loader of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module, and only to check or time the oracle; the product never does.
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "build" / "liboracle.so"

_F = C.POINTER(C.c_float)
_D = C.POINTER(C.c_double)
_U = C.POINTER(C.c_uint32)
_B = C.POINTER(C.c_uint8)
_SIGS = {
    "orc_isaac_init": (None, [_U, C.c_int]),
    "orc_isaac_round": (None, [_U]),
    "orc_isaac_draw": (None, [_U, _U, C.c_int]),
    "orc_seed_walkers": (None, [C.c_uint32, C.c_uint32, C.c_int, _U]),
    "orc_rand_transforms": (None, [_U, C.c_int, _F, _F]),
    "orc_chaos": (None, [_F, _U, C.c_int, _U, _F, _U, C.c_uint32, C.c_uint32, _F, C.c_int,
                         C.c_int]),
    "orc_chaos_follow": (None, [_F, _U, C.c_int, _U, _F, _U, _F, C.c_int, C.c_int, _F]),
    "orc_chaos_capture": (None, [_F, _U, C.c_int, _U, _F, _U, _U, C.c_uint32, C.c_uint32]),
    "orc_chaos_mt": (None, [_F, _U, C.c_int, _U, _F, _U, C.c_int]),
    "orc_kernel_norm_lut": (None, [_F]),
    "orc_density": (None, [_U, _U, _F, _F, C.c_int]),
    "orc_tonemap": (None, [_U, _F, _F, _U, _F, _F, _B, C.c_int]),
    "orc_spatial_kernel": (C.c_int, [C.c_double, C.c_int, _F]),
    "orc_calibration": (None, [C.c_double, C.c_double, C.c_double, C.c_int, _D, _D]),
    "orc_build_distrib": (C.c_int, [C.c_int, _D, _D, _U]),
    "orc_variation_f64": (C.c_int, [C.c_int, C.c_double, C.c_double, _D, _D, _D,
                                    C.POINTER(C.c_uint32), C.c_int, _D]),
    "orc_variation_f32": (C.c_int, [C.c_int, C.c_float, C.c_float, _F, _F, _F,
                                    C.POINTER(C.c_uint32), C.c_int, _F]),
    "orc_variation_supported": (C.c_int, [C.c_int]),
    "orc_dc_color": (C.c_int, [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, C.c_float,
                               _F]),
}

_lib = None


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB


def load() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        lib = C.CDLL(str(LIB))
        for n, (res, args) in _SIGS.items():
            f = getattr(lib, n)
            f.restype, f.argtypes = res, args
        _lib = lib
    return _lib


def p(a: np.ndarray, kind):
    return a.ctypes.data_as(kind)


# ------------------------------------------------------------- wrappers --

def isaac_init(state36: np.ndarray, from_rsl=True) -> np.ndarray:
    s = np.ascontiguousarray(state36, dtype=np.uint32).copy()
    load().orc_isaac_init(p(s, _U), 1 if from_rsl else 0)
    return s


def isaac_round(state36: np.ndarray) -> np.ndarray:
    s = np.ascontiguousarray(state36, dtype=np.uint32).copy()
    load().orc_isaac_round(p(s, _U))
    return s


def isaac_draw(state36: np.ndarray, n: int):
    s = np.ascontiguousarray(state36, dtype=np.uint32).copy()
    out = np.zeros(n, dtype=np.uint32)
    load().orc_isaac_draw(p(s, _U), p(out, _U), n)
    return out, s


def seed_walkers(seed: int, n: int, base: int = 0) -> np.ndarray:
    out = np.zeros((n, 36), dtype=np.uint32)
    load().orc_seed_walkers(seed & 0xFFFFFFFF, base, n, p(out, _U))
    return out


def rand_transforms(raw: np.ndarray):
    r = np.ascontiguousarray(raw, dtype=np.uint32)
    a = np.zeros(r.size, dtype=np.float32)
    b = np.zeros(r.size, dtype=np.float32)
    load().orc_rand_transforms(p(r, _U), r.size, p(a, _F), p(b, _F))
    return a, b


def chaos(xforms, distrib, uniforms, palette, hist=None, wb=0, we=None, trace_shape=None):
    """Run walkers [wb, we) single-threaded; returns (hist, trace or None)."""
    xf = np.ascontiguousarray(xforms, dtype=np.float32)
    ds = np.ascontiguousarray(distrib, dtype=np.uint32)
    u = np.ascontiguousarray(uniforms, dtype=np.uint32)
    pal = np.ascontiguousarray(palette, dtype=np.float32).reshape(-1)
    W, H = int(u[0]), int(u[1])
    if hist is None:
        hist = np.zeros((H, W, 4), dtype=np.uint32)
    we = int(u[12]) if we is None else we
    tr = None
    tw = ti = 0
    if trace_shape:
        tw, ti = trace_shape
        tr = np.zeros((tw, ti, 16), dtype=np.float32)
    load().orc_chaos(p(xf, _F), p(ds, _U), ds.shape[0], p(u, _U), p(pal, _F), p(hist, _U), wb, we,
                     p(tr, _F) if tr is not None else None, tw, ti)
    return hist, tr


def chaos_capture(xforms, distrib, uniforms, palette):
    """All walkers single-threaded with the colour-index capture buffer;
    returns (hist, idx_sum (H, W) u32)."""
    xf = np.ascontiguousarray(xforms, dtype=np.float32)
    ds = np.ascontiguousarray(distrib, dtype=np.uint32)
    u = np.ascontiguousarray(uniforms, dtype=np.uint32)
    pal = np.ascontiguousarray(palette, dtype=np.float32).reshape(-1)
    W, H = int(u[0]), int(u[1])
    hist = np.zeros((H, W, 4), dtype=np.uint32)
    idx = np.zeros((H, W), dtype=np.uint32)
    load().orc_chaos_capture(p(xf, _F), p(ds, _U), ds.shape[0], p(u, _U), p(pal, _F),
                             p(hist, _U), p(idx, _U), 0, int(u[12]))
    return hist, idx


def chaos_follow(xforms, distrib, uniforms, palette, force):
    """Replay traced walkers turn by turn from another run's recorded states."""
    xf = np.ascontiguousarray(xforms, dtype=np.float32)
    ds = np.ascontiguousarray(distrib, dtype=np.uint32)
    u = np.ascontiguousarray(uniforms, dtype=np.uint32)
    pal = np.ascontiguousarray(palette, dtype=np.float32).reshape(-1)
    fc = np.ascontiguousarray(force, dtype=np.float32)
    tw, ti = fc.shape[0], fc.shape[1]
    hist = np.zeros((int(u[1]), int(u[0]), 4), dtype=np.uint32)
    tr = np.zeros((tw, ti, 16), dtype=np.float32)
    load().orc_chaos_follow(p(xf, _F), p(ds, _U), ds.shape[0], p(u, _U), p(pal, _F), p(hist, _U),
                            p(tr, _F), tw, ti, p(fc, _F))
    return tr


def chaos_mt(xforms, distrib, uniforms, palette, threads: int):
    xf = np.ascontiguousarray(xforms, dtype=np.float32)
    ds = np.ascontiguousarray(distrib, dtype=np.uint32)
    u = np.ascontiguousarray(uniforms, dtype=np.uint32)
    pal = np.ascontiguousarray(palette, dtype=np.float32).reshape(-1)
    hist = np.zeros((int(u[1]), int(u[0]), 4), dtype=np.uint32)
    load().orc_chaos_mt(p(xf, _F), p(ds, _U), ds.shape[0], p(u, _U), p(pal, _F), p(hist, _U),
                        threads)
    return hist


def kernel_norm_lut() -> np.ndarray:
    out = np.zeros(257, dtype=np.float32)
    load().orc_kernel_norm_lut(p(out, _F))
    return out


def density(hist, de_uniforms, threads: int = 8) -> np.ndarray:
    h = np.ascontiguousarray(hist, dtype=np.uint32)
    u = np.ascontiguousarray(de_uniforms, dtype=np.uint32)
    out = np.zeros(h.shape, dtype=np.float32)
    lut = kernel_norm_lut()
    load().orc_density(p(h, _U), p(u, _U), p(lut, _F), p(out, _F), threads)
    return out


def tonemap(viz_uniforms, taps, filtered=None, hist=None, threads: int = 8, curves=None):
    u = np.ascontiguousarray(viz_uniforms, dtype=np.uint32)
    if u[16] != 0 and curves is None:
        raise ValueError("curves mask set without a curves LUT")
    cl = np.ascontiguousarray(curves, dtype=np.float32).reshape(-1) if curves is not None else None
    if cl is not None and cl.size != 5 * 256:
        raise ValueError("curves LUT must hold 5 x 256 floats")
    t = np.ascontiguousarray(taps, dtype=np.float32)
    W, H = int(u[0]), int(u[1])
    of = np.zeros((H, W, 4), dtype=np.float32)
    o8 = np.zeros((H, W, 4), dtype=np.uint8)
    f = np.ascontiguousarray(filtered, dtype=np.float32) if filtered is not None else None
    h = np.ascontiguousarray(hist, dtype=np.uint32) if hist is not None else None
    load().orc_tonemap(p(u, _U), p(t, _F), p(f, _F) if f is not None else None,
                       p(h, _U) if h is not None else None, p(cl, _F) if cl is not None else None,
                       p(of, _F), p(o8, _B), threads)
    return of, o8


def spatial_kernel(radius: float, oversample: int) -> np.ndarray:
    out = np.zeros(17, dtype=np.float32)
    n = load().orc_spatial_kernel(radius, oversample, p(out, _F))
    return out[:n].copy()


def calibration(scale, samples, brightness, oversample):
    k1, k2 = C.c_double(), C.c_double()
    load().orc_calibration(scale, samples, brightness, oversample, C.byref(k1), C.byref(k2))
    return k1.value, k2.value


def build_distrib(weights, xaos=None) -> np.ndarray:
    w = np.ascontiguousarray(weights, dtype=np.float64)
    n = w.size
    out = np.zeros((n + 1, 16384), dtype=np.uint32)
    x = np.ascontiguousarray(xaos, dtype=np.float64) if xaos is not None else None
    rc = load().orc_build_distrib(n, p(w, _D), p(x, _D) if x is not None else None, p(out, _U))
    if rc != 0:
        raise ValueError("negative or non-finite weight")
    return out


def variation(idx, x, y, weight, params=(), affine=(1, 0, 0, 0, 1, 0), rng_words=(),
              dtype=np.float64):
    """One variation evaluation; RNG draws come from `rng_words` in order."""
    par = np.zeros(10, dtype=dtype)
    par[:len(params)] = params
    a, b, c, d, e, f = affine
    a0 = np.array([a, b, c, 0.0], dtype=dtype)
    a1 = np.array([d, e, f, 0.0], dtype=dtype)
    words = np.ascontiguousarray(np.array(list(rng_words) or [0], dtype=np.uint32))
    out = np.zeros(2, dtype=dtype)
    if dtype == np.float64:
        rc = load().orc_variation_f64(idx, x, y, p(np.array([weight], dtype=dtype), _D),
                                      p(par, _D), p(np.concatenate([a0, a1]), _D),
                                      p(words, _U), len(rng_words), p(out, _D))
    else:
        rc = load().orc_variation_f32(idx, np.float32(x), np.float32(y),
                                      p(np.array([weight], dtype=dtype), _F), p(par, _F),
                                      p(np.concatenate([a0, a1]), _F), p(words, _U),
                                      len(rng_words), p(out, _F))
    if rc < 0:
        raise ValueError(f"variation {idx} not in the oracle")
    return float(out[0]), float(out[1]), rc


def dc_color(idx, x, y, params=(0.0, 0.0, 0.0)):
    """Direct-colour RGB of entry idx at (x, y), or None when it has none."""
    rgb = np.zeros(3, dtype=np.float32)
    q = list(params) + [0.0] * (3 - len(params))
    hit = load().orc_dc_color(idx, np.float32(x), np.float32(y), np.float32(q[0]),
                              np.float32(q[1]), np.float32(q[2]), p(rgb, _F))
    return tuple(float(v) for v in rgb) if hit else None


def variation_supported(idx: int) -> bool:
    return bool(load().orc_variation_supported(idx))


def histogram_to_linear_rgba(hist, width: int, height: int, oversample: int, k1: float,
                             k2: float) -> np.ndarray:
    """src/export-linear.ts:24-47: box sum of each os x os block of the raw
    histogram (exact integers), times (k1 k2) / os^2 in f64, stored as f32;
    alpha = clamp(count sum x scale, 0, 1)."""
    h = np.asarray(hist, dtype=np.uint64).reshape(height, oversample, width, oversample, 4)
    sums = h.sum(axis=(1, 3)).astype(np.float64)          # exact below 2^53
    scale = (k1 * k2) / (oversample * oversample)
    out = sums * scale
    out[..., 3] = np.clip(out[..., 3], 0.0, 1.0)
    return out.astype(np.float32)
