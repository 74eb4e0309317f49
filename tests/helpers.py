"""Shared builders for the render-path tests (genomes, packed blobs)."""
import numpy as np

from pyr3_amd import pack
from pyr3_amd.genome import Affine, ColorStop, Density, Genome, Palette, Tonemap, Variation, Xform


def gray_palette():
    return Palette("gray", [ColorStop(0, 0.2, 0.4, 0.6), ColorStop(1, 1, 0.8, 0.5)])


def xform(weight=1.0, affine=(1, 0, 0, 0, 1, 0), vars=((0, 1.0),), color=0.0, speed=0.5,
          opacity=1.0, post=None, params=None):
    vs = []
    for j, (idx, w) in enumerate(vars):
        v = Variation(idx, w)
        if params and j < len(params) and params[j]:
            v.params[:len(params[j])] = params[j]
        vs.append(v)
    x = Xform(weight=weight, color=color, color_speed=speed, affine=Affine(*affine),
              variations=vs, opacity=opacity)
    if post is not None:
        x.post = Affine(*post)
    return x


def genome(xforms, scale=30.0, cx=0.0, cy=0.0, final=None, density=True, **kw):
    g = Genome(name="t", xforms=list(xforms), scale=scale, cx=cx, cy=cy, palette=gray_palette(),
               finalxform=final, density=Density(9.0, 0.0, 0.4) if density else None,
               tonemap=Tonemap(gamma=4.0, vibrancy=1.0))
    for k, v in kw.items():
        setattr(g, k, v)
    return g


def blobs(g, W, H, os_=1, iters=100, fuse=20, seed=1, walkers=64, jitter=pack.DEFAULT_WALKER_JITTER):
    e = pack.expand_genome_for_gpu(g)
    xf = pack.pack_xforms(e)
    ds = pack.pack_xform_distrib(e)
    u = pack.chaos_uniforms(e, W * os_, H * os_, os_, iters, fuse, seed, walkers, jitter)
    pal = pack.pack_palette(e)
    return xf, ds, u, pal


def sierpinski():
    """Three contractive linear maps (attractor inside the unit square)."""
    return [xform(1.0, (0.5, 0, 0, 0, 0.5, 0), color=0.0),
            xform(1.0, (0.5, 0, 0.5, 0, 0.5, 0), color=0.5),
            xform(1.0, (0.5, 0, 0.25, 0, 0.5, 0.5), color=1.0)]


def psnr(a, b):
    a = a[..., :3].astype(np.float64)
    b = b[..., :3].astype(np.float64)
    mse = np.mean((a - b) ** 2)
    return float("inf") if mse == 0 else 10 * np.log10(255.0 ** 2 / mse)
