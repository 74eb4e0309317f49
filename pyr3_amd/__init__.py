"""pyr3_amd — MI355X (gfx950) render path of the pyr3 fractal-flame renderer.

chaos game -> u32 RGBA histogram scatter -> density estimation -> tonemap,
as hand-written HIP kernels in libpyr3hip.so behind the C ABI of
include/pyr3hip.h. This package is the host side: the .pyr3.json genome codec
(genome.py), the byte-compatible packers (pack.py), the renderer facade over
the C ABI (renderer.py) and seeded synthetic genomes (synth.py).
"""
from .genome import (Affine, ColorStop, Density, Genome, GenomeError, Palette, Tonemap,
                     Variation, Xform, genome_from_json, genome_from_text, genome_to_json,
                     genome_to_text)

__all__ = [
    "Affine", "ColorStop", "Density", "Genome", "GenomeError", "Palette", "Tonemap",
    "Variation", "Xform", "genome_from_json", "genome_from_text", "genome_to_json",
    "genome_to_text", "create_renderer", "Renderer",
]


def __getattr__(name):
    # The renderer needs the built library; import it lazily so the codec and
    # packers stay usable for host-only work (tests, packing) without a GPU.
    if name in ("create_renderer", "Renderer"):
        from . import renderer
        return getattr(renderer, name)
    raise AttributeError(name)
