"""Catalog indices the gfx950 kernels implement (mirrors p3_variation_supported)."""
SUPPORTED = frozenset({0, 1, 2, 97})
