"""Catalog indices the gfx950 kernels implement (mirrors p3_variation_supported)."""
SUPPORTED = frozenset(range(0, 326))
