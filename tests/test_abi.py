"""The C-ABI library loads and exports every entry point include/pyr3hip.h declares."""
import re
from pathlib import Path

from conftest import ROOT


def _declared():
    text = "".join((ROOT / "include" / h).read_text() for h in ("pyr3hip.h", "pyr3nn.h"))
    return sorted(set(re.findall(r"\b(p3(?:nn)?_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_stage_interface():
    names = _declared()
    for n in ("p3_create", "p3_reset", "p3_iterate", "p3_present", "p3_read_histogram",
              "p3_destroy", "p3_last_error"):
        assert n in names


def test_library_exports_every_declared_symbol():
    from pyr3_amd import _lib
    lib = _lib.load()
    for n in _declared():
        assert hasattr(lib, n), n
    assert set(_declared()) == set(_lib.SIGNATURES)


def test_supported_catalog_matches_host_list():
    from pyr3_amd import _lib
    from pyr3_amd.variations_supported import SUPPORTED
    lib = _lib.load()
    got = {i for i in range(400) if lib.p3_variation_supported(i)}
    assert got == set(SUPPORTED)


def test_create_error_is_reported_without_gpu_calls():
    import ctypes as C
    from pyr3_amd import _lib
    lib = _lib.load()
    h = C.c_void_p()
    # invalid size is rejected before any device work
    taps = (C.c_float * 3)(0.0, 1.0, 0.0)
    rc = lib.p3_create(0, 10, 1, taps, 3, C.byref(h))
    assert rc != 0 and not h
