"""CPU checks of the C ABI boundary: librwig_hip.so loads and exports every
entry point include/rwig_hip.h declares (no device calls)."""
import ctypes

from rwig_amd import _lib


def test_header_declares_the_drop_in_entry_points():
    names = _lib.header_symbols()
    for n in ["rwig_available", "rwig_last_error", "rwig_sinkhorn_vanilla",
              "rwig_sinkhorn_log", "rwig_barycenter_parallel", "rwig_barycenter_log"]:
        assert n in names


def test_library_exports_every_declared_symbol():
    handle = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _lib.header_symbols() if not hasattr(handle, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert set(_lib.header_symbols()) <= set(_lib.SIGNATURES)
