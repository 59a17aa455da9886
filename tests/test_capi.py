"""CPU checks of the C ABI: libblsamd.so loads, exports every entry point include/bls_amd.h declares, and the
Python binding covers them. No compute call is made (no GPU here)."""
import ctypes as ct
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "bls_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bls_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = declared_symbols()
    for must in ("bls_schur_create", "bls_reduce_schur_sparse_f64", "bls_reduce_schur_dense_f64",
                 "bls_reset_schur_sparse_f64", "bls_copyto_csc_f64", "bls_backsubstitute_schur_f64"):
        assert must in syms


def test_library_exports_every_declared_symbol(blsamd):
    lib = ct.CDLL(blsamd.device.LIBPATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header(blsamd):
    assert set(declared_symbols()) <= set(blsamd.device.SIGNATURES)
    blsamd.device.lib()   # binds every signature
    assert blsamd.device.lib().bls_version().decode().startswith("blsamd")


def test_no_gpu_fallback_in_product(blsamd):
    """The product path has no CPU route: the binding raises when the library is absent."""
    dev = blsamd.device
    saved_path, saved_lib = dev.LIBPATH, dev._lib
    try:
        dev.LIBPATH, dev._lib = os.path.join(ROOT, "nonexistent.so"), None
        try:
            dev.lib()
            raise AssertionError("expected BLSError")
        except dev.BLSError:
            pass
    finally:
        dev.LIBPATH, dev._lib = saved_path, saved_lib
