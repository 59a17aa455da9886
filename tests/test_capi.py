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


def test_backsubstitute_arguments_rejected_before_launch(blsamd):
    """backsubstitute! on either path checks the element type and length of every vector it indexes, on the host,
    before any launch: a Float32 x_e, a short x_r or C_data of the wrong type would otherwise send the kernels past
    the end of a buffer."""
    import pytest
    import torch
    dev = blsamd.device
    p = blsamd.generators.config2(scale=0.0005)
    st = p.structure
    f64 = lambda n: torch.zeros(n, dtype=torch.float64)
    f32 = lambda n: torch.zeros(n, dtype=torch.float32)
    nc = st.coupling_storage_length()
    dev._check_backsub(st, "f64", f64(st.ne), f64(st.nr), f64(nc), None, f64(st.ne))
    dev._check_backsub(st, "mixed", f64(st.ne), f64(st.nr), f32(nc), f64(st.nr), None)
    bad = [("f64", f32(st.ne), f64(st.nr), f64(nc), None, None),
           ("f64", f64(st.ne), f64(st.nr - 1), f64(nc), None, None),
           ("f64", f64(st.ne), f64(st.nr), f32(nc), None, None),
           ("f32", f64(st.ne), f64(st.nr), f64(nc), None, None),
           ("f64", f64(st.ne), f64(st.nr), f64(nc - 3), None, None),
           ("f64", f64(st.ne), f64(st.nr), f64(nc), f64(st.nr + 1), None),
           ("f64", f64(st.ne), f64(st.nr), f64(nc), None, f32(st.ne)),
           ("f64", None, f64(st.nr), f64(nc), None, None)]
    for args in bad:
        with pytest.raises(dev.BLSError):
            dev._check_backsub(st, *args)
    q = blsamd.generators.config4(scale=0.0001, ncam=12)
    jst = q.jst

    class WS:
        structure = jst
        factored = f64(jst.raw_length())
    dev._check_qr_backsub(jst, WS, f64(jst.schur.ne), f64(jst.schur.nr))
    for xe, xr in ((f32(jst.schur.ne), f64(jst.schur.nr)), (f64(jst.schur.ne), f64(jst.schur.nr + 6))):
        with pytest.raises(dev.BLSError):
            dev._check_qr_backsub(jst, WS, xe, xr)

    class WS2(WS):
        factored = f64(jst.raw_length() - 1)
    with pytest.raises(dev.BLSError):
        dev._check_qr_backsub(jst, WS2, f64(jst.schur.ne), f64(jst.schur.nr))


def test_julia_binding_calls_only_exported_symbols(blsamd):
    """Every C symbol the Julia binding (julia/BLSAmd.jl) calls is exported by the library and declared in the
    header, and it defines a method for each of the reference's entry points on the path."""
    lib = ct.CDLL(blsamd.device.LIBPATH)
    text = open(os.path.join(ROOT, "blocklinearsystems.jl_amd", "julia", "BLSAmd.jl")).read()
    called = set(re.findall(r"ccall\(\(:(bls_[a-z0-9_]+)", text)) | set(re.findall(r":(bls_[a-z0-9_]+)", text))
    declared = set(declared_symbols())
    assert called and all(hasattr(lib, s) for s in called), sorted(s for s in called if not hasattr(lib, s))
    assert called <= declared, sorted(called - declared)
    for method in ("function BLS.reset_schur_sparse!", "function BLS.reset_schur_dense!",
                   "function BLS.reduce_schur_sparse!", "function BLS.reduce_schur_dense!",
                   "function BLS.reduce_qr!", "function Base.copyto!(dest::SparseMatrixCSC{Float64}"):
        assert method in text, method
    assert text.count("function BLS.backsubstitute!") == 2   # Schur and QR
