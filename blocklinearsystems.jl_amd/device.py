"""ctypes binding of libblsamd.so (the HIP kernels) with the reference's entry-point names.

Julia's ``reduce_schur_sparse!`` is ``reduce_schur_sparse_`` here (a trailing underscore for the bang); arguments keep
the reference's order and meaning (src/schur_reduction.jl:465-507, src/SparseBlockLinearSystem.jl:356-362). Value
arrays are torch tensors on the GPU (float64), index arrays are the host mirror's 1-based Int64 numpy arrays.
There is no CPU fallback: if the library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes as ct
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libblsamd.so")
P64 = ct.POINTER(ct.c_int64)
VP = ct.c_void_p

BLS_OK, BLS_FAILED = 0, 1


class BLSError(RuntimeError):
    pass


_lib = None

# exported symbols and their signatures (restype, argtypes)
SIGNATURES = {
    "bls_last_error": (ct.c_char_p, []),
    "bls_version": (ct.c_char_p, []),
    "bls_alloc": (ct.c_int, [ct.POINTER(VP), ct.c_int64]),
    "bls_free": (ct.c_int, [VP]),
    "bls_upload": (ct.c_int, [VP, VP, ct.c_int64, VP]),
    "bls_download": (ct.c_int, [VP, VP, ct.c_int64, VP]),
    "bls_stream_synchronize": (ct.c_int, [VP]),
    "bls_schur_create": (ct.c_int, [ct.c_int64] * 4 + [P64] * 10 + [ct.POINTER(VP)]),
    "bls_schur_destroy": (None, [VP]),
    "bls_schur_set_sparse": (ct.c_int, [VP, ct.c_int64, P64, ct.c_int64, P64, P64, P64, P64, ct.c_int64, P64, P64]),
    "bls_schur_set_dense": (ct.c_int, [VP]),
    "bls_reset_schur_sparse_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, VP, VP]),
    "bls_reset_schur_dense_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, VP, VP]),
    "bls_reduce_schur_sparse_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, VP, VP, P64, VP]),
    "bls_reduce_schur_dense_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, VP, VP, P64, VP]),
    "bls_reduce_schur_sparse_f64_async": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_reduce_schur_sparse_f64_noalloc": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, VP, VP]),
    "bls_reduce_schur_dense_f64_async": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_backsubstitute_schur_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_copyto_csc_f64": (ct.c_int, [ct.c_int64, VP, VP, VP, VP]),
    "bls_scatter_map_create": (ct.c_int, [ct.c_int64, P64, ct.POINTER(VP)]),
    "bls_schur_set_timing": (ct.c_int, [VP, ct.c_int]),
    "bls_schur_set_tuning": (ct.c_int, [VP, ct.c_int64, ct.c_int64]),
    "bls_schur_get_tuning": (ct.c_int, [VP, P64, P64]),
    "bls_schur_set_phase2_mode": (ct.c_int, [VP, ct.c_int]),
    "bls_schur_phase2_info": (ct.c_int, [VP, P64]),
    "bls_schur_autotune_f64": (ct.c_int, [VP, ct.c_int64, P64, P64] + [VP] * 5 + [ct.c_double] + [VP] * 6
                               + [ct.POINTER(ct.c_float), VP]),
    "bls_pcg_create": (ct.c_int, [ct.c_int64, ct.c_int64, P64, P64, ct.c_int64, P64, ct.c_int64, P64, P64, P64, P64,
                                  ct.c_int64, P64, ct.POINTER(VP)]),
    "bls_pcg_destroy": (None, [VP]),
    "bls_pcg_values_size": (ct.c_int, [VP, P64, P64]),
    "bls_pcg_solve_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, ct.c_int64, VP, P64, ct.POINTER(ct.c_double), VP]),
    "bls_qr_create": (ct.c_int, [ct.c_int64] * 3 + [P64] * 2 + [ct.c_int64] + [P64] * 4 + [ct.POINTER(VP)]),
    "bls_qr_destroy": (None, [VP]),
    "bls_jfill_create": (ct.c_int, [ct.c_int64, ct.c_int64, P64, P64, ct.c_int64] + [P64] * 8 + [ct.c_int64] + [P64] * 5
                         + [P64, P64, ct.POINTER(VP)]),
    "bls_jfill_destroy": (None, [VP]),
    "bls_jfill_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_jfill_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_reduce_qr_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, P64, VP]),
    "bls_reduce_qr_f64_async": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP]),
    "bls_qr_set_timing": (ct.c_int, [VP, ct.c_int]),
    "bls_qr_kernel_time": (ct.c_int, [VP, ct.POINTER(ct.c_float)]),
    "bls_qr_set_reduced": (ct.c_int, [VP, P64, P64, P64, ct.c_int64] + [P64] * 8),
    "bls_reduced_mul_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_reduced_tmul_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_reduced_residual_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_reduced_diagonal_blocks_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_pcg_relres_f64": (ct.c_int, [VP, VP, VP, VP, ct.POINTER(ct.c_double), VP]),
    "bls_pcg_residual_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.POINTER(ct.c_double), VP]),
    "bls_pcg_set_groups": (ct.c_int, [VP, ct.c_int]),
    "bls_pcg_group_blocks_size": (ct.c_int, [VP, ct.POINTER(ct.c_int64)]),
    "bls_pcg_group_blocks_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_pcg_solve_groups_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, ct.c_int64, VP, P64, ct.POINTER(ct.c_double), VP]),
    "bls_axpy_f64": (ct.c_int, [ct.c_int64, ct.c_double, VP, VP, VP]),
    "bls_dense_residual_f64": (ct.c_int, [ct.c_int64, VP, VP, VP, VP, VP]),
    "bls_reduce_schur_dense_mfma_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, ct.c_int64,
                                                   ct.POINTER(ct.c_int64), VP]),
    "bls_reduce_schur_dense_mfma_mixed": (ct.c_int, [VP, VP, VP, VP, VP, VP, ct.c_double, VP, VP, VP, VP, ct.c_int64,
                                                     ct.POINTER(ct.c_int64), VP]),
    "bls_schur_factor_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, VP, VP, VP, ct.POINTER(ct.c_int64), VP]),
    "bls_schur_retained_residual_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_schur_retained_residual_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_qr_set_sparse": (ct.c_int, [VP, ct.c_int64, P64, ct.c_int64, P64, P64, P64, P64, P64, P64]),
    "bls_reduced_normal_sparse_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_reduced_normal_sparse_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_reduced_normal_dense_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_reduced_normal_dense_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_qr_backsubstitute_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_qr_backsubstitute_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_diag_create": (ct.c_int, [ct.c_int64, P64, ct.POINTER(VP)]),
    "bls_asm_create": (ct.c_int, [ct.c_int64, ct.c_int64, P64, P64, P64, P64, P64, P64, P64, P64, P64, P64, P64,
                                  ct.c_int64, P64, P64, P64, ct.POINTER(VP)]),
    "bls_asm_destroy": (None, [VP]),
    "bls_asm_lengths": (ct.c_int, [VP, P64]),
    "bls_asm_hessian_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_asm_dot_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_asm_diag_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_asm_dot_f32": (ct.c_int, [VP, VP, VP, VP, VP, VP, VP, VP]),
    "bls_asm_diag_f32": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_jtr_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_jtj_diag_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_jtj_quadform_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_reduced_jacobian_nnz": (ct.c_int, [VP, P64]),
    "bls_build_reduced_jacobian_pattern": (ct.c_int, [VP, P64, P64]),
    "bls_check_reduced_jacobian_pattern": (ct.c_int, [VP, ct.c_int64, ct.c_int64, P64, P64]),
    "bls_reduced_jacobian_sparse_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_reduced_jacobian_dense_f64": (ct.c_int, [VP, VP, VP, VP, VP, VP]),
    "bls_chol_create": (ct.c_int, [ct.c_int64, ct.POINTER(VP)]),
    "bls_chol_destroy": (None, [VP]),
    "bls_chol_factor_f64": (ct.c_int, [VP, VP, P64, VP]),
    "bls_chol_factor_f64_async": (ct.c_int, [VP, VP, VP]),
    "bls_chol_factor_f32": (ct.c_int, [VP, VP, P64, VP]),
    "bls_chol_factor_f32_async": (ct.c_int, [VP, VP, VP]),
    "bls_chol_solve_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_chol_factor_copy_f64": (ct.c_int, [VP, VP, VP]),
    "bls_ldl_create": (ct.c_int, [ct.c_int64, ct.POINTER(VP)]),
    "bls_ldl_destroy": (None, [VP]),
    "bls_ldl_factor_f64": (ct.c_int, [VP, VP, P64, P64, VP]),
    "bls_ldl_solve_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_ldl_factor_copy_f64": (ct.c_int, [VP, VP, VP, VP]),
    "bls_cond_create": (ct.c_int, [ct.c_int64, ct.POINTER(VP)]),
    "bls_cond_destroy": (None, [VP]),
    "bls_cond_is_stale": (ct.c_int, [VP]),
    "bls_cond_markstale": (ct.c_int, [VP]),
    "bls_prepare_scale_f64": (ct.c_int, [VP, VP, VP, VP, ct.c_double, ct.c_int, VP]),
    "bls_cond_nrescale": (ct.c_int, [VP, P64, VP]),
    "bls_unscale_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_diag_destroy": (None, [VP]),
    "bls_factor_diagonal_blocks_f64": (ct.c_int, [VP, VP, P64, VP]),
    "bls_factor_diagonal_blocks_f32": (ct.c_int, [VP, VP, P64, VP]),
    "bls_solve_diagonal_blocks_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_solve_diagonal_blocks_f32": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_pcg_features_f32": (ct.c_int, [VP, VP, ct.c_double, VP, VP]),
    "bls_pcg_shift_blocks_f64": (ct.c_int, [VP, VP, VP, VP, VP]),
    "bls_mlp_dims": (ct.c_int, [ct.POINTER(ct.c_int)] * 4),
    "bls_mlp_forward_f32": (ct.c_int, [VP, VP, VP, VP, ct.c_int64, VP, VP]),
    "bls_mlp_grad_f32": (ct.c_int, [VP, VP, VP, VP, VP, ct.c_int64, VP, VP, VP, VP]),
    "bls_mlp_adam_f32": (ct.c_int, [VP, VP, VP, VP, ct.c_int64, ct.c_float, VP]),
    "bls_schur_kernel_times": (ct.c_int, [VP, ct.POINTER(ct.c_float), ct.c_int]),
}
# Float32 (SchurBlockHessian{Float32} / SchurBlockJacobian{Float32}) and mixed variants: the same argument lists
for _v in ("f32", "mixed"):
    for _n in ("bls_reset_schur_sparse", "bls_reset_schur_dense", "bls_reduce_schur_sparse", "bls_reduce_schur_dense",
               "bls_reduce_schur_sparse_%s_async", "bls_backsubstitute_schur"):
        _src = _n.replace("%s", "f64") if "%s" in _n else _n + "_f64"
        SIGNATURES[_n % _v if "%s" in _n else _n + "_" + _v] = SIGNATURES[_src]
for _n in ("bls_copyto_csc", "bls_reduce_qr", "bls_reduced_mul", "bls_reduced_tmul", "bls_reduced_diagonal_blocks",
           "bls_jtr", "bls_jtj_diag", "bls_jtj_quadform", "bls_reduced_jacobian_sparse", "bls_reduced_jacobian_dense"):
    SIGNATURES[_n + "_f32"] = SIGNATURES[_n + "_f64"]
SIGNATURES["bls_reduce_qr_f32_async"] = SIGNATURES["bls_reduce_qr_f64_async"]


def _variant(s_dtype, e_dtype):
    """The precision variant of a call from the element types of the reduced system and of the blocks:
    f64 (both Float64), f32 (both Float32, SchurBlockHessian{Float32}), mixed (Float32 blocks, Float64 S)."""
    import torch
    if s_dtype == torch.float64 and e_dtype == torch.float64:
        return "f64"
    if s_dtype == torch.float32 and e_dtype == torch.float32:
        return "f32"
    if s_dtype == torch.float64 and e_dtype == torch.float32:
        return "mixed"
    raise BLSError(f"no device variant for a reduced system in {s_dtype} from blocks in {e_dtype}")


def lib():
    """Load libblsamd.so (built by __graft_entry__.build() / csrc/Makefile). Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise BLSError(f"{LIBPATH} is missing: build it with `make -C blocklinearsystems.jl_amd/csrc`")
        L = ct.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        raise BLSError(lib().bls_last_error().decode())
    return rc


def _p64(a, keep):
    a = np.ascontiguousarray(a, dtype=np.int64)
    keep.append(a)
    return a.ctypes.data_as(P64)


def _dp(t):
    """Device pointer of a torch tensor (None -> NULL)."""
    if t is None:
        return None
    import torch
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.is_contiguous()):
        raise BLSError("expected a contiguous GPU tensor")
    return ct.c_void_p(t.data_ptr())


def _stream(stream):
    if stream is None:
        import torch
        return ct.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ct.c_void_p(stream)


class SchurHandle:
    """A registered SchurStructure (+ its workspace index and reduced-system layout) on the device:
    the C ABI's bls_schur. ``pattern`` (build_schur_pattern's result) selects the sparse S; ``dense=True`` the dense
    nr x nr S. ``r_to_s`` is build_r_to_s_map's result (sparse only)."""

    def __init__(self, st, ws_index, pattern=None, r_to_s=None, dense=False):
        keep = []
        h = VP()
        _check(lib().bls_schur_create(st.nr, st.nrblocks, st.neblocks, st.ncoup, _p64(st.r_dims, keep),
                                      _p64(st.e_dims, keep), _p64(st.boffsets_r, keep), _p64(st.boffsets_e, keep),
                                      _p64(st.eoffsets, keep), _p64(st.eliminated_ptr, keep),
                                      _p64(st.entry_retained, keep), _p64(st.coffsets, keep),
                                      _p64(ws_index.retained_ptr, keep), _p64(ws_index.retained_entry, keep), ct.byref(h)))
        self.h = h
        self.st = st
        self.dense = dense
        try:
            if dense:
                _check(lib().bls_schur_set_dense(h))
            else:
                if pattern is None:
                    raise BLSError("a sparse handle needs the pattern of build_schur_pattern")
                rts = np.zeros(0, np.int64) if r_to_s is None else r_to_s
                _check(lib().bls_schur_set_sparse(h, pattern.nnz, _p64(pattern.S_colptr, keep), len(pattern.block_rowval),
                                                  _p64(pattern.block_colptr, keep), _p64(pattern.block_rowval, keep),
                                                  _p64(pattern.block_row_offsets, keep), _p64(pattern.col_heights, keep),
                                                  len(rts), _p64(rts, keep), _p64(pattern.diag_inds, keep)))
                self.nnz = pattern.nnz
        except Exception:
            lib().bls_schur_destroy(h)
            self.h = None
            raise

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_schur_destroy(self.h)
            self.h = None

    def set_timing(self, on=True):
        _check(lib().bls_schur_set_timing(self.h, 1 if on else 0))

    def set_tuning(self, part, lanes):
        """The phase-2 work-list shape (landmarks per part, rows per run of block rows); rebuilds the work list."""
        _check(lib().bls_schur_set_tuning(self.h, int(part), int(lanes)))

    def tuning(self):
        a, b = ct.c_int64(0), ct.c_int64(0)
        _check(lib().bls_schur_get_tuning(self.h, ct.byref(a), ct.byref(b)))
        return int(a.value), int(b.value)

    def set_phase2_mode(self, mode):
        """0: the tiles of S on the matrix cores when the layout fits them, else the work list; 1: the work list;
        2: the tiles (an error when the layout does not fit)."""
        _check(lib().bls_schur_set_phase2_mode(self.h, int(mode)))

    def phase2_info(self):
        """(kernel in use: 1 work list / 2 tiles, tiles, landmark entries, matrix-core products)."""
        arr = (ct.c_int64 * 4)()
        _check(lib().bls_schur_phase2_info(self.h, arr))
        return tuple(int(v) for v in arr)

    def autotune(self, candidates, S, g_S, b_r, b_e, scaling_e, lam, E_data, C_data, ws, s_r=None, s_e=None,
                 stream=None):
        """Time one Float64 reduction per (part, lanes) candidate and keep the fastest; returns (part, lanes, ms)."""
        keep = []
        parts = np.array([c[0] for c in candidates], dtype=np.int64)
        lanes = np.array([c[1] for c in candidates], dtype=np.int64)
        best = ct.c_float(0)
        _check(lib().bls_schur_autotune_f64(self.h, len(parts), _p64(parts, keep), _p64(lanes, keep), _dp(S), _dp(g_S),
                                            _dp(b_r), _dp(b_e), _dp(scaling_e), float(lam), _dp(E_data), _dp(C_data),
                                            _dp(ws.E_inv_data), _dp(ws.z_data), _dp(s_r), _dp(s_e), ct.byref(best),
                                            _stream(stream)))
        return self.tuning() + (float(best.value),)

    def kernel_times(self):
        arr = (ct.c_float * 4)()
        _check(lib().bls_schur_kernel_times(self.h, arr, 4))
        return list(arr)


class SchurReductionWorkspace:
    """Device counterpart of SchurReductionWorkspace{T} (src/schur_reduction.jl:16-27): E_inv, z and the failure
    flag of the last reduction. variant "f64" / "mixed": E_inv in Float64; "f32": in Float32 (z is Float64)."""

    def __init__(self, handle, variant="f64"):
        import torch
        st = handle.st
        self.variant = variant
        edt = torch.float32 if variant == "f32" else torch.float64
        self.E_inv_data = torch.empty(st.eliminated_storage_length(), dtype=edt, device="cuda")
        self.z_data = torch.empty(st.ne, dtype=torch.float64, device="cuda")
        self.d_failed = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.failed_block = 0


def reset_schur_sparse_(nzval, R_nzval, handle, scaling_r, lam, s_r=None, stream=None, check=True):
    """reset_schur_sparse! (src/schur_reduction.jl:416-436). check=False skips the host-side check_damping of
    scaling_r (a device reduction and a synchronization) for a caller that validated it already."""
    if check:
        _check_damping(lam, scaling_r, "scaling_r")
    v = _variant(nzval.dtype, R_nzval.dtype if R_nzval is not None else nzval.dtype)
    f = getattr(lib(), "bls_reset_schur_sparse_" + v)
    _check(f(handle.h, _dp(nzval), _dp(R_nzval), _dp(scaling_r), float(lam), _dp(s_r), _stream(stream)))


def reset_schur_dense_(S, R, handle, scaling_r, lam, s_r=None, stream=None, check=True):
    """reset_schur_dense! (src/schur_reduction.jl:389-407); S, R column-major flat (nr*nr)."""
    if check:
        _check_damping(lam, scaling_r, "scaling_r")
    f = getattr(lib(), "bls_reset_schur_dense_" + _variant(S.dtype, R.dtype))
    _check(f(handle.h, _dp(S), _dp(R), _dp(scaling_r), float(lam), _dp(s_r), _stream(stream)))


def _check_damping(lam, scaling, name):
    """check_damping (src/utils.jl:100-107), on the host before any work."""
    import math
    if not (math.isfinite(lam) and lam >= 0):
        raise ValueError(f"lambda must be finite and non-negative, got {lam}")
    if scaling is not None:
        import torch
        bad = ~(torch.isfinite(scaling) & (scaling >= 0))
        if bool(bad.any()):
            i = int(torch.nonzero(bad)[0])
            raise ValueError(f"{name} must be finite and non-negative, but entry {i + 1} is {float(scaling[i])}")


def reduce_schur_sparse_(nzval, g_S, b_r, b_e, scaling_e, lam, handle, E_data, C_data, ws, s_r=None, s_e=None,
                         stream=None, sync=True, check=True):
    """reduce_schur_sparse! (src/schur_reduction.jl:489-507). Returns True, or False with ws.failed_block set.
    With sync=False the status is left on the device (ws.d_failed) and None is returned."""
    if check:
        _check_damping(lam, scaling_e, "scaling_e")
    return _reduce(handle, False, nzval, g_S, b_r, b_e, scaling_e, lam, E_data, C_data, ws, s_r, s_e, stream, sync)


def reduce_schur_dense_(S, g_S, b_r, b_e, scaling_e, lam, handle, E_data, C_data, ws, s_r=None, s_e=None,
                        stream=None, sync=True, check=True):
    """reduce_schur_dense! (src/schur_reduction.jl:465-479); S column-major flat (nr*nr), upper triangle written."""
    if check:
        _check_damping(lam, scaling_e, "scaling_e")
    return _reduce(handle, True, S, g_S, b_r, b_e, scaling_e, lam, E_data, C_data, ws, s_r, s_e, stream, sync)


def reduce_schur_dense_mfma_(S, g_S, b_r, b_e, scaling_e, lam, handle, E_data, C_data, ws, nslices=0, stream=None,
                             check=True):
    """reduce_schur_dense! with phase 2 as a SYRK on the matrix cores (bls_reduce_schur_dense_mfma_*): Float64 E, C
    and FP64 MFMA (C_data Float64), or Float32 couplings C_data with FP32 MFMA and Float64 cross-slice
    accumulation (C_data Float32); E and the landmark factorisation are Float64 in both (workspace "f64").
    Returns True, or False with ws.failed_block set."""
    import torch
    if check:
        _check_damping(lam, scaling_e, "scaling_e")
    if not handle.dense:
        raise BLSError("the MFMA reduction needs a dense handle")
    v = "f64" if C_data.dtype == torch.float64 else "mixed"
    if ws.variant != "f64":
        raise BLSError("the MFMA reduction keeps E_inv in Float64: use a Float64 workspace")
    st = handle.st
    _require("S", S, torch.float64, st.nr * st.nr)
    _require("g_S", g_S, torch.float64, st.nr)
    _require("E_data", E_data, torch.float64, st.eliminated_storage_length())
    if C_data.dtype not in (torch.float64, torch.float32):
        raise BLSError("C_data must be Float64 or Float32")
    _require("C_data", C_data, C_data.dtype, st.coupling_storage_length())
    fb = ct.c_int64(0)
    f = getattr(lib(), "bls_reduce_schur_dense_mfma_" + v)
    rc = _check(f(handle.h, _dp(S), _dp(g_S), _dp(b_r), _dp(b_e), _dp(scaling_e), float(lam), _dp(E_data), _dp(C_data),
                  _dp(ws.E_inv_data), _dp(ws.z_data), int(nslices), ct.byref(fb), _stream(stream)))
    ws.failed_block = int(fb.value)
    return rc == BLS_OK


def _reduce(handle, dense, S, g, b_r, b_e, sc, lam, E, C, ws, s_r, s_e, stream, sync):
    L = lib()
    v = _variant(S.dtype, E.dtype)
    if C.dtype != E.dtype or ws.variant != v:
        raise BLSError(f"the blocks ({E.dtype}, {C.dtype}), S ({S.dtype}) and the workspace ({ws.variant}) disagree")
    kind = "dense" if dense else "sparse"
    if sync:
        fb = ct.c_int64(0)
        f = getattr(L, f"bls_reduce_schur_{kind}_{v}")
        rc = _check(f(handle.h, _dp(S), _dp(g), _dp(b_r), _dp(b_e), _dp(sc), float(lam), _dp(E), _dp(C),
                      _dp(ws.E_inv_data), _dp(ws.z_data), _dp(s_r), _dp(s_e), ct.byref(fb), _stream(stream)))
        ws.failed_block = int(fb.value)
        return rc == BLS_OK
    if dense and v != "f64":
        raise BLSError("the asynchronous dense reduction is Float64 only")
    f = getattr(L, f"bls_reduce_schur_{kind}_{v}_async")
    _check(f(handle.h, _dp(S), _dp(g), _dp(b_r), _dp(b_e), _dp(sc), float(lam), _dp(E), _dp(C), _dp(ws.E_inv_data),
             _dp(ws.z_data), _dp(s_r), _dp(s_e), _dp(ws.d_failed), _stream(stream)))
    return None


def _require(name, t, dtype, n, optional=False):
    """Check, before a launch, that a vector argument has the element type and length the kernels index it with
    (a wrong one would make them read or write past the buffer); raises BLSError otherwise."""
    if t is None:
        if optional:
            return
        raise BLSError(f"{name} is required")
    if t.dtype != dtype:
        raise BLSError(f"{name} must be {dtype}, got {t.dtype}")
    if t.numel() != n:
        raise BLSError(f"{name} has length {t.numel()}, expected {n}")


def _check_backsub(st, variant, x_e, x_r, C_data, s_r, s_e):
    import torch
    _require("x_e", x_e, torch.float64, st.ne)
    _require("x_r", x_r, torch.float64, st.nr)
    _require("C_data", C_data, torch.float64 if variant == "f64" else torch.float32, st.coupling_storage_length())
    _require("s_r", s_r, torch.float64, st.nr, optional=True)
    _require("s_e", s_e, torch.float64, st.ne, optional=True)


def axpy_(y, alpha, x, stream=None):
    """y += alpha x (Float64 device vectors)."""
    import torch
    _require("x", x, torch.float64, y.numel())
    _check(lib().bls_axpy_f64(y.numel(), float(alpha), _dp(x), _dp(y), _stream(stream)))
    return y


def factor_eliminated_(handle, ws, E_data, b_e, scaling_e, lam, s_e=None, stream=None, check=True):
    """factor_all_eliminated! (src/schur_reduction.jl:439-455): E_inv and z of every eliminated block into the
    Float64 workspace ws; returns True, or False with ws.failed_block set. check=False skips the host-side
    check_damping of scaling_e (a device reduction and a synchronization) for a caller that validated it."""
    import torch
    if check:
        _check_damping(lam, scaling_e, "scaling_e")
    st = handle.st
    if ws.variant != "f64":
        raise BLSError("factor_eliminated_ is Float64 only")
    _require("E_data", E_data, torch.float64, st.eliminated_storage_length())
    _require("b_e", b_e, torch.float64, st.ne)
    fb = ct.c_int64(0)
    rc = _check(lib().bls_schur_factor_f64(handle.h, _dp(E_data), _dp(b_e), _dp(scaling_e), float(lam),
                                           _dp(ws.E_inv_data), _dp(ws.z_data), _dp(s_e), ct.byref(fb), _stream(stream)))
    ws.failed_block = int(fb.value)
    return rc == BLS_OK


def retained_residual_(out, handle, b_r, x_e, C_data, s_r=None, s_e=None, stream=None):
    """out = b_r .* s_r - sum_c scaled(C_c)' x_e per retained block (fixed-order sums): g_S when x_e = z, and
    g_S - (S_red - R_damped) x_r when x_e = backsubstitute!(x_r)."""
    import torch
    st = handle.st
    _require("out", out, torch.float64, st.nr)
    _require("b_r", b_r, torch.float64, st.nr)
    _require("x_e", x_e, torch.float64, st.ne)
    if C_data.dtype not in (torch.float64, torch.float32) or C_data.numel() != st.coupling_storage_length():
        raise BLSError("C_data must be Float64 or Float32 of the structure's coupling storage length")
    f = lib().bls_schur_retained_residual_f64 if C_data.dtype == torch.float64 else lib().bls_schur_retained_residual_f32
    _check(f(handle.h, _dp(out), _dp(b_r), _dp(x_e), _dp(C_data), _dp(s_r), _dp(s_e), _stream(stream)))
    return out


def reduce_schur_sparse_noalloc_(nzval, g_S, b_r, b_e, scaling_e, lam, handle, E_data, C_data, ws, stream=None):
    """The Float64 sparse reduction through bls_reduce_schur_sparse_f64_noalloc: the entry point that allocates no
    device memory (its scratch is the library's static array; no conditioning scales). The status is left on the
    device in ws.d_failed, as with sync=False."""
    import math
    if not (math.isfinite(lam) and lam >= 0):
        raise ValueError(f"lambda must be finite and non-negative, got {lam}")
    if ws.variant != "f64" or E_data.dtype != C_data.dtype or nzval.dtype != E_data.dtype:
        raise BLSError("the entry point that allocates nothing is Float64 only")
    _check(lib().bls_reduce_schur_sparse_f64_noalloc(handle.h, _dp(nzval), _dp(g_S), _dp(b_r), _dp(b_e),
                                                      _dp(scaling_e), float(lam), _dp(E_data), _dp(C_data),
                                                      _dp(ws.E_inv_data), _dp(ws.z_data), _dp(ws.d_failed),
                                                      _stream(stream)))


def backsubstitute_(x_e, x_r, handle, C_data, ws, s_r=None, s_e=None, stream=None):
    """backsubstitute! (src/schur_reduction.jl:519-576)."""
    _check_backsub(handle.st, ws.variant, x_e, x_r, C_data, s_r, s_e)
    f = getattr(lib(), "bls_backsubstitute_schur_" + ws.variant)
    _check(f(handle.h, _dp(x_e), _dp(x_r), _dp(C_data), _dp(ws.E_inv_data), _dp(ws.z_data), _dp(s_r), _dp(s_e),
             _stream(stream)))


def copyto_(dest, src, scatter_map, stream=None):
    """copyto!(dest, src, scatter_map) (src/SparseBlockLinearSystem.jl:356-362): dest[i] = src[scatter_map[i]].
    ``scatter_map``: a GPU int64 tensor of 1-based indices (see scatter_map_to_device)."""
    import torch
    if dest.numel() != scatter_map.numel():
        raise ValueError(f"Destination nzval length ({dest.numel()}) does not match scatter map length ({scatter_map.numel()})")
    if scatter_map.dtype != torch.int64:
        raise BLSError("scatter map must be int64")
    if dest.dtype != src.dtype or dest.dtype not in (torch.float64, torch.float32):
        raise BLSError("copyto!: dest and src must both be Float64 or both Float32")
    f = lib().bls_copyto_csc_f64 if dest.dtype == torch.float64 else lib().bls_copyto_csc_f32
    _check(f(dest.numel(), _dp(dest), _dp(src), _dp(scatter_map), _stream(stream)))
    return dest


def scatter_map_to_device(scatter_map):
    import torch
    return torch.as_tensor(np.ascontiguousarray(scatter_map, dtype=np.int64)).cuda()


class JacobianFill:
    """The row blocks of a SchurJacobianStructure built from a description (the C ABI's bls_jfill), for filling its
    tables on the device: block(J, k, blockentry(jst, k, j)) .= J_kj and residual_block(J, k) .= r_k for every row
    block k (src/schur_jacobian.jl:330-387). ``nvals`` / ``nres``: the lengths of the value and residual arrays
    (row block after row block, its blocks in ascending column block, each m x d column-major)."""

    def __init__(self, jst):
        if jst.rowblocks is None:
            raise BLSError("the structure has no row blocks (build it with SchurJacobianStructure.from_description)")
        keep = []
        st, rb = jst.schur, jst.rowblocks
        h = VP()
        nv, nr = ct.c_int64(0), ct.c_int64(0)
        _check(lib().bls_jfill_create(st.nrblocks, st.neblocks, _p64(st.r_dims, keep), _p64(st.e_dims, keep),
                                      jst.ntables, _p64(jst.table_nrows, keep), _p64(jst.table_ncols, keep),
                                      _p64(jst.table_top, keep), _p64(jst.raw_offsets, keep),
                                      _p64(jst.residual_offsets, keep), _p64(jst.table_ptr, keep),
                                      _p64(jst.entry_retained, keep), _p64(jst.entry_col, keep), len(rb["table"]),
                                      _p64(rb["table"], keep), _p64(rb["first"], keep), _p64(rb["nrows"], keep),
                                      _p64(rb["ptr"], keep), _p64(rb["blocks"], keep), ct.byref(nv), ct.byref(nr),
                                      ct.byref(h)))
        self.h, self.jst, self.nvals, self.nres = h, jst, int(nv.value), int(nr.value)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_jfill_destroy(self.h)
            self.h = None


def fill_jacobian_(raw, r, fill, vals, res, stream=None):
    """Write every row block's Jacobian blocks ``vals`` and residual rows ``res`` into the raw tables ``raw``
    (Float64 or Float32, as ``vals``) and the Float64 residual ``r`` (device tensors)."""
    import torch
    jst = fill.jst
    if vals is None or vals.dtype not in (torch.float64, torch.float32):
        raise BLSError("vals must be a Float64 or Float32 device tensor")
    _require("vals", vals, vals.dtype, fill.nvals)
    _require("raw", raw, vals.dtype, jst.raw_length())
    _require("res", res, torch.float64, fill.nres)
    _require("r", r, torch.float64, jst.residual_length())
    f = lib().bls_jfill_f64 if vals.dtype == torch.float64 else lib().bls_jfill_f32
    _check(f(fill.h, _dp(raw), _dp(r), _dp(vals), _dp(res), _stream(stream)))


class QRHandle:
    """A registered SchurJacobianStructure on the device (the C ABI's bls_qr)."""

    def __init__(self, jst):
        keep = []
        st = jst.schur
        h = VP()
        _check(lib().bls_qr_create(st.nr, st.nrblocks, st.neblocks, _p64(st.e_dims, keep), _p64(st.boffsets_e, keep),
                                   jst.ntables, _p64(jst.table_nrows, keep), _p64(jst.table_ncols, keep),
                                   _p64(jst.raw_offsets, keep), _p64(jst.residual_offsets, keep), ct.byref(h)))
        self.h = h
        self.jst = jst
        _check(lib().bls_qr_set_reduced(h, _p64(jst.table_top, keep), _p64(jst.reduced_offsets, keep),
                                        _p64(jst.table_ptr, keep), len(jst.entry_retained),
                                        _p64(jst.entry_retained, keep), _p64(jst.entry_col, keep),
                                        _p64(jst.entry_table, keep), _p64(jst.retained_ptr, keep),
                                        _p64(jst.retained_entry, keep), _p64(st.r_dims, keep),
                                        _p64(st.boffsets_r, keep), _p64(jst.doffsets, keep)))

    def set_sparse(self, pattern):
        """Register the sparse reduced-system layout (build_schur_pattern's result) for reduced_normal_sparse!."""
        keep = []
        jst = self.jst
        _check(lib().bls_qr_set_sparse(self.h, pattern.nnz, _p64(pattern.S_colptr, keep), len(pattern.block_rowval),
                                       _p64(pattern.block_colptr, keep), _p64(pattern.block_rowval, keep),
                                       _p64(pattern.block_row_offsets, keep), _p64(pattern.col_heights, keep),
                                       _p64(jst.table_ptr, keep), _p64(jst.entry_retained, keep)))
        self.pattern = pattern

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_qr_destroy(self.h)
            self.h = None

    def set_timing(self, on=True):
        _check(lib().bls_qr_set_timing(self.h, 1 if on else 0))

    def kernel_time(self):
        v = ct.c_float(0)
        _check(lib().bls_qr_kernel_time(self.h, ct.byref(v)))
        return v.value


class QRReductionWorkspace:
    """Device counterpart of QRReductionWorkspace{T} (src/qr_reduction.jl:11-23): the factored tables (T, Float64
    or Float32), the Float64 factored residual and the failure flag."""

    def __init__(self, handle, dtype=None):
        import torch
        jst = handle.jst
        self.structure = jst
        self.factored = torch.empty(jst.raw_length(), dtype=dtype or torch.float64, device="cuda")
        self.fres = torch.empty(jst.residual_length(), dtype=torch.float64, device="cuda")
        self.d_failed = torch.zeros(1, dtype=torch.int64, device="cuda")
        self.failed_block = 0


def reduce_qr_(ws, handle, J_data, J_r, scaling_e, lam, s=None, stream=None, sync=True, check=True):
    """reduce_qr!(ws, J, scaling_e, lambda, s) (src/qr_reduction.jl:162-187); J given as its raw tables and Float64
    residual (SchurBlockJacobian's data and r)."""
    jst = handle.jst
    if ws.structure is not jst:
        raise ValueError("The Jacobian and the workspace were built for different structures")
    if scaling_e.numel() != jst.schur.ne:
        raise ValueError(f"scaling_e has length {scaling_e.numel()}, expected {jst.schur.ne}")
    if check:
        _check_damping(lam, scaling_e, "scaling_e")
    L = lib()
    if J_data.dtype != ws.factored.dtype:
        raise BLSError(f"the tables ({J_data.dtype}) and the workspace ({ws.factored.dtype}) disagree")
    sfx = _fsfx(J_data)
    if sync:
        fb = ct.c_int64(0)
        rc = _check(getattr(L, "bls_reduce_qr_" + sfx)(handle.h, _dp(ws.factored), _dp(ws.fres), _dp(J_data), _dp(J_r), _dp(scaling_e),
                                        float(lam), _dp(s), ct.byref(fb), _stream(stream)))
        ws.failed_block = int(fb.value)
        return rc == BLS_OK
    _check(getattr(L, f"bls_reduce_qr_{sfx}_async")(handle.h, _dp(ws.factored), _dp(ws.fres), _dp(J_data), _dp(J_r), _dp(scaling_e),
                                     float(lam), _dp(s), _dp(ws.d_failed), _stream(stream)))
    return None


class PCGHandle:
    """Block-Jacobi PCG on reduced systems in build_schur_pattern's layout, batched over independent problems
    (prob_block_ptr: 1-based first retained block of each problem, plus nrblocks + 1)."""

    def __init__(self, st, pattern, prob_block_ptr=None):
        keep = []
        pbp = np.array([1, st.nrblocks + 1], dtype=np.int64) if prob_block_ptr is None else np.asarray(prob_block_ptr, np.int64)
        h = VP()
        _check(lib().bls_pcg_create(st.nr, st.nrblocks, _p64(st.r_dims, keep), _p64(st.boffsets_r, keep), pattern.nnz,
                                    _p64(pattern.S_colptr, keep), len(pattern.block_rowval), _p64(pattern.block_colptr, keep),
                                    _p64(pattern.block_rowval, keep), _p64(pattern.block_row_offsets, keep),
                                    _p64(pattern.col_heights, keep), len(pbp) - 1, _p64(pbp, keep), ct.byref(h)))
        self.h = h
        self.nprob = len(pbp) - 1
        self.st = st

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_pcg_destroy(self.h)
            self.h = None

    def factor_size(self):
        nv, nf = ct.c_int64(0), ct.c_int64(0)
        _check(lib().bls_pcg_values_size(self.h, ct.byref(nv), ct.byref(nf)))
        return int(nf.value)

    def relres(self, S_nzval, g, x, stream=None):
        """||S x - g|| / ||g|| per problem (FP64, on the device), as a numpy array."""
        out = (ct.c_double * self.nprob)()
        _check(lib().bls_pcg_relres_f64(self.h, _dp(S_nzval), _dp(g), _dp(x), out, _stream(stream)))
        return np.array(out[:])

    def group_blocks(self, S_nzval, omega, group, stream=None):
        """The group-corrected preconditioner blocks (device Float64) for groups of `group` consecutive retained blocks
        and coupling weights omega (Float32 per retained block, read at each group's first block)."""
        import torch
        _check(lib().bls_pcg_set_groups(self.h, int(group)))
        n = ct.c_int64(0)
        _check(lib().bls_pcg_group_blocks_size(self.h, ct.byref(n)))
        M = torch.empty(int(n.value), dtype=torch.float64, device="cuda")
        _require("omega", omega, torch.float32, self.st.nrblocks)
        _check(lib().bls_pcg_group_blocks_f64(self.h, _dp(S_nzval), _dp(omega.contiguous()), _dp(M), _stream(stream)))
        return M

    def solve_groups(self, S_nzval, g, x, group_blocks, tol=1e-8, maxit=10000, stream=None):
        """solve() with the group-corrected preconditioner (group_blocks from group_blocks(), same group size)."""
        iters = np.zeros(self.nprob, dtype=np.int64)
        rel = np.zeros(self.nprob)
        rc = _check(lib().bls_pcg_solve_groups_f64(self.h, _dp(S_nzval), _dp(g), _dp(x), float(tol), int(maxit),
                                                   _dp(group_blocks), iters.ctypes.data_as(P64),
                                                   rel.ctypes.data_as(ct.POINTER(ct.c_double)), _stream(stream)))
        self.last_iters = iters
        return rc == BLS_OK, iters, rel

    def residual(self, S_nzval, t, x, gref, r, stream=None):
        """r = t - S x on the device, and ||r|| / ||gref|| per problem as a numpy array."""
        out = (ct.c_double * self.nprob)()
        _check(lib().bls_pcg_residual_f64(self.h, _dp(S_nzval), _dp(t), _dp(x), _dp(gref), _dp(r), out,
                                          _stream(stream)))
        return np.array(out[:])

    def solve(self, S_nzval, g, x, tol=1e-8, maxit=10000, precond_blocks=None, stream=None):
        """x = S^-1 g per problem to ||S x - g|| <= tol ||g||. Returns (all_converged, iters[nprob], relres[nprob])."""
        iters = np.zeros(self.nprob, dtype=np.int64)
        rel = np.zeros(self.nprob)
        rc = _check(lib().bls_pcg_solve_f64(self.h, _dp(S_nzval), _dp(g), _dp(x), float(tol), int(maxit),
                                            _dp(precond_blocks), iters.ctypes.data_as(P64),
                                            rel.ctypes.data_as(ct.POINTER(ct.c_double)), _stream(stream)))
        self.last_iters = iters
        return rc == BLS_OK, iters, rel


# ---- products with the reduced Jacobian J~ (src/qr_reduced.jl:161-243, :484-512) ------------------------------
def _fsfx(t):
    import torch
    if t.dtype == torch.float64:
        return "f64"
    if t.dtype == torch.float32:
        return "f32"
    raise BLSError(f"tables must be Float64 or Float32, got {t.dtype}")


def reduced_mul_(y, handle, factored, v, stream=None):
    """reduced_mul!: y = J~ v."""
    _check(getattr(lib(), "bls_reduced_mul_" + _fsfx(factored))(handle.h, _dp(y), _dp(factored), _dp(v), _stream(stream)))


def reduced_tmul_(g, handle, factored, w, stream=None):
    """reduced_tmul!: g = J~' w."""
    _check(getattr(lib(), "bls_reduced_tmul_" + _fsfx(factored))(handle.h, _dp(g), _dp(factored), _dp(w), _stream(stream)))


def reduced_residual_(rt, handle, fres, stream=None):
    """reduced_residual!: r~."""
    _check(lib().bls_reduced_residual_f64(handle.h, _dp(rt), _dp(fres), _stream(stream)))


def reduced_diagonal_blocks_(D, handle, factored, stream=None):
    """reduced_diagonal_blocks!: the diagonal blocks of J~'J~ (full, column-major, at doffsets)."""
    if D.dtype != factored.dtype:
        raise BLSError("D and the factored tables must have one element type")
    _check(getattr(lib(), "bls_reduced_diagonal_blocks_" + _fsfx(factored))(handle.h, _dp(D), _dp(factored), _stream(stream)))


# ---- quantities of the full problem from the raw (or factored) tables (src/qr_reduced.jl:574-645) -------------
def jtr_(g, handle, tables, r, stream=None):
    """jtr!(g, J): g = J'r, [g_r; g_e] (Float64), from the tables (J.data or ws.factored) and their Float64 residual."""
    _check(getattr(lib(), "bls_jtr_" + _fsfx(tables))(handle.h, _dp(g), _dp(tables), _dp(r), _stream(stream)))
    return g


def jtj_diag_(d, handle, tables, stream=None):
    """jtj_diag!(d, J): d = diag(J'J), the squared column norms, [d_r; d_e] (Float64)."""
    _check(getattr(lib(), "bls_jtj_diag_" + _fsfx(tables))(handle.h, _dp(d), _dp(tables), _stream(stream)))
    return d


def jtj_quadform(handle, tables, v, stream=None):
    """jtj_quadform(J, v): ||J v||^2, a Python float."""
    import torch
    out = torch.empty(1, dtype=torch.float64, device=v.device)
    _check(getattr(lib(), "bls_jtj_quadform_" + _fsfx(tables))(handle.h, _dp(tables), _dp(v), _dp(out),
                                                               _stream(stream)))
    return float(out.item())


def build_reduced_jacobian_pattern(handle):
    """build_reduced_jacobian_pattern(jst) (src/qr_reduced.jl:289-310): (colptr, rowval), 1-based numpy Int64, of J~'s
    CSC (reduced_nrows x nr)."""
    cnt = np.zeros(1, dtype=np.int64)
    _check(lib().bls_reduced_jacobian_nnz(handle.h, cnt.ctypes.data_as(P64)))
    nnz = int(cnt[0])
    colptr = np.zeros(handle.jst.schur.nr + 1, dtype=np.int64)
    rowval = np.zeros(max(nnz, 1), dtype=np.int64)
    _check(lib().bls_build_reduced_jacobian_pattern(handle.h, colptr.ctypes.data_as(P64), rowval.ctypes.data_as(P64)))
    return colptr, rowval[:nnz]


def check_reduced_jacobian_pattern(handle, m, n, colptr, rowval):
    """check_reduced_jacobian_pattern (src/qr_reduced.jl:312-332): raises BLSError with the reference's reason."""
    keep = []
    _check(lib().bls_check_reduced_jacobian_pattern(handle.h, m, n, _p64(colptr, keep), _p64(rowval, keep)))


def reduced_jacobian_(Jt, rt, handle, ws, sparse, stream=None):
    """reduced_jacobian!(Jt, rt, ws) (src/qr_reduced.jl:247-270, :334-370): J~ into Jt (sparse: the nzval of
    build_reduced_jacobian_pattern, checked by the caller; dense: reduced_nrows x nr column-major, flat) and
    rt = r~ (Float64)."""
    if Jt.dtype != ws.factored.dtype:
        raise BLSError("Jt and the factored tables must have one element type")
    name = "bls_reduced_jacobian_%s_%s" % ("sparse" if sparse else "dense", _fsfx(ws.factored))
    _check(getattr(lib(), name)(handle.h, _dp(Jt), _dp(rt), _dp(ws.factored), _dp(ws.fres), _stream(stream)))
    return Jt


def _check_qr_backsub(jst, ws, x_e, x_r):
    import torch
    if ws.structure is not jst:
        raise BLSError("the workspace was built for another structure")
    _require("x_e", x_e, torch.float64, jst.schur.ne)
    _require("x_r", x_r, torch.float64, jst.schur.nr)
    if ws.factored.dtype not in (torch.float64, torch.float32) or ws.factored.numel() != jst.raw_length():
        raise BLSError("the workspace's factored tables do not match the structure")


def qr_backsubstitute_(x_e, x_r, handle, ws, stream=None):
    """backsubstitute!(x_e, x_r, ws::QRReductionWorkspace) (src/qr_reduced.jl:24-110): x_e = U^-1 (q - B x_r)."""
    _check_qr_backsub(handle.jst, ws, x_e, x_r)
    _check(getattr(lib(), "bls_qr_backsubstitute_" + _fsfx(ws.factored))(handle.h, _dp(x_e), _dp(x_r),
                                                                         _dp(ws.factored), _dp(ws.fres),
                                                                         _stream(stream)))


class DiagonalBlocks:
    """The block sizes of a flat array of square blocks on the device (the C ABI's bls_diag), for
    factor_diagonal_blocks! / solve_diagonal_blocks! (src/diagonal_blocks.jl:6-61)."""

    def __init__(self, dims):
        keep = []
        h = VP()
        dims = np.ascontiguousarray(dims, dtype=np.int64)
        _check(lib().bls_diag_create(len(dims), _p64(dims, keep), ct.byref(h)))
        self.h = h
        self.dims = dims

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_diag_destroy(self.h)
            self.h = None


class AssemblyHandle:
    """The fill pattern of a Schur Hessian's addblock! calls for a batch of row blocks (the C ABI's bls_asm,
    src/schur_assembly.jl:140-246). st: a SchurStructure; rowblocks: per row block, its 1-based layout block numbers
    in its Jacobian's column order (or the pair (rb_ptr, rb_blocks) of that list in 1-based CSR); rb_rows: its row
    counts; R: dense (R_layout None) or the stored lower blocks of a
    SparseBlockLinearSystem as (colptr, rowval, offsets), 1-based."""

    def __init__(self, st, rowblocks, rb_rows, R_layout=None):
        keep = []
        h = VP()
        i64 = lambda a: np.ascontiguousarray(a, dtype=np.int64)
        if isinstance(rowblocks, tuple):   # (rb_ptr, rb_blocks): CSR, 1-based, as the C ABI takes them
            rb_ptr, rb_blocks = i64(rowblocks[0]), i64(rowblocks[1])
            rowblocks = range(len(rb_ptr) - 1)
        else:
            rb_ptr = i64(np.concatenate([[1], 1 + np.cumsum([len(b) for b in rowblocks])]))
            rb_blocks = i64(np.concatenate([np.asarray(b, dtype=np.int64) for b in rowblocks]) if len(rowblocks) else [])
        rb_rows = i64(rb_rows)
        rc_, rr_, ro_ = (None, None, None) if R_layout is None else (i64(x) for x in R_layout)
        p = lambda a: _p64(a, keep) if a is not None else None
        _check(lib().bls_asm_create(st.nrblocks, st.neblocks, p(i64(st.r_dims)), p(i64(st.e_dims)),
                                    p(i64(st.boffsets_r)), p(i64(st.boffsets_e)), p(i64(st.eoffsets)),
                                    p(i64(st.eliminated_ptr)), p(i64(st.entry_retained)), p(i64(st.coffsets)),
                                    p(rc_), p(rr_), p(ro_), len(rowblocks), p(rb_ptr), p(rb_blocks), p(rb_rows),
                                    ct.byref(h)))
        self.h = h
        lens = np.zeros(4, dtype=np.int64)
        _check(lib().bls_asm_lengths(h, p(lens)))
        self.E_len, self.C_len, self.R_len, self.nr = (int(x) for x in lens)
        self.dense = R_layout is None

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_asm_destroy(self.h)
            self.h = None


def assemble_hessian_(E_data, C_data, R, b_r, b_e, handle, J, r, stream=None):
    """H = sum_k J_k' J_k into E_data, C_data, R and b = sum_k J_k' r_k into b_r, b_e (Float64 device tensors):
    the addblock! calls of the handle's pattern on a zero Hessian, in row-block order."""
    _check(lib().bls_asm_hessian_f64(handle.h, _dp(J), _dp(r), _dp(E_data), _dp(C_data), _dp(R), _dp(b_r), _dp(b_e),
                                     _stream(stream)))
    return True


def hessian_dot(x, H_handle, y, E_data, C_data, R, stream=None):
    """dot(x, H, y) (src/linalg.jl:133-260) over the storage of the handle's layout: a Python float."""
    import torch
    out = torch.empty(1, dtype=torch.float64, device=x.device)
    _check(getattr(lib(), "bls_asm_dot_" + _fsfx(E_data))(H_handle.h, _dp(x), _dp(y), _dp(E_data), _dp(C_data), _dp(R),
                                                          _dp(out), _stream(stream)))
    return float(out.item())


def hessian_diag_(d, H_handle, E_data, R, stream=None):
    """getdiag!(d, ls) of a Schur system (src/schur_linear_systems.jl:239-270): the diagonal of H as filled."""
    _check(getattr(lib(), "bls_asm_diag_" + _fsfx(E_data))(H_handle.h, _dp(E_data), _dp(R), _dp(d), _stream(stream)))
    return d


# ---- the conditioning scale (src/conditioning.jl:24-87) ---------------------------------------------------------
class EliminationConditioning:
    """EliminationConditioning(n, on) (src/conditioning.jl:47-55) on the device: s (ones) and d (zeros), n Float64
    each, with the stale flag and the recomputation count held by the C ABI's bls_cond. `on` is the internal switch:
    an unconditioned solve passes no scale (UnitScale) to the kernels."""

    def __init__(self, n, on=True, device="cuda"):
        import torch
        h = VP()
        _check(lib().bls_cond_create(n, ct.byref(h)))
        self.h = h
        self.n = n
        self.on = on
        self.s = torch.ones(n, dtype=torch.float64, device=device)
        self.d = torch.zeros(n, dtype=torch.float64, device=device)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_cond_destroy(self.h)
            self.h = None

    @property
    def stale(self):
        return bool(_check(lib().bls_cond_is_stale(self.h)))

    def markstale(self):
        """zero! / markstale!: the next solve computes the scale from the diagonal afresh."""
        _check(lib().bls_cond_markstale(self.h))

    def nrescale(self, stream=None):
        out = np.zeros(1, dtype=np.int64)
        _check(lib().bls_cond_nrescale(self.h, out.ctypes.data_as(P64), _stream(stream)))
        return int(out[0])


def prepare_scale_(c, diag_, lam, sc, elem_dtype=None, stream=None):
    """prepare_scale!(c, diag!, lambda, sc, T) (src/conditioning.jl:60-87): the scale for a solve at the damping
    lambda Diagonal(sc) (sc a device Float64 vector, or None for ones); diag_(d) writes the undamped diagonal into
    c.d, and is called only when the scale is stale. elem_dtype: the system's element type (torch.float32 raises the
    scale's floor to floatmin(Float32)). Returns c.s."""
    import torch
    if c.stale:
        diag_(c.d)
    f32 = 1 if elem_dtype == torch.float32 else 0
    _check(lib().bls_prepare_scale_f64(c.h, _dp(c.s), _dp(c.d), _dp(sc), float(lam), f32, _stream(stream)))
    return c.s


class DenseCholesky:
    """The :cholesky branch of solve_reduced_dense! (src/schur_linear_systems.jl:312-319) for a dense reduced system
    of order n (the C ABI's bls_chol): factor(S) -> (True, 0) or (False, first failing column, 1-based);
    solve(g, x) -> x = S^-1 g with the last factor."""

    def __init__(self, n):
        h = VP()
        _check(lib().bls_chol_create(n, ct.byref(h)))
        self.h = h
        self.n = n

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_chol_destroy(self.h)
            self.h = None

    def factor(self, S, stream=None):
        fc = ct.c_int64(0)
        rc = _check(getattr(lib(), "bls_chol_factor_" + _fsfx(S))(self.h, _dp(S), ct.byref(fc), _stream(stream)))
        return rc == BLS_OK, int(fc.value)

    def solve(self, g, x, stream=None):
        _check(lib().bls_chol_solve_f64(self.h, _dp(g), _dp(x), _stream(stream)))
        return x

    def factor_copy(self, U, stream=None):
        _check(lib().bls_chol_factor_copy_f64(self.h, _dp(U), _stream(stream)))
        return U


class DenseLDL:
    """The :ldl branch of solve_reduced_dense! (src/schur_linear_systems.jl:312-319): Bunch-Kaufman of a copy of the
    dense reduced system of order n (the C ABI's bls_ldl). factor(S) -> (True, negative eigenvalues of S) or
    (False, first zero pivot, 1-based); solve(g, x) -> x = S^-1 g with the last factor."""

    def __init__(self, n):
        h = VP()
        _check(lib().bls_ldl_create(n, ct.byref(h)))
        self.h = h
        self.n = n

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.bls_ldl_destroy(self.h)
            self.h = None

    def factor(self, S, stream=None):
        import torch
        _require("S", S, torch.float64, self.n * self.n)
        zc, neg = ct.c_int64(0), ct.c_int64(0)
        rc = _check(lib().bls_ldl_factor_f64(self.h, _dp(S), ct.byref(zc), ct.byref(neg), _stream(stream)))
        return (True, int(neg.value)) if rc == BLS_OK else (False, int(zc.value))

    def solve(self, g, x, stream=None):
        import torch
        _require("g", g, torch.float64, self.n)
        _require("x", x, torch.float64, self.n)
        _check(lib().bls_ldl_solve_f64(self.h, _dp(g), _dp(x), _stream(stream)))
        return x

    def factor_copy(self, A, ipiv, stream=None):
        import torch
        _require("A", A, torch.float64, self.n * self.n)
        _require("ipiv", ipiv, torch.int64, self.n)
        _check(lib().bls_ldl_factor_copy_f64(self.h, _dp(A), _dp(ipiv), _stream(stream)))
        return A, ipiv


def dense_residual_(r, S, t, x, stream=None):
    """r = t - S x for a dense symmetric S stored by its upper triangle (column-major, n x n)."""
    import torch
    n = r.numel()
    _require("S", S, torch.float64, n * n)
    _require("t", t, torch.float64, n)
    _require("x", x, torch.float64, n)
    _check(lib().bls_dense_residual_f64(n, _dp(S), _dp(t), _dp(x), _dp(r), _stream(stream)))
    return r


def unscale_(c, x, s, nonfinite, stream=None):
    """unscale!(x, s) (src/conditioning.jl:90-95) over c's n variables (s None: UnitScale), and the device int32
    nonfinite[0] = 1 if x then holds a value that is not finite."""
    _check(lib().bls_unscale_f64(c.h, _dp(x), _dp(s), _dp(nonfinite), _stream(stream)))
    return x


def factor_diagonal_blocks_(D, blocks, stream=None):
    """factor_diagonal_blocks!(D, dims): True, or False with the lowest failing block as second value."""
    fb = ct.c_int64(0)
    rc = _check(getattr(lib(), "bls_factor_diagonal_blocks_" + _fsfx(D))(blocks.h, _dp(D), ct.byref(fb), _stream(stream)))
    return rc == BLS_OK, int(fb.value)


def solve_diagonal_blocks_(y, D, blocks, x, stream=None):
    """solve_diagonal_blocks!(y, D, dims, x): y = (L L')^-1 x per block (Float64 y, x)."""
    _check(getattr(lib(), "bls_solve_diagonal_blocks_" + _fsfx(D))(blocks.h, _dp(y), _dp(D), _dp(x), _stream(stream)))
    return y


def reduced_normal_sparse_(S_nzval, g_S, handle, ws, stream=None):
    """reduced_normal_sparse!(S, g_S, ws, pattern) (src/qr_reduced.jl:437-480), the pattern of handle.set_sparse."""
    _check(getattr(lib(), "bls_reduced_normal_sparse_" + _fsfx(ws.factored))(handle.h, _dp(S_nzval), _dp(g_S),
                                                                             _dp(ws.factored), _dp(ws.fres),
                                                                             _stream(stream)))
    return True


def reduced_normal_dense_(S, g_S, handle, ws, stream=None):
    """reduced_normal_dense!(S, g_S, ws) (src/qr_reduced.jl:409-435); S column-major flat (nr*nr)."""
    _check(getattr(lib(), "bls_reduced_normal_dense_" + _fsfx(ws.factored))(handle.h, _dp(S), _dp(g_S),
                                                                            _dp(ws.factored), _dp(ws.fres),
                                                                            _stream(stream)))
    return True
