"""ctypes binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; the product library never
loads it. Builds the C structs from the host mirror's layout objects (1-based arrays passed verbatim).
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liborc.so")
P64 = ct.POINTER(ct.c_int64)
PD = ct.POINTER(ct.c_double)
PF = ct.POINTER(ct.c_float)


class OrcSchur(ct.Structure):
    _fields_ = [("nr", ct.c_int64), ("nrblocks", ct.c_int64), ("neblocks", ct.c_int64), ("ncoup", ct.c_int64),
                ("r_dims", P64), ("e_dims", P64), ("boffsets_r", P64), ("boffsets_e", P64), ("eoffsets", P64),
                ("eliminated_ptr", P64), ("entry_retained", P64), ("coffsets", P64),
                ("nranges", ct.c_int64), ("chunk_lo", P64), ("chunk_hi", P64), ("chunk_run", P64),
                ("retained_ptr", P64), ("retained_entry", P64), ("entry_eliminated", P64),
                ("ncol_chunks", ct.c_int64), ("col_lo", P64), ("col_hi", P64)]


class OrcPattern(ct.Structure):
    _fields_ = [("S_colptr", P64), ("S_rowval", P64), ("block_colptr", P64), ("block_rowval", P64),
                ("block_row_offsets", P64), ("col_heights", P64)]


class OrcJacobian(ct.Structure):
    _fields_ = [("ntables", ct.c_int64), ("table_nrows", P64), ("table_ncols", P64), ("table_top", P64),
                ("raw_offsets", P64), ("residual_offsets", P64), ("reduced_offsets", P64), ("doffsets", P64),
                ("table_ptr", P64), ("entry_retained", P64), ("entry_col", P64), ("entry_table", P64),
                ("retained_ptr", P64), ("retained_entry", P64), ("ncol_chunks", ct.c_int64),
                ("col_lo", P64), ("col_hi", P64)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None
NATIVE_FLAGS = ["-O3", "-march=native", "-fPIC", "-std=c11", "-ffp-contract=off", "-fopenmp"]


def use_native(outdir):
    """Compile the same oracle source with -O3 -march=native (still -ffp-contract=off, so every result is the same
    as the -O2 build's) into `outdir` on the host that will run it, and make it the library the functions below
    call. Used only by the CPU-baseline timing; returns the compiler command, or None when the compiler fails (the
    -O2 build then stays in use)."""
    so =os.path.join(outdir, "liborc_native.so")
    cmd = [os.environ.get("CC", "gcc")] + NATIVE_FLAGS + ["-shared", "-o", so, os.path.join(HERE, "bls_oracle.c"), "-lm"]
    try:
        subprocess.run(cmd, check=True, capture_output=True)
    except (OSError, subprocess.CalledProcessError):
        return None
    _bind(ct.CDLL(so))
    return " ".join(cmd[:len(NATIVE_FLAGS) + 1]) + " bls_oracle.c"


def lib():
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _bind(ct.CDLL(LIB))
    return _lib


def _bind(handle):
    global _lib
    for s in ("f64", "f32"):
        getattr(handle, f"orc_reduce_schur_{s}").restype = ct.c_int
        getattr(handle, f"orc_reduce_qr_{s}").restype = ct.c_int
        getattr(handle, f"orc_inv_sym_block_{s}").restype = ct.c_int
        getattr(handle, f"orc_inv_sym_block_chol_{s}").restype = ct.c_int
        getattr(handle, f"orc_factor_diagonal_blocks_{s}").restype = ct.c_int
    handle.orc_pcg_f64.restype = ct.c_int64
    _lib = handle


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(P64)


def _ptr(a, dtype):
    if a is None:
        return None
    assert a.dtype == dtype and a.flags.c_contiguous, (a.dtype, dtype)
    return a.ctypes.data_as(PD if dtype == np.float64 else PF)


class SchurArgs:
    """Keeps the numpy arrays alive for the lifetime of the C structs."""

    def __init__(self, st, ws_index, pattern=None):
        keep = []

        def p(a):
            arr, ptr = _i64(a)
            keep.append(arr)
            return ptr

        lo, hi, run = [], [], []
        for k, grp in enumerate(st.chunks):
            for (a, b) in grp:
                lo.append(a); hi.append(b); run.append(k)
        col_lo = [a for (a, b) in ws_index.col_chunks]
        col_hi = [b for (a, b) in ws_index.col_chunks]
        self.s = OrcSchur(st.nr, st.nrblocks, st.neblocks, st.ncoup, p(st.r_dims), p(st.e_dims), p(st.boffsets_r),
                          p(st.boffsets_e), p(st.eoffsets), p(st.eliminated_ptr), p(st.entry_retained),
                          p(st.coffsets), len(lo), p(lo), p(hi), p(run), p(ws_index.retained_ptr),
                          p(ws_index.retained_entry), p(ws_index.entry_eliminated), len(col_lo), p(col_lo), p(col_hi))
        self.pat = None
        if pattern is not None:
            self.pat = OrcPattern(p(pattern.S_colptr), p(pattern.S_rowval), p(pattern.block_colptr),
                                  p(pattern.block_rowval), p(pattern.block_row_offsets), p(pattern.col_heights))
        self.keep = keep


class JacobianArgs:
    def __init__(self, jst):
        keep = []

        def p(a):
            arr, ptr = _i64(a)
            keep.append(arr)
            return ptr

        self.j = OrcJacobian(jst.ntables, p(jst.table_nrows), p(jst.table_ncols), p(jst.table_top), p(jst.raw_offsets),
                             p(jst.residual_offsets), p(jst.reduced_offsets), p(jst.doffsets), p(jst.table_ptr),
                             p(jst.entry_retained), p(jst.entry_col), p(jst.entry_table), p(jst.retained_ptr),
                             p(jst.retained_entry), len(jst.col_chunks), p([a for a, b in jst.col_chunks]),
                             p([b for a, b in jst.col_chunks]))
        self.keep = keep


def _sfx(dtype):
    return "f64" if np.dtype(dtype) == np.float64 else "f32"


def inv_sym_block(H, damp, chol=False):
    """inv_sym_block(H, one(T), damp): closed forms for d <= 3 (chol=False) or the generic Cholesky method."""
    H = np.asarray(H)
    d = H.shape[0]
    dt = H.dtype
    Hc = np.ascontiguousarray(H.T.reshape(-1))   # column-major flat
    out = np.zeros(d * d, dtype=dt)
    f = getattr(lib(), f"orc_inv_sym_block{'_chol' if chol else ''}_{_sfx(dt)}")
    ok = f(ct.c_int64(d), _ptr(Hc, dt), _ptr(np.ascontiguousarray(damp, dtype=dt), dt), _ptr(out, dt))
    return bool(ok), out.reshape(d, d).T.copy()


def reset_schur_dense(S, R, scaling_r, lam, s_r=None):
    """S, R: column-major flat arrays (length nr*nr) of dtype T."""
    nr = len(scaling_r)
    getattr(lib(), f"orc_reset_schur_dense_{_sfx(S.dtype)}")(ct.c_int64(nr), _ptr(S, S.dtype), _ptr(R, S.dtype),
                                                             _ptr(scaling_r, np.float64), ct.c_double(lam), _ptr(s_r, np.float64))


def reset_schur_sparse(nzval, R_nzval, r_to_s, pattern, scaling_r, lam, s_r=None):
    dt = nzval.dtype
    rts, prts = _i64(r_to_s)
    di, pdi = _i64(pattern.diag_inds)
    cp, pcp = _i64(pattern.S_colptr)
    rv, prv = _i64(pattern.S_rowval)
    getattr(lib(), f"orc_reset_schur_sparse_{_sfx(dt)}")(ct.c_int64(len(nzval)), _ptr(nzval, dt), ct.c_int64(len(R_nzval)),
                                                          _ptr(R_nzval, dt), prts, ct.c_int64(len(scaling_r)), pdi,
                                                          _ptr(scaling_r, np.float64), ct.c_double(lam), _ptr(s_r, np.float64), pcp, prv)


def reduce_schur(sparse, args, S, g_S, b_r, b_e, scaling_e, lam, E_data, C_data, E_inv, z, W, s_r=None, s_e=None,
                 nthreads=1):
    dt = E_data.dtype
    fb = ct.c_int64(0)
    ok = getattr(lib(), f"orc_reduce_schur_{_sfx(dt)}")(
        ct.c_int(1 if sparse else 0), ct.byref(args.s), ct.byref(args.pat) if args.pat is not None else None,
        _ptr(S, dt), _ptr(g_S, np.float64), _ptr(b_r, np.float64), _ptr(b_e, np.float64), _ptr(scaling_e, np.float64),
        ct.c_double(lam), _ptr(E_data, dt), _ptr(C_data, dt), _ptr(E_inv, dt), _ptr(z, np.float64), _ptr(W, dt),
        _ptr(s_r, np.float64), _ptr(s_e, np.float64), ct.byref(fb), ct.c_int(nthreads))
    return bool(ok), int(fb.value)


def backsubstitute_schur(args, x_e, x_r, C_data, E_inv, z, s_r=None, s_e=None):
    dt = C_data.dtype
    getattr(lib(), f"orc_backsubstitute_schur_{_sfx(dt)}")(ct.byref(args.s), _ptr(x_e, np.float64), _ptr(x_r, np.float64),
                                                            _ptr(C_data, dt), _ptr(E_inv, dt), _ptr(z, np.float64),
                                                            _ptr(s_r, np.float64), _ptr(s_e, np.float64))


def copyto(dest, src, scatter_map):
    dt = dest.dtype
    m, pm = _i64(scatter_map)
    getattr(lib(), f"orc_copyto_{_sfx(dt)}")(ct.c_int64(len(dest)), _ptr(dest, dt), _ptr(src, dt), pm)


def reduce_qr(sargs, jargs, F, Fr, Rw, Rr, scaling_e, lam, cs=None, nthreads=1):
    dt = F.dtype
    fb = ct.c_int64(0)
    ok = getattr(lib(), f"orc_reduce_qr_{_sfx(dt)}")(ct.byref(sargs.s), ct.byref(jargs.j), _ptr(F, dt), _ptr(Fr, np.float64),
                                                      _ptr(Rw, dt), _ptr(Rr, np.float64), _ptr(scaling_e, np.float64),
                                                      ct.c_double(lam), _ptr(cs, np.float64), ct.byref(fb), ct.c_int(nthreads))
    return bool(ok), int(fb.value)


def qr_backsubstitute(sargs, jargs, x_e, x_r, F, Fr):
    dt = F.dtype
    getattr(lib(), f"orc_qr_backsubstitute_{_sfx(dt)}")(ct.byref(sargs.s), ct.byref(jargs.j), _ptr(x_e, np.float64),
                                                         _ptr(x_r, np.float64), _ptr(F, dt), _ptr(Fr, np.float64))


def reduced_normal(sparse, sargs, jargs, S, g_S, F, Fr):
    dt = F.dtype
    getattr(lib(), f"orc_reduced_normal_{_sfx(dt)}")(ct.c_int(1 if sparse else 0), ct.byref(sargs.s), ct.byref(jargs.j),
                                                      ct.byref(sargs.pat) if sargs.pat is not None else None,
                                                      _ptr(S, dt), _ptr(g_S, np.float64), _ptr(F, dt), _ptr(Fr, np.float64))


def reduced_diagonal_blocks(sargs, jargs, D, F):
    dt = F.dtype
    getattr(lib(), f"orc_reduced_diagonal_blocks_{_sfx(dt)}")(ct.byref(sargs.s), ct.byref(jargs.j), _ptr(D, dt), _ptr(F, dt))


def reduced_mul(sargs, jargs, y, F, v):
    getattr(lib(), f"orc_reduced_mul_{_sfx(F.dtype)}")(ct.byref(sargs.s), ct.byref(jargs.j), _ptr(y, np.float64), _ptr(F, F.dtype), _ptr(v, np.float64))


def reduced_tmul(sargs, jargs, g, F, w):
    getattr(lib(), f"orc_reduced_tmul_{_sfx(F.dtype)}")(ct.byref(sargs.s), ct.byref(jargs.j), _ptr(g, np.float64), _ptr(F, F.dtype), _ptr(w, np.float64))


def reduced_residual(jargs, rt, Fr):
    lib().orc_reduced_residual_f64(ct.byref(jargs.j), _ptr(rt, np.float64), _ptr(Fr, np.float64))


def factor_diagonal_blocks(dims, D):
    d, pd = _i64(dims)
    return bool(getattr(lib(), f"orc_factor_diagonal_blocks_{_sfx(D.dtype)}")(ct.c_int64(len(dims)), pd, _ptr(D, D.dtype)))


def solve_diagonal_blocks(dims, y, D, x):
    d, pd = _i64(dims)
    getattr(lib(), f"orc_solve_diagonal_blocks_{_sfx(D.dtype)}")(ct.c_int64(len(dims)), pd, _ptr(y, np.float64), _ptr(D, D.dtype), _ptr(x, np.float64))


def sym_spmv(st, args, nzval, x):
    y = np.zeros(st.nr)
    rd, prd = _i64(st.r_dims)
    bo, pbo = _i64(st.boffsets_r)
    lib().orc_sym_spmv_f64(ct.c_int64(st.nr), ct.c_int64(st.nrblocks), prd, pbo, ct.byref(args.pat), _ptr(nzval, np.float64),
                           _ptr(x, np.float64), _ptr(y, np.float64))
    return y


def pcg(st, args, nzval, g, tol=1e-8, maxit=10000, precond_blocks=None):
    x = np.zeros(st.nr)
    rd, prd = _i64(st.r_dims)
    bo, pbo = _i64(st.boffsets_r)
    rel = ct.c_double(0)
    it = lib().orc_pcg_f64(ct.c_int64(st.nr), ct.c_int64(st.nrblocks), prd, pbo, ct.byref(args.pat), _ptr(nzval, np.float64),
                           _ptr(g, np.float64), _ptr(x, np.float64), ct.c_double(tol), ct.c_int64(maxit),
                           _ptr(precond_blocks, np.float64), ct.byref(rel))
    return x, int(it), float(rel.value)


def diag_blocks_from_sparse(st, args, nzval):
    D = np.zeros(int((st.r_dims ** 2).sum()))
    rd, prd = _i64(st.r_dims)
    bo, pbo = _i64(st.boffsets_r)
    lib().orc_diag_blocks_from_sparse_f64(ct.c_int64(st.nrblocks), prd, pbo, ct.byref(args.pat), _ptr(nzval, np.float64), _ptr(D, np.float64))
    return D


def schur_reduce_problem(p, dtype=np.float64, nthreads=1, lam=None):
    """Run reset + reduce on a generated HessianProblem; returns dict with S (nzval or dense col-major), g_S,
    E_inv, z, ok, failed_block."""
    st = p.structure
    lam = p.lam if lam is None else lam
    args = SchurArgs(st, p.ws_index, p.pattern)
    E = p.E_data.astype(dtype)
    C = p.C_data.astype(dtype)
    E_inv = np.zeros(len(E), dtype=dtype)
    z = np.zeros(st.ne)
    W = np.zeros(len(C), dtype=dtype)
    g = np.zeros(st.nr)
    if p.dense:
        S = np.zeros(st.nr * st.nr, dtype=dtype)
        R = np.asfortranarray(p.R).T.reshape(-1).astype(dtype)  # column-major flat
        reset_schur_dense(S, R, p.scaling_r, lam)
    else:
        S = np.zeros(p.pattern.nnz, dtype=dtype)
        reset_schur_sparse(S, p.R_data.astype(dtype), p.r_to_s, p.pattern, p.scaling_r, lam)
    ok, fb = reduce_schur(not p.dense, args, S, g, p.b_r, p.b_e, p.scaling_e, lam, E, C, E_inv, z, W, nthreads=nthreads)
    return dict(S=S, g=g, E_inv=E_inv, z=z, W=W, ok=ok, failed_block=fb, args=args)
