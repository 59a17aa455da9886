"""Seeded synthetic bundle-adjustment instances for the elimination path.

RNG: numpy's PCG64 (``numpy.random.default_rng(seed)``); normals from ``Generator.standard_normal``. Every
instance is described in the caller's numbering (cameras 1..ncam, then landmarks) and laid out by
SchurStructure.from_description / SchurJacobianStructure.from_description, exactly as a caller of the reference
would (src/schur_description.jl:97-101, 137-189); the Hessian blocks are then filled as ``addblock!`` would
(src/schur_assembly.jl:140-246): E_e = sum J_l'J_l, C_er = J_l'J_c stored d_e x d_r column-major, R's diagonal
blocks sum J_c'J_c, b = J'r.

Configurations (sizes may be scaled down for tests with ``scale``):
  cfg1  10 cams (d 6), 200 landmarks (d 3), k ~ U{3..6} distinct cams, dense R, lambda 1e-3
  cfg2  1000 cams, 200000 landmarks, k = clip(Geometric(0.2), 2, 40), contiguous camera window, lambda 1e-4
  cfg3  256 cams, 1000000 landmarks, k ~ U{2..64} distinct cams, J scaled per camera by 10^U(-2, 2)
  cfg4  500 cams of dim 6 or 9 (50/50), 500000 landmarks, k ~ U{2..30}; Jacobian tables
  cfg5  64 problems: cams U{50..2000}, landmarks U{10000..300000}, tracks as cfg2, lambda 1e-4
"""
from __future__ import annotations

import numpy as np

from .structure import (I64, SchurJacobianStructure, SchurReductionWorkspaceIndex, SchurStructure,
                        build_r_to_s_map, build_schur_pattern, hessian_retained_storage)


def _tracks_uniform_random(rng, ncam, nland, kmin, kmax):
    """k ~ U{kmin..kmax} distinct cameras per landmark, uniform without replacement (0-based cams)."""
    k = rng.integers(kmin, kmax + 1, size=nland)
    # Gumbel-top-k per landmark, done in groups of equal k to stay vectorised
    cams = np.empty(int(k.sum()), dtype=I64)
    ptr = np.zeros(nland + 1, dtype=I64)
    np.cumsum(k, out=ptr[1:])
    for kk in np.unique(k):
        idx = np.nonzero(k == kk)[0]
        for s in range(0, len(idx), 20000):
            sub = idx[s:s + 20000]
            keys = rng.random((len(sub), ncam))
            top = np.argpartition(keys, int(kk) - 1, axis=1)[:, :int(kk)]
            cams[(ptr[sub][:, None] + np.arange(kk)[None, :]).ravel()] = top.ravel()
    return k, ptr, cams


def _tracks_window(rng, ncam, nland, p=0.2, kmin=2, kmax=40):
    """k = clip(Geometric(p) on {1, 2, ...}, kmin, kmax); cams = contiguous window at U{0..ncam-k} (0-based)."""
    k = np.clip(rng.geometric(p, size=nland), kmin, min(kmax, ncam)).astype(I64)
    start = (rng.random(nland) * (ncam - k + 1)).astype(I64)
    ptr = np.zeros(nland + 1, dtype=I64)
    np.cumsum(k, out=ptr[1:])
    own = np.repeat(np.arange(nland, dtype=I64), k)
    cams = start[own] + (np.arange(ptr[-1], dtype=I64) - ptr[own])
    return k, ptr, cams


class HessianProblem:
    """A Schur-path instance: structure, pattern, workspace index and the filled storage (all layout order)."""


def _hessian_from_tracks(rng, ncam, cam_dims, nland, k, ptr, cams, lam, dense, cam_scale=None, nchunks=1,
                         de=3, extra_pattern=True):
    nobs = int(ptr[-1])
    own = np.repeat(np.arange(nland, dtype=I64), k)
    blocksizes = np.concatenate([cam_dims, np.full(nland, de, dtype=I64)])
    eliminated = np.concatenate([np.zeros(ncam, bool), np.ones(nland, bool)])
    # caller numbering: camera c -> c+1, landmark l -> ncam+l+1; one row block per observation
    rb_ptr = np.arange(0, 2 * nobs + 1, 2, dtype=I64)
    rb_flat = np.empty(2 * nobs, dtype=I64)
    rb_flat[0::2] = cams + 1
    rb_flat[1::2] = ncam + own + 1
    st = SchurStructure.from_description(blocksizes, (rb_ptr, rb_flat), eliminated, nchunks=nchunks)
    pos = np.empty(len(blocksizes), dtype=I64)
    pos[st.order - 1] = np.arange(1, len(blocksizes) + 1, dtype=I64)
    r_of = pos[cams] - 0                      # layout retained block (1-based) of each observation's camera
    e_of = pos[ncam + own] - st.nrblocks      # layout eliminated block (1-based)
    # coupling entry of each observation: position of (e, r) among the sorted entries
    owner_c = np.repeat(np.arange(1, st.neblocks + 1, dtype=I64), np.diff(st.eliminated_ptr))
    key_sorted = owner_c * (st.nrblocks + 1) + st.entry_retained
    c_of = np.searchsorted(key_sorted, e_of * (st.nrblocks + 1) + r_of) + 1
    dmax = int(cam_dims.max())
    Jc = rng.standard_normal((nobs, 2, dmax))
    Jl = rng.standard_normal((nobs, 2, de))
    rr = rng.standard_normal((nobs, 2))
    dcam = cam_dims[cams]
    Jc *= (np.arange(dmax)[None, None, :] < dcam[:, None, None])
    if cam_scale is not None:
        f = cam_scale[cams][:, None, None]
        Jc = Jc * f
        Jl = Jl * f
        rr = rr * f[:, :, 0]
    # E_data: per landmark sum of J_l'J_l, column-major de x de at eoffsets (a landmark's observations are
    # consecutive, so the per-landmark sums are segment sums in observation order)
    JlTJl = Jl[:, 0, :, None] * Jl[:, 0, None, :] + Jl[:, 1, :, None] * Jl[:, 1, None, :]   # (nobs, de, de), [i, j]
    E = np.zeros((st.neblocks, de, de))
    first = ptr[:-1]
    E[e_of[first] - 1] = np.add.reduceat(JlTJl, first, axis=0)
    E_data = E.transpose(0, 2, 1).reshape(-1).copy()    # column-major per block
    # C_data: per coupling J_l'J_c, de x dr column-major: C[co + (ir-1)*de + i-1] = (J_l'J_c)[i, ir]
    C_data = np.zeros(st.coupling_storage_length())
    dr_c = st.r_dims[st.entry_retained - 1]
    for d in np.unique(dr_c):
        m = dcam == d
        jl, jc = Jl[m], Jc[m][:, :, :d]
        blk = jc[:, 0, :, None] * jl[:, 0, None, :] + jc[:, 1, :, None] * jl[:, 1, None, :]   # (n, dr, de): [ir, i]
        co = st.coffsets[c_of[m] - 1] - 1
        C_data[(co[:, None] + np.arange(d * de)[None, :]).ravel()] = blk.reshape(len(co), -1).ravel()
    # b = J'r in layout order
    b_r = np.zeros(st.nr)
    b_e = np.zeros(st.ne)
    be = Jl[:, 0, :] * rr[:, 0, None] + Jl[:, 1, :] * rr[:, 1, None]
    b_e[(st.boffsets_e[e_of[first] - 1] - 1)[:, None] + np.arange(de)[None, :]] = np.add.reduceat(be, first, axis=0)
    bc = Jc[:, 0, :] * rr[:, 0, None] + Jc[:, 1, :] * rr[:, 1, None]
    ro = st.boffsets_r[r_of - 1] - 1
    for d in np.unique(dr_c):
        m = dcam == d
        np.add.at(b_r, ro[m][:, None] + np.arange(d)[None, :], bc[m][:, :d])
    # R: diagonal blocks sum J_c'J_c
    JcTJc = Jc[:, 0, :, None] * Jc[:, 0, None, :] + Jc[:, 1, :, None] * Jc[:, 1, None, :]
    Rblk = np.zeros((st.nrblocks, dmax, dmax))
    np.add.at(Rblk, r_of - 1, JcTJc)
    p = HessianProblem()
    p.structure, p.lam = st, lam
    p.E_data, p.C_data, p.b_r, p.b_e = E_data, C_data, b_r, b_e
    p.scaling_r, p.scaling_e = np.ones(st.nr), np.ones(st.ne)
    p.Rblocks = Rblk
    p.ws_index = SchurReductionWorkspaceIndex(st, nchunks)
    p.dense = dense
    if dense:
        R = np.zeros((st.nr, st.nr))
        for r in range(st.nrblocks):
            o, d = st.boffsets_r[r] - 1, st.r_dims[r]
            R[o:o + d, o:o + d] = Rblk[r, :d, :d]
        p.R = R
        p.pattern = None
    else:
        pat = build_schur_pattern(st)
        Rs = hessian_retained_storage(st, pat)
        Rdata = np.zeros(Rs.nnzvals)
        for r in range(st.nrblocks):
            s0 = Rs.block_start(r + 1, r + 1)
            d = st.r_dims[r]
            Rdata[s0 - 1:s0 - 1 + d * d] = Rblk[r, :d, :d].T.reshape(-1)   # column-major (symmetric anyway)
        p.pattern, p.Rsys, p.R_data = pat, Rs, Rdata
        p.r_to_s = build_r_to_s_map(Rs, pat, st.boffsets_r)
    p.cam_of_obs, p.e_of_obs = r_of, e_of
    return p


def config1(seed=1, scale=1.0, nchunks=1):
    rng = np.random.default_rng(seed)
    ncam, nland = 10, max(2, int(round(200 * scale)))
    k, ptr, cams = _tracks_uniform_random(rng, ncam, nland, 3, 6)
    return _hessian_from_tracks(rng, ncam, np.full(ncam, 6, dtype=I64), nland, k, ptr, cams, 1e-3, True, nchunks=nchunks)


def config2(seed=2, scale=1.0, nchunks=1, ncam=None, nland=None, dense=False):
    rng = np.random.default_rng(seed)
    ncam = ncam or max(4, int(round(1000 * min(1.0, scale * 10))))
    nland = nland or max(4, int(round(200000 * scale)))
    k, ptr, cams = _tracks_window(rng, ncam, nland)
    return _hessian_from_tracks(rng, ncam, np.full(ncam, 6, dtype=I64), nland, k, ptr, cams, 1e-4, dense, nchunks=nchunks)


def window_problem(seed=7, ncam=60, nland=1500, de=3, dr=6, kmin=2, kmax=40, unobserved=(), lam=1e-4, dense=False):
    """Sequential-SLAM tracks (contiguous camera windows, k = clip(Geometric(0.2), kmin, kmax)) with landmark blocks
    of size `de` and camera blocks of size `dr`. Cameras listed in `unobserved` (0-based) are seen by no landmark:
    windows that would cover one are moved to the longest free run of cameras."""
    rng = np.random.default_rng(seed)
    bad = np.zeros(ncam, bool)
    bad[list(unobserved)] = True
    free = np.nonzero(~bad)[0]
    # maximal runs of observed cameras
    runs = np.split(free, np.nonzero(np.diff(free) != 1)[0] + 1) if len(free) else []
    longest = max(len(r) for r in runs)
    k = np.clip(rng.geometric(0.2, size=nland), kmin, min(kmax, longest)).astype(I64)
    ptr = np.zeros(nland + 1, dtype=I64)
    np.cumsum(k, out=ptr[1:])
    cams = np.empty(int(ptr[-1]), dtype=I64)
    for l in range(nland):
        fits = [r for r in runs if len(r) >= k[l]]
        r = fits[int(rng.integers(len(fits)))]
        st = r[0] + int(rng.integers(len(r) - k[l] + 1))
        cams[ptr[l]:ptr[l + 1]] = np.arange(st, st + k[l])
    return _hessian_from_tracks(rng, ncam, np.full(ncam, dr, dtype=I64), nland, k, ptr, cams, lam, dense, de=de)


def config3(seed=3, scale=1.0, nchunks=1, ncam=256, nland=None, kmax=64):
    rng = np.random.default_rng(seed)
    nland = nland or max(4, int(round(1000000 * scale)))
    kmax = min(kmax, ncam)
    cam_scale = 10.0 ** rng.uniform(-2.0, 2.0, size=ncam)
    k, ptr, cams = _tracks_uniform_random(rng, ncam, nland, 2, kmax)
    p = _hessian_from_tracks(rng, ncam, np.full(ncam, 6, dtype=I64), nland, k, ptr, cams, 1e-4, True,
                             cam_scale=cam_scale, nchunks=nchunks)
    p.cam_scale = cam_scale
    return p


def config5_sizes(seed=5, nprob=64):
    rng = np.random.default_rng(seed)
    ncams = rng.integers(50, 2001, size=nprob)
    nlands = rng.integers(10000, 300001, size=nprob)
    return ncams, nlands


def config5_problem(i, seed=5, scale=1.0, ncam=None, nland=None):
    ncams, nlands = config5_sizes(seed)
    ncam = ncam or int(ncams[i])
    nland = nland or max(4, int(round(int(nlands[i]) * scale)))
    rng = np.random.default_rng([seed, i])
    k, ptr, cams = _tracks_window(rng, ncam, nland)
    return _hessian_from_tracks(rng, ncam, np.full(ncam, 6, dtype=I64), nland, k, ptr, cams, 1e-4, False)


def config5_union(seed=5, scale=1.0, nprob=64, ncams=None, nlands=None):
    """The batched-serving configuration as one structure: the nprob problems' cameras and landmarks side by side,
    so that S is block diagonal over problems and one reduction (one launch per kernel) and one batched PCG cover
    them all. Problem i has config5_sizes' camera count and scale x its landmark count; its tracks come from the
    stream [seed, i] as in config5_problem, the Jacobian values from one stream [seed, nprob]. Adds
    `prob_block_ptr`: the 1-based first camera block of each problem, then nrblocks + 1. `ncams` / `nlands`
    replace the generated sizes (small test instances)."""
    c5, l5 = config5_sizes(seed, nprob)
    ncams = c5 if ncams is None else np.asarray(ncams)
    nlands = l5 if nlands is None else np.asarray(nlands)
    nprob = len(ncams)
    ks, cams_all, coff = [], [], 0
    for i in range(nprob):
        rng = np.random.default_rng([seed, i])
        k, _, cams = _tracks_window(rng, int(ncams[i]), max(4, int(round(int(nlands[i]) * scale))))
        ks.append(k)
        cams_all.append(cams + coff)
        coff += int(ncams[i])
    k = np.concatenate(ks)
    ptr = np.zeros(len(k) + 1, dtype=I64)
    np.cumsum(k, out=ptr[1:])
    rng = np.random.default_rng([seed, nprob])
    p = _hessian_from_tracks(rng, coff, np.full(coff, 6, dtype=I64), len(k), k, ptr, np.concatenate(cams_all),
                             1e-4, False)
    p.prob_block_ptr = np.cumsum([1] + [int(c) for c in ncams]).astype(I64)
    p.nprob = nprob
    return p


class JacobianProblem:
    """A Jacobian-path instance: SchurJacobianStructure and the raw tables + Float64 residual."""


def config4(seed=4, scale=1.0, nchunks=1, ncam=None, nland=None, kmax=30):
    rng = np.random.default_rng(seed)
    ncam = ncam or max(4, int(round(500 * min(1.0, scale * 10))))
    nland = nland or max(4, int(round(500000 * scale)))
    cam_dims = np.where(rng.random(ncam) < 0.5, 6, 9).astype(I64)
    k, ptr, cams = _tracks_uniform_random(rng, ncam, nland, 2, min(kmax, ncam))
    return jacobian_from_tracks(rng, ncam, cam_dims, nland, k, ptr, cams, 1e-4, nchunks=nchunks)


def jacobian_from_tracks(rng, ncam, cam_dims, nland, k, ptr, cams, lam, nchunks=1, de=3):
    nobs = int(ptr[-1])
    own = np.repeat(np.arange(nland, dtype=I64), k)
    blocksizes = np.concatenate([cam_dims, np.full(nland, de, dtype=I64)])
    eliminated = np.concatenate([np.zeros(ncam, bool), np.ones(nland, bool)])
    rb_ptr = np.arange(0, 2 * nobs + 1, 2, dtype=I64)
    rb_flat = np.empty(2 * nobs, dtype=I64)
    rb_flat[0::2] = cams + 1
    rb_flat[1::2] = ncam + own + 1
    rowsz = np.full(nobs, 2, dtype=I64)
    jst = SchurJacobianStructure.from_description(blocksizes, (rb_ptr, rb_flat), rowsz, eliminated, nchunks=nchunks)
    st = jst.schur
    rb = jst.rowblocks
    data = np.zeros(jst.raw_length())
    res = np.zeros(jst.residual_length())
    dmax = int(cam_dims.max())
    Jc = rng.standard_normal((nobs, 2, dmax))
    Jl = rng.standard_normal((nobs, 2, de))
    rr = rng.standard_normal((nobs, 2))
    pos = np.empty(len(blocksizes), dtype=I64)
    pos[st.order - 1] = np.arange(1, len(blocksizes) + 1, dtype=I64)
    t = rb["table"]                       # table of each row block (observation)
    first = rb["first"]                   # first row in it
    r_of = pos[cams]
    M = jst.table_nrows[t - 1]
    off = jst.raw_offsets[t - 1] - 1
    # coupling entry of the observation's camera in its table
    owner_c = np.repeat(np.arange(1, jst.ntables + 1, dtype=I64), np.diff(jst.table_ptr))
    key_sorted = owner_c * (st.nrblocks + 1) + jst.entry_retained
    u = np.searchsorted(key_sorted, t * (st.nrblocks + 1) + r_of) + 1
    col0 = jst.entry_col[u - 1] - 1        # 0-based first column of the camera block
    dcam = st.r_dims[r_of - 1]
    ecol0 = jst.table_ncols[t - 1] - 1 - de  # 0-based first eliminated column
    for row in range(2):
        rowi = first - 1 + row
        for j in range(dmax):
            m = j < dcam
            data[off[m] + (col0[m] + j) * M[m] + rowi[m]] = Jc[m, row, j]
        for j in range(de):
            data[off + (ecol0 + j) * M + rowi] = Jl[:, row, j]
        res[jst.residual_offsets[t - 1] - 1 + rowi] = rr[:, row]
    p = JacobianProblem()
    p.jst, p.structure, p.lam = jst, st, lam
    p.data, p.r = data, res
    p.scaling_e = np.ones(st.ne)
    p.Jc, p.Jl, p.rr, p.cam_of_obs, p.e_of_obs = Jc, Jl, rr, r_of, pos[ncam + own] - st.nrblocks
    return p
