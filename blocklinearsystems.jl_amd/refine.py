"""Float64 iterative refinement of the mixed-precision reduction (no reference counterpart).

The mixed reduction forms the reduced system S_m, g_m of the Float32-stored blocks (Float64 factorisation and
accumulation). Solving it leaves the error of the inputs' rounding: its solution's residual against the reduced
system of the Float64 blocks, ||g - S x|| / ||g||, is ~1e-6 on the bundle-adjustment workloads. Refinement removes it
without forming S from the Float64 blocks: the residual of that system is evaluated from the blocks themselves,

    x_e = z - E_inv sum_c C_c x            (backsubstitute! with the Float64 E_inv, z and C)
    r   = b_r - sum_c C_c' x_e - R_d x     (retained rows of the full system's residual; R_d = R + damping)

which equals g - S x exactly (g = b_r - C' z, S = R_d - C' E_inv C), and the correction solves S_m d = r by PCG and
adds d. Each evaluation streams C twice in Float64; the bulk reduction streamed it once in Float32. The reference
documents the same trade on the CPU (docs/src/tutorials/dampedlinearsolve.md:41-49: Float64 refinement of the
Float32 path removed its error but cost more than it saved).
"""
from __future__ import annotations

import numpy as np

from . import device as dev


class RefinementData:
    """The Float64 side of a refined mixed solve: the Float64 blocks (device tensors), the reset S holding R plus
    damping in the reduced pattern, the Float64 workspace (E_inv, z) and scratch vectors."""

    def __init__(self, handle, E64, C64, b_r, b_e, scaling_e, SR, lam):
        import torch
        st = handle.st
        self.h, self.E, self.C, self.b_r, self.b_e, self.sc, self.SR, self.lam = handle, E64, C64, b_r, b_e, scaling_e, SR, lam
        self.ws = dev.SchurReductionWorkspace(handle, "f64")
        self.x_e = torch.empty(st.ne, dtype=torch.float64, device="cuda")
        self.t = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        self.r = torch.empty(st.nr, dtype=torch.float64, device="cuda")
        self.d = torch.empty(st.nr, dtype=torch.float64, device="cuda")


def factor(rd, check=True):
    """factor_all_eliminated! on the Float64 blocks (E_inv and z for the residual evaluations). check=False when
    lambda and scaling_e were validated already (by the mixed reduction of the same step)."""
    return dev.factor_eliminated_(rd.h, rd.ws, rd.E, rd.b_e, rd.sc, rd.lam, check=check)


def residual(pcg, rd, x, gref):
    """r = g - S x of the Float64 blocks' reduced system, into rd.r; returns ||r|| / ||gref|| per problem."""
    dev.backsubstitute_(rd.x_e, x, rd.h, rd.C, rd.ws)
    dev.retained_residual_(rd.t, rd.h, rd.b_r, rd.x_e, rd.C)
    return pcg.residual(rd.SR, rd.t, x, gref, rd.r)


def refined_solve(pcg, S_m, g_m, x, rd, tol=1e-8, inner_tol=1e-4, maxsteps=6, maxit=10000, group_blocks=None,
                  factored=False):
    """Solve the Float64 blocks' reduced system to ||g - S x|| <= tol ||g|| (judged against g_m, which equals g to
    the rounding of the inputs) using only PCG on the mixed system S_m: x = 0, then repeat {d = S_m^-1 r to
    inner_tol (with the learned group correction when group_blocks is given); x += d; r = g - S x (Float64, from the
    blocks)} until the residual is met. Returns
    (converged, steps, pcg_iterations_total per problem, relres per problem)."""
    if not factored and not factor(rd):
        raise dev.BLSError(f"Float64 factorisation failed at eliminated block {rd.ws.failed_block}")
    x.zero_()
    rhs = g_m
    iters = np.zeros(pcg.nprob, dtype=np.int64)
    rel = np.ones(pcg.nprob)
    for step in range(1, maxsteps + 1):
        rd.d.zero_()
        if group_blocks is None:
            pcg.solve(S_m, rhs, rd.d, tol=inner_tol, maxit=maxit)
        else:
            pcg.solve_groups(S_m, rhs, rd.d, group_blocks, tol=inner_tol, maxit=maxit)
        it = pcg.last_iters
        iters += np.maximum(it, 0)
        dev.axpy_(x, 1.0, rd.d)
        rel = residual(pcg, rd, x, g_m)
        if np.all(rel <= tol):
            return True, step, iters, rel
        rhs = rd.r
    return False, maxsteps, iters, rel


def refined_dense_solve(chol, g_m, x, rd, steps=1):
    """The dense reduced system of the Float64 blocks solved through the Cholesky factor of the mixed system
    (chol, factored from S_m): x = S_m^-1 g_m, then `steps` refinement steps x += S_m^-1 (g - S x), the residual
    from the Float64 blocks (backsubstitute!, the retained residual, and R + damping as the dense reset S in
    rd.SR). rd.ws must hold the Float64 factorisation (factor(rd)). Returns rd.r, the last residual computed
    (g - S x before the last update)."""
    chol.solve(g_m, x)
    for _ in range(steps):
        dev.backsubstitute_(rd.x_e, x, rd.h, rd.C, rd.ws)
        dev.retained_residual_(rd.t, rd.h, rd.b_r, rd.x_e, rd.C)
        dev.dense_residual_(rd.r, rd.SR, rd.t, x)
        chol.solve(rd.r, rd.d)
        dev.axpy_(x, 1.0, rd.d)
    return rd.r
