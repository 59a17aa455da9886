"""The dense Schur system's solve on the device: a host-side mirror of DenseSchurLinearSystem
(src/schur_linear_systems.jl:29-90) and its solve (conditioned_schur_solve!, schur_solve!, :292-335) over the C ABI's
entry points, with the reference's names, argument meaning and status values.

A solve runs, on one stream, with no host round trip until the status is read:
  prepare_scale! (src/conditioning.jl:60-87) from getdiag! of H as filled (conditioned systems only);
  build_reduced!: reset_schur_dense! + reduce_schur_dense! in the scaled variables (:301-307);
  solve_reduced_dense! (:312-319) on the reduced system: Cholesky, falling back to Bunch-Kaufman (:ldl) when it
  fails ("Indefinite" when S has a negative eigenvalue; :ldl alone when asked for);
  backsubstitute! (src/schur_reduction.jl:519-614), then unscale! and the "x finite" rule (:333-334).
The reduced system's own conditioning inside DampedLinearSolve is a power-of-two scaling, which commutes with every
operation of the Cholesky factorization and solve, so it is not repeated here (Bunch-Kaufman's pivot choices
compare magnitudes of the scaled matrix; the factorization it finds is the same up to rounding for a power-of-two
scale).
"""
from __future__ import annotations

import numpy as np

from . import device as dev

SUCCESS, NOT_POSITIVE_DEFINITE, INDEFINITE = "Success", "NotPositiveDefinite", "Indefinite"


class DenseSchurLinearSystem:
    """A Float64 SchurBlockHessian with a dense R (device tensors E_data, C_data, R (nr x nr column-major, flat)),
    the right-hand side b = [b_r; b_e] and the damping's scaling vectors, all resident on the device. `condition`
    is the reference's internal switch (src/schur_linear_systems.jl:58)."""

    def __init__(self, st, E_data, C_data, R, b_r, b_e, scaling_r, scaling_e, condition=True, nchunks=1):
        import torch
        from .structure import SchurReductionWorkspaceIndex
        self.st = st
        self.E_data, self.C_data, self.R = E_data, C_data, R
        self.b_r, self.b_e = b_r, b_e
        self.scaling_r, self.scaling_e = scaling_r, scaling_e
        self.nr, self.ne = int(st.nr), int(st.ne)
        n = self.nr + self.ne
        self.handle = dev.SchurHandle(st, SchurReductionWorkspaceIndex(st, nchunks), dense=True)
        self.ws = dev.SchurReductionWorkspace(self.handle)
        self.layout = dev.AssemblyHandle(st, [], [], None)          # the storage layout, for getdiag!
        self.chol = dev.DenseCholesky(self.nr)
        self.ldl = dev.DenseLDL(self.nr)
        self.cond = dev.EliminationConditioning(n, on=condition)
        f = lambda k: torch.empty(k, dtype=torch.float64, device="cuda")
        self.S, self.g_S = f(self.nr * self.nr), f(self.nr)
        self.x = f(n)
        self.scaling = torch.cat([scaling_r, scaling_e])
        self.lam = 0.0
        self._nonfinite = torch.zeros(1, dtype=torch.int32, device="cuda")

    def dampen(self, lam):
        """dampen!(ls, lambda): the damping lambda Diagonal([scaling_r; scaling_e]) of the next solve."""
        self.lam = float(lam)

    def zero(self):
        """zero! after a new fill of H: the next solve computes the conditioning scale afresh."""
        self.cond.markstale()

    @property
    def x_r(self):
        return self.x[:self.nr]

    @property
    def x_e(self):
        return self.x[self.nr:]

    def solve(self, solver="auto", stream=None):
        """solve!(ls, solver) for solver in ("auto", "cholesky", "ldl"): "Success" or "Indefinite" with
        x = [x_r; x_e] solved, or "NotPositiveDefinite"."""
        if solver not in ("auto", "cholesky", "ldl"):
            raise ValueError(f"{solver!r} is not a reduced solver")
        nr = self.nr
        if self.cond.on:
            s = dev.prepare_scale_(self.cond, lambda d: dev.hessian_diag_(d, self.layout, self.E_data, self.R,
                                                                          stream=stream),
                                   self.lam, self.scaling, stream=stream)
            s_r, s_e = s[:nr], s[nr:]
        else:
            s = s_r = s_e = None
        dev.reset_schur_dense_(self.S, self.R, self.handle, self.scaling_r, self.lam, s_r=s_r, stream=stream)
        ok = dev.reduce_schur_dense_(self.S, self.g_S, self.b_r, self.b_e, self.scaling_e, self.lam, self.handle,
                                     self.E_data, self.C_data, self.ws, s_r=s_r, s_e=s_e, stream=stream)
        if not ok:
            return NOT_POSITIVE_DEFINITE
        status = SUCCESS
        ok = solver != "ldl" and self.chol.factor(self.S, stream=stream)[0]
        if ok:
            self.chol.solve(self.g_S, self.x_r, stream=stream)
        else:   # Bunch-Kaufman on a copy of S (:ldl; the fallback of :auto and :cholesky)
            ok, neg = self.ldl.factor(self.S, stream=stream)
            if not ok:
                return NOT_POSITIVE_DEFINITE
            status = INDEFINITE if neg > 0 else SUCCESS
            self.ldl.solve(self.g_S, self.x_r, stream=stream)
        dev.backsubstitute_(self.x_e, self.x_r, self.handle, self.C_data, self.ws, s_r=s_r, s_e=s_e, stream=stream)
        dev.unscale_(self.cond, self.x, s, self._nonfinite, stream=stream)
        return status if int(self._nonfinite.item()) == 0 else NOT_POSITIVE_DEFINITE
