"""Restatement of the conditioning scale (TEST INFRASTRUCTURE ONLY): the checker of the device's prepare_scale!
(bls_prepare_scale_f64).

Written from src/conditioning.jl:24-87: conditioning_scale (:29-35), EliminationConditioning's state (:47-55) and
prepare_scale! (:60-87), in numpy Float64 (IEEE-rounded sqrt and division, exact frexp / ldexp, no fused
operations), so the scale is the reference's bit for bit on the same diagonal.

Parity pinning (tests/test_conditioning.py): what the mathematics fixes for the rule -- s is a power of two with
q / s in [sqrt(1/2), sqrt(2)) for q = 1 / (1e-12 + sqrt|d|), s(0) = 2^40, a non-finite d gives 1, Float32's floor
2^-126 -- and the reference test's sequence of dampings (test/conditioning.jl:72-80): the same scale within the
band, a recomputation far outside it, and the first scale again on returning.
"""
import numpy as np

CONDITIONING_EPS = 1e-12
RESCALE_BOUND = 2.0 ** 16
SQRT_HALF = np.sqrt(0.5)
FLOATMIN = {np.float64: 2.0 ** -1022, np.float32: 2.0 ** -126}


def conditioning_scale(d, T=np.float64):
    """src/conditioning.jl:29-35, elementwise over an array of damped diagonal entries."""
    dd = np.abs(np.asarray(d, dtype=np.float64))
    fin = np.isfinite(dd)
    q = 1.0 / (CONDITIONING_EPS + np.sqrt(np.where(fin, dd, 0.0)))
    f, e = np.frexp(q)
    s = np.where(f >= SQRT_HALF, np.ldexp(1.0, e), np.ldexp(1.0, e - 1))
    s = np.maximum(s, FLOATMIN[T])
    return np.where(fin, s, 1.0)


class EliminationConditioning:
    """src/conditioning.jl:47-55: s = ones, d = zeros, stale, nrescale = 0."""

    def __init__(self, n):
        self.s = np.ones(n)
        self.d = np.zeros(n)
        self.stale = True
        self.nrescale = 0


def prepare_scale(c, diag, lam, sc, T=np.float64):
    """src/conditioning.jl:60-87: diag(d) writes the undamped diagonal (called only when stale)."""
    lam = max(lam, 0.0)
    sc = np.ones(len(c.s)) if sc is None else np.asarray(sc, dtype=np.float64)
    if c.stale:
        diag(c.d)
        c.s[:] = conditioning_scale(c.d + lam * sc, T)
        c.stale = False
    else:
        dj = c.d + lam * sc
        v = np.abs(dj) * (c.s * c.s)
        outside = ~((1.0 / RESCALE_BOUND <= v) & (v <= RESCALE_BOUND))
        if np.any(outside & (conditioning_scale(dj, T) != c.s)):
            c.s[:] = conditioning_scale(dj, T)
            c.nrescale += 1
    return c.s
