"""Learned correction of the block-Jacobi preconditioner for the batched PCG (no reference counterpart).

Per retained block b of a reduced system S, an MLP (26 FP32 features -> 64 -> 64 -> 1, ReLU) reads the features of
`bls_pcg_features_f32` and emits a shift mu_b = alpha MU_MAX sigmoid(o_b) (alpha, an output scale chosen on
held-out problems, is 0 when the network does not beat plain block-Jacobi there); the preconditioner block becomes
M_b = D_b + mu_b m_b I, where D_b is S's diagonal block and m_b the mean of its diagonal (mu_b = 0 is plain
block-Jacobi, BlockJacobi of ext/KrylovExt.jl:88-95). Features, forward, backward, Adam and the corrected blocks are
the library's HIP kernels; this module only holds the parameters and drives them. Plain block-Jacobi stays the
fallback: a learned solve that does not converge is redone with the plain preconditioner, and the FP64 true
residual ||S x - g|| / ||g|| judges both.
"""
import ctypes as ct
import os

import numpy as np

from . import device as dev

MU_MAX = 0.3
WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "precond_mlp.npz")


def dims():
    v = [ct.c_int(0) for _ in range(4)]
    dev._check(dev.lib().bls_mlp_dims(*[ct.byref(x) for x in v]))
    return tuple(int(x.value) for x in v)   # nfeatures, nhidden, nparams, work per sample


def features(pcg, S, lam, stream=None):
    """(nrblocks, 26) FP32 device tensor of the MLP features of reduced system S (device, FP64)."""
    import torch
    nf = dims()[0]
    F = torch.empty((pcg.st.nrblocks, nf), dtype=torch.float32, device="cuda")
    dev._check(dev.lib().bls_pcg_features_f32(pcg.h, dev._dp(S), float(lam), dev._dp(F), dev._stream(stream)))
    return F


def shift_blocks(pcg, S, mu, stream=None):
    """Corrected preconditioner blocks M_b = D_b + mu_b m_b I (device FP64, the solve's precond_blocks)."""
    import torch
    M = torch.empty(pcg.factor_size(), dtype=torch.float64, device="cuda")
    dev._check(dev.lib().bls_pcg_shift_blocks_f64(pcg.h, dev._dp(S), dev._dp(mu.contiguous()), dev._dp(M),
                                                  dev._stream(stream)))
    return M


class MLP:
    """Parameters (flat FP32: W1 [26][64], b1, W2 [64][64], b2, W3 [64], b3) and feature statistics, on the device."""

    def __init__(self, params, fmean, fstd, alpha=1.0):
        import torch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
        self.params, self.fmean, self.fstd = t(params), t(fmean), t(fstd)
        self.alpha = float(alpha)   # output scale chosen on held-out problems (0: the correction is off)

    @classmethod
    def initial(cls, fmean, fstd, seed=0):
        """He-initialised weights, zero biases."""
        nf, nh, npar, _ = dims()
        rng = np.random.default_rng(seed)
        w1 = rng.standard_normal((nf, nh)) * np.sqrt(2.0 / nf)
        w2 = rng.standard_normal((nh, nh)) * np.sqrt(2.0 / nh)
        w3 = rng.standard_normal(nh) * np.sqrt(1.0 / nh)
        p = np.concatenate([w1.ravel(), np.zeros(nh), w2.ravel(), np.zeros(nh), w3, [0.0]])
        assert len(p) == npar
        return cls(p, fmean, fstd)

    @classmethod
    def load(cls, path=WEIGHTS):
        with np.load(path) as z:
            return cls(z["params"], z["fmean"], z["fstd"], float(z["alpha"]) if "alpha" in z else 1.0)

    def save(self, path=WEIGHTS, **meta):
        np.savez(path, params=self.params.cpu().numpy(), fmean=self.fmean.cpu().numpy(),
                 fstd=self.fstd.cpu().numpy(), alpha=np.array(self.alpha), **meta)

    def forward(self, X, stream=None):
        import torch
        out = torch.empty(X.shape[0], dtype=torch.float32, device="cuda")
        dev._check(dev.lib().bls_mlp_forward_f32(dev._dp(self.params), dev._dp(self.fmean), dev._dp(self.fstd),
                                                 dev._dp(X.contiguous()), X.shape[0], dev._dp(out), dev._stream(stream)))
        return out

    def grad(self, X, y, stream=None):
        """(outputs, gradient of mean (o - y)^2 / 2)."""
        import torch
        n = X.shape[0]
        _, _, npar, wps = dims()
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        G = torch.empty(npar, dtype=torch.float32, device="cuda")
        work = torch.empty(n * wps, dtype=torch.float32, device="cuda")
        dev._check(dev.lib().bls_mlp_grad_f32(dev._dp(self.params), dev._dp(self.fmean), dev._dp(self.fstd),
                                              dev._dp(X.contiguous()), dev._dp(y.contiguous()), n, dev._dp(out),
                                              dev._dp(G), dev._dp(work), dev._stream(stream)))
        return out, G

    def mu(self, F):
        """Shifts alpha MU_MAX sigmoid(o) for features F (zero-feature rows: blocks not of dimension 6 -> 0)."""
        import torch
        o = self.forward(F)
        mu = (self.alpha * MU_MAX) * torch.sigmoid(o)
        return torch.where((F != 0).any(dim=1), mu, torch.zeros_like(mu)).contiguous()


def train(model, X, y, steps=400, batch=16384, lr=1e-3, seed=0, log=None):
    """Adam on mean (o - y)^2 / 2 over random mini-batches of (X, y) (device tensors); returns the losses."""
    import torch
    _, _, npar, _ = dims()
    m = torch.zeros(npar, dtype=torch.float32, device="cuda")
    v = torch.zeros(npar, dtype=torch.float32, device="cuda")
    g = torch.Generator(device="cpu").manual_seed(seed)
    n = X.shape[0]
    losses = []
    for step in range(1, steps + 1):
        idx = torch.randint(0, n, (min(batch, n),), generator=g).cuda()
        xb, yb = X[idx].contiguous(), y[idx].contiguous()
        out, G = model.grad(xb, yb)
        dev._check(dev.lib().bls_mlp_adam_f32(dev._dp(model.params), dev._dp(G), dev._dp(m), dev._dp(v), step,
                                              float(lr), None))
        if step % 50 == 0 or step == 1:
            losses.append(float(0.5 * ((out - yb) ** 2).mean()))
            if log:
                log(f"step {step}: loss {losses[-1]:.4f}")
    return losses


def target_logit(mu_best):
    """Regression target for a chosen shift: logit(mu / MU_MAX), with mu clipped into (0, MU_MAX)."""
    q = np.clip(np.asarray(mu_best, dtype=np.float64) / MU_MAX, 1e-3, 1 - 1e-3)
    return np.log(q / (1 - q))


def solve(pcg, S, g, x, lam, model=None, tol=1e-8, maxit=10000):
    """PCG with the learned correction (model given) or plain block-Jacobi; a learned solve that fails falls back to
    plain block-Jacobi. Returns (ok, iters, relres, used_learned). A model with output scale 0 shifts no block
    (mu = 0), so its solve is the plain one and neither features nor the network are evaluated."""
    if model is not None and model.alpha != 0.0:
        mu = model.mu(features(pcg, S, lam))
        M = shift_blocks(pcg, S, mu)
        x.zero_()
        ok, iters, rel = pcg.solve(S, g, x, tol=tol, maxit=maxit, precond_blocks=M)
        if ok:
            return ok, iters, rel, True
    x.zero_()
    ok, iters, rel = pcg.solve(S, g, x, tol=tol, maxit=maxit)
    return ok, iters, rel, False
