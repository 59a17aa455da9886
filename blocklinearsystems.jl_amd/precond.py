"""Learned correction of the block-Jacobi preconditioner for the batched PCG (no reference counterpart).

The retained blocks of each problem are grouped in order, GROUP at a time ((b0, .., b0 + 3), (b0 + 4, ..), ...); per
retained block b an MLP (26 FP32 features -> 64 -> 64 -> 1, ReLU) reads the features of `bls_pcg_features_f32` and
emits a coupling weight w_b = alpha sigmoid(o_b) in [0, 1] (alpha, an output scale chosen on held-out problems, is 0
when the network does not beat plain block-Jacobi there). The preconditioner block of the group that b starts
becomes S's principal block submatrix over the group with its off-diagonal blocks scaled by w_b: w = 0 is plain
block-Jacobi (BlockJacobi of ext/KrylovExt.jl:88-95), w = 1 the group's exact block of S, and every w in between
keeps it SPD (a convex combination of the two).
Features, forward, backward, Adam and the corrected blocks are the library's HIP kernels; this module only holds the
parameters and drives them. Plain block-Jacobi stays the fallback: a learned solve that does not converge is redone
with the plain preconditioner, and the FP64 true residual ||S x - g|| / ||g|| judges both.

(An earlier family, a diagonal shift M_b = D_b + mu_b m_b I, stays available as `shift_blocks` / `MLP.mu`: even its
per-problem best shift saved only 2-5 % of the iterations on the training generator.)
"""
import ctypes as ct
import os

import numpy as np

from . import device as dev

MU_MAX = 0.3
OMEGA_MAX = 1.0
GROUP = 4   # retained blocks per preconditioner group of the correction
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


def group_blocks(pcg, S, omega, group=GROUP, stream=None):
    """The group-corrected preconditioner blocks for coupling weights omega (device Float32, one per retained block,
    read at each group's first block); the solve_groups preconditioner."""
    return pcg.group_blocks(S, omega, group, stream)


def shift_blocks(pcg, S, mu, stream=None):
    """Corrected preconditioner blocks M_b = D_b + mu_b m_b I (device FP64, the solve's precond_blocks)."""
    import torch
    M = torch.empty(pcg.factor_size(), dtype=torch.float64, device="cuda")
    dev._check(dev.lib().bls_pcg_shift_blocks_f64(pcg.h, dev._dp(S), dev._dp(mu.contiguous()), dev._dp(M),
                                                  dev._stream(stream)))
    return M


class MLP:
    """Parameters (flat FP32: W1 [26][64], b1, W2 [64][64], b2, W3 [64], b3) and feature statistics, on the device."""

    def __init__(self, params, fmean, fstd, alpha=1.0, group=GROUP):
        import torch
        t = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float32)).cuda()
        self.params, self.fmean, self.fstd = t(params), t(fmean), t(fstd)
        self.alpha = float(alpha)   # output scale chosen on held-out problems (0: the correction is off)
        self.group = int(group)     # retained blocks per preconditioner group the network was trained for

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
            return cls(z["params"], z["fmean"], z["fstd"], float(z["alpha"]) if "alpha" in z else 1.0,
                       int(z["group"]) if "group" in z else GROUP)

    def save(self, path=WEIGHTS, **meta):
        np.savez(path, params=self.params.cpu().numpy(), fmean=self.fmean.cpu().numpy(),
                 fstd=self.fstd.cpu().numpy(), alpha=np.array(self.alpha), group=np.array(self.group), **meta)

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

    def omega(self, F):
        """Coupling weights alpha OMEGA_MAX sigmoid(o) for features F (zero-feature rows: blocks not of dimension 6
        -> 0, their group stays uncoupled)."""
        import torch
        o = self.forward(F)
        w = (self.alpha * OMEGA_MAX) * torch.sigmoid(o)
        return torch.where((F != 0).any(dim=1), w, torch.zeros_like(w)).contiguous()

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


def target_logit(w_best, scale=OMEGA_MAX):
    """Regression target for a chosen coupling weight (or shift, with scale MU_MAX): logit(w / scale), w clipped
    into (0, scale)."""
    q = np.clip(np.asarray(w_best, dtype=np.float64) / scale, 1e-3, 1 - 1e-3)
    return np.log(q / (1 - q))


def solve(pcg, S, g, x, lam, model=None, tol=1e-8, maxit=10000):
    """PCG with the learned correction (model given) or plain block-Jacobi; a learned solve that fails falls back to
    plain block-Jacobi. Returns (ok, iters, relres, used_learned). A model with output scale 0 couples no group
    (w = 0), so its solve is the plain one and neither features nor the network are evaluated."""
    if model is not None and model.alpha != 0.0:
        w = model.omega(features(pcg, S, lam))
        M = group_blocks(pcg, S, w, model.group)
        x.zero_()
        ok, iters, rel = pcg.solve_groups(S, g, x, M, tol=tol, maxit=maxit)
        if ok:
            return ok, iters, rel, True
    x.zero_()
    ok, iters, rel = pcg.solve(S, g, x, tol=tol, maxit=maxit)
    return ok, iters, rel, False
