"""CPU oracle of the learned-preconditioner pieces (TEST INFRASTRUCTURE ONLY; no reference counterpart).

numpy restatements, in float64, of the MLP forward pass, the gradient of L = mean (o - y)^2 / 2, one Adam step, the
per-block features and the corrected blocks M_b = D_b + mu_b m_b I -- what the library's HIP kernels compute. The
gradient is pinned by central finite differences (tests/test_oracle_mlp.py), the features and blocks by their
definitions on a dense S.
"""
import numpy as np

NF, NH = 26, 64


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    o = 0
    W1 = p[o:o + NF * NH].reshape(NF, NH); o += NF * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def forward(p, fmean, fstd, X):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xn = (np.asarray(X, np.float64) - fmean) / fstd
    a1 = xn @ W1 + b1
    h1 = np.maximum(a1, 0)
    a2 = h1 @ W2 + b2
    h2 = np.maximum(a2, 0)
    return h2 @ W3 + b3, (xn, a1, h1, a2, h2)


def loss(p, fmean, fstd, X, y):
    o, _ = forward(p, fmean, fstd, X)
    return 0.5 * np.mean((o - y) ** 2)


def grad(p, fmean, fstd, X, y):
    W1, b1, W2, b2, W3, b3 = unpack(p)
    o, (xn, a1, h1, a2, h2) = forward(p, fmean, fstd, X)
    n = len(y)
    dO = (o - y) / n
    d2 = (dO[:, None] * W3[None, :]) * (a2 > 0)
    d1 = (d2 @ W2.T) * (a1 > 0)
    return np.concatenate([(xn.T @ d1).ravel(), d1.sum(0), (h1.T @ d2).ravel(), d2.sum(0), h2.T @ dO, [dO.sum()]])


def adam(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    return p - lr * (m / (1 - b1 ** step)) / (np.sqrt(v / (1 - b2 ** step)) + eps), m, v


def features(Sd, r_dims, boffsets_r, blocks_in_row, lam):
    """Features of every retained block of a dense symmetric S (blocks_in_row: stored blocks of each block row,
    the diagonal one included)."""
    nb = len(r_dims)
    F = np.zeros((nb, NF))
    for b in range(nb):
        d, o = r_dims[b], boffsets_r[b] - 1
        if d != 6:
            continue
        D = Sd[o:o + 6, o:o + 6]
        m = np.trace(D) / 6
        row = Sd[o:o + 6, :].copy()
        row[:, o:o + 6] = 0
        iu = [(r, c) for c in range(6) for r in range(c + 1)]
        F[b, :21] = [D[r, c] / m for (r, c) in iu]
        F[b, 21] = np.log10(m)
        F[b, 22] = np.log(1.0 + blocks_in_row[b] - 1)
        F[b, 23] = np.log10(lam)
        F[b, 24] = np.log10(m / lam)
        F[b, 25] = np.log10((np.linalg.norm(row) + 1e-300) / (np.linalg.norm(D) + 1e-300))
    return F


def shift_block(D, mu):
    return D + mu * (np.trace(D) / D.shape[0]) * np.eye(D.shape[0])


def group_block(Sd_group, dims, w):
    """The group correction's preconditioner block: the principal block submatrix of S over a group of consecutive
    retained blocks (dims), its off-diagonal blocks multiplied by w (the definition the device's
    bls_pcg_group_blocks_f64 follows; pinned by construction in the tests: w = 0 gives block-Jacobi's blocks, w = 1
    the group's exact submatrix of S)."""
    M = w * np.array(Sd_group, dtype=np.float64)
    o = 0
    for d in dims:
        M[o:o + d, o:o + d] = Sd_group[o:o + d, o:o + d]
        o += d
    return M
