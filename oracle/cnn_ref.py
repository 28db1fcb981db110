"""FP64 CPU reference of the learned closure's network: a 3-layer 3x3x3
convolutional network on the coarse grid's cells (numpy).

TEST INFRASTRUCTURE ONLY: the tests (tests/) compare the GPU's BF16 forward
and backward kernels with it; it is never the product.  The reference has no learned
closure; this network corrects the explicit advective term of Hc_op
(pllcart3d.c:1539-1644), N = -(u c_x + v c_y + w c_z), on a 2x coarser grid.

Network (channels-last arrays [nx][ny][nz][C]):
    a1 = tanh(conv(x,  W1) + b1)     x: 16 channels (5 used: c, u, v, w, mu
                                        at cell centres, normalised; the
                                        other 11 are zero)
    a2 = tanh(conv(a1, W2) + b2)     16 -> 16
    y  = conv(a2, W3) + b3           16 -> 1
conv(a, W)[p][co] = sum_{o in {-1,0,1}^3} sum_ci a[p + o][ci] W[o][ci][co],
with zero padding outside the grid; W has shape [27][16][Cout], offset
o = (oi, oj, ok) at index 9 (oi + 1) + 3 (oj + 1) + (ok + 1).
Loss: sum over the masked cells of (y - t)^2 / (number of masked cells).
"""
import numpy as np

OFFSETS = [(oi, oj, ok) for oi in (-1, 0, 1) for oj in (-1, 0, 1) for ok in (-1, 0, 1)]


def conv(a, W):
    """a: [nx][ny][nz][Ci], W: [27][Ci][Co] -> [nx][ny][nz][Co] (zero padding)."""
    nx, ny, nz, _ = a.shape
    ap = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nx, ny, nz, W.shape[2]))
    for q, (oi, oj, ok) in enumerate(OFFSETS):
        out += ap[1 + oi:1 + oi + nx, 1 + oj:1 + oj + ny, 1 + ok:1 + ok + nz] @ W[q]
    return out


def conv_adjoint_input(g, W):
    """d/da of sum(g * conv(a, W)): [nx][ny][nz][Co] -> [nx][ny][nz][Ci]."""
    nx, ny, nz, _ = g.shape
    gp = np.pad(g, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.zeros((nx, ny, nz, W.shape[1]))
    for q, (oi, oj, ok) in enumerate(OFFSETS):
        # a[p + o] meets g[p]: a[r] meets g[r - o]
        out += gp[1 - oi:1 - oi + nx, 1 - oj:1 - oj + ny, 1 - ok:1 - ok + nz] @ W[q].T
    return out


def conv_adjoint_weight(a, g):
    """d/dW of sum(g * conv(a, W)): [27][Ci][Co]."""
    nx, ny, nz, ci = a.shape
    ap = np.pad(a, ((1, 1), (1, 1), (1, 1), (0, 0)))
    gw = np.zeros((27, ci, g.shape[3]))
    gf = g.reshape(-1, g.shape[3])
    for q, (oi, oj, ok) in enumerate(OFFSETS):
        gw[q] = ap[1 + oi:1 + oi + nx, 1 + oj:1 + oj + ny, 1 + ok:1 + ok + nz].reshape(-1, ci).T @ gf
    return gw


def forward(x, P):
    """P: dict W1 [27][16][16], b1 [16], W2, b2, W3 [27][16][1], b3 [1]."""
    a1 = np.tanh(conv(x, P["W1"]) + P["b1"])
    a2 = np.tanh(conv(a1, P["W2"]) + P["b2"])
    y = conv(a2, P["W3"]) + P["b3"]
    return y[..., 0], (a1, a2)


def loss_and_grads(x, t, mask, P):
    """Loss and its gradient with respect to every parameter."""
    y, (a1, a2) = forward(x, P)
    m = mask.astype(np.float64)
    nm = m.sum()
    r = (y - t) * m
    L = float((r * r).sum() / nm)
    dy = (2.0 / nm) * r[..., None]
    G = {"W3": conv_adjoint_weight(a2, dy), "b3": dy.sum(axis=(0, 1, 2))}
    dz2 = conv_adjoint_input(dy, P["W3"]) * (1.0 - a2 * a2)
    G["W2"] = conv_adjoint_weight(a1, dz2)
    G["b2"] = dz2.sum(axis=(0, 1, 2))
    dz1 = conv_adjoint_input(dz2, P["W2"]) * (1.0 - a1 * a1)
    G["W1"] = conv_adjoint_weight(x, dz1)
    G["b1"] = dz1.sum(axis=(0, 1, 2))
    return L, G, y


def init_params(seed, cin_used=5):
    """Seeded parameters: W scaled by 1/sqrt(fan-in), inputs beyond
    cin_used zero, biases zero, the last layer small."""
    rng = np.random.default_rng(seed)
    P = {"W1": rng.normal(size=(27, 16, 16)) / np.sqrt(27 * cin_used),
         "W2": rng.normal(size=(27, 16, 16)) / np.sqrt(27 * 16),
         "W3": 0.1 * rng.normal(size=(27, 16, 1)) / np.sqrt(27 * 16),
         "b1": np.zeros(16), "b2": np.zeros(16), "b3": np.zeros(1)}
    P["W1"][:, cin_used:, :] = 0.0
    return P
