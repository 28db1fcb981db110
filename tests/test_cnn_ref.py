"""Pins of the FP64 reference network (oracle/cnn_ref.py), the reference for
the GPU closure kernels: the convolution against a brute-force sum, its
two adjoints by the inner-product identity, and the parameter gradients by
central finite differences."""
import numpy as np

from oracle import cnn_ref as R


def _brute_conv(a, W):
    nx, ny, nz, ci = a.shape
    out = np.zeros((nx, ny, nz, W.shape[2]))
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for q, (oi, oj, ok) in enumerate(R.OFFSETS):
                    ii, jj, kk = i + oi, j + oj, k + ok
                    if 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz:
                        out[i, j, k] += a[ii, jj, kk] @ W[q]
    return out


def test_conv_equals_brute_force_sum():
    rng = np.random.default_rng(1)
    a = rng.normal(size=(3, 4, 5, 3))
    W = rng.normal(size=(27, 3, 2))
    assert np.allclose(R.conv(a, W), _brute_conv(a, W), rtol=0, atol=1e-12)
    # the offset table: index 9 (oi+1) + 3 (oj+1) + (ok+1)
    assert R.OFFSETS[9 * 2 + 3 * 0 + 1] == (1, -1, 0)


def test_adjoints_satisfy_inner_product_identity():
    rng = np.random.default_rng(2)
    a = rng.normal(size=(4, 3, 5, 6))
    W = rng.normal(size=(27, 6, 4))
    g = rng.normal(size=(4, 3, 5, 4))
    lhs = np.sum(g * R.conv(a, W))
    assert np.isclose(lhs, np.sum(a * R.conv_adjoint_input(g, W)), rtol=1e-12)
    assert np.isclose(lhs, np.sum(W * R.conv_adjoint_weight(a, g)), rtol=1e-12)


def test_gradients_match_central_differences():
    rng = np.random.default_rng(3)
    shape = (5, 4, 6)
    x = np.zeros(shape + (16,))
    x[..., :5] = rng.normal(size=shape + (5,))
    t = rng.normal(size=shape)
    mask = rng.uniform(size=shape) > 0.3
    P = R.init_params(7)
    P["W3"] *= 10.0
    P["b1"] = 0.1 * rng.normal(size=16)
    L0, G, _ = R.loss_and_grads(x, t, mask, P)
    h = 1e-6
    for name in ("W1", "W2", "W3", "b1", "b2", "b3"):
        flat = P[name].reshape(-1)
        idx = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        if name == "W1":   # only the five used input channels carry gradient
            idx = [q * 256 + ci * 16 + co for q, ci, co in zip(rng.integers(0, 27, 6), rng.integers(0, 5, 6),
                                                               rng.integers(0, 16, 6))]
        for m in idx:
            old = flat[m]
            flat[m] = old + h
            Lp = R.loss_and_grads(x, t, mask, P)[0]
            flat[m] = old - h
            Lm = R.loss_and_grads(x, t, mask, P)[0]
            flat[m] = old
            fd = (Lp - Lm) / (2 * h)
            an = G[name].reshape(-1)[m]
            assert abs(fd - an) <= 1e-6 * max(abs(fd), 1e-3), (name, m, fd, an)
