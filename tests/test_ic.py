"""Initial-condition recipes of pllc_ic_fill (host code, no GPU)."""
import numpy as np
from scipy.special import erf

import pllcart3d_amd as P


def centres(prm):
    nx, ny, nz = prm.nx, prm.ny, prm.nz
    dx, dy, dz = (prm.X2 - prm.X1) / nx, (prm.Y2 - prm.Y1) / ny, (prm.Z2 - prm.Z1) / nz
    xs = prm.X1 + (0 - 3) * dx + dx / 2.0
    ys = prm.Y1 - 1 * dy + dy / 2.0
    zs = prm.Z1 - 1 * dz + dz / 2.0
    return xs + np.arange(nx + 6) * dx, ys + np.arange(ny + 2) * dy, zs + np.arange(nz + 2) * dz


def splitmix64(seed):
    s = seed & (2 ** 64 - 1)
    M = 2 ** 64 - 1
    while True:
        s = (s + 0x9E3779B97F4A7C15) & M
        z = s
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        yield (z ^ (z >> 31)) >> 11


def test_erf_profile_matches_initial_condition():
    """c = 0.5 + 0.5 erf((Xc - Inf_pos)/Inf_thick) on every cell (pllcart3d.c:501-508)."""
    prm = P.make_params(20, 6, 5, Inf_thick=0.05)
    c = P.ic_fill(prm, P.IC_ERF)
    xc, _, _ = centres(prm)
    ref = 0.5 + 0.5 * erf((xc - 1.0) / 0.05)
    assert np.allclose(c, ref[:, None, None], rtol=0, atol=2e-16)


def test_noise_recipe():
    prm = P.make_params(16, 4, 4, Inf_thick=0.2)
    c = P.ic_fill(prm, P.IC_NOISE, seed=1, amp=0.01)
    base = P.ic_fill(prm, P.IC_ERF)
    g = splitmix64(1)
    ref = base.copy()
    for i in range(3, 19):
        for j in range(1, 5):
            for k in range(1, 5):
                u = 2.0 * (next(g) * (1.0 / 9007199254740992.0)) - 1.0
                v = ref[i, j, k]
                if 0.01 < v < 0.99:
                    ref[i, j, k] = min(1.0, max(0.0, v + 0.01 * u))
    ref[3:19, 0, 1:5] = ref[3:19, 1, 1:5]
    ref[3:19, 5, 1:5] = ref[3:19, 4, 1:5]
    ref[3:19, :, 0] = ref[3:19, :, 1]
    ref[3:19, :, 5] = ref[3:19, :, 4]
    assert np.array_equal(c, ref)
    assert c.min() >= 0.0 and c.max() <= 1.0


def test_multimode_front():
    prm = P.make_params(24, 10, 8, Inf_thick=0.02)
    c = P.ic_fill(prm, P.IC_MULTIMODE, seed=2, amp=0.01)
    xc, yc, zc = centres(prm)
    g = splitmix64(2)
    ph = np.empty((16, 16)); ps = np.empty((16, 16))
    for m in range(16):
        for n in range(16):
            ph[m, n] = 2 * np.pi * (next(g) * 2.0 ** -53)
            ps[m, n] = 2 * np.pi * (next(g) * 2.0 ** -53)
    mm = np.arange(1, 17)
    Y = (yc + 0.5)[:, None, None, None]
    Z = (zc + 0.5)[None, :, None, None]
    S = np.sum(np.cos(mm[:, None] * np.pi * Y + ph) * np.cos(mm[None, :] * np.pi * Z + ps), axis=(2, 3))
    eta = 0.01 * S / np.abs(S[1:-1, 1:-1]).max()
    ref = 0.5 + 0.5 * erf((xc[:, None, None] - 1.0 - eta[None]) / 0.02)
    assert np.allclose(c, ref, rtol=0, atol=1e-12)
    # the front is displaced by at most the amplitude
    assert abs(np.abs(eta[1:-1, 1:-1]).max() - 0.01) < 1e-15
