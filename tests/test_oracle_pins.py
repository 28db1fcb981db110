"""Pins of the FP64 oracle (CPU).  Each check is independent of the
oracle's own code: the reference's compiled pll_tridiag (fixture), FFTW
known-answer vectors, library routines (numpy, scipy), closed forms on
polynomial data, and discrete invariants of the scheme.
"""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.fft

from tests.conftest import rel_l2

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def O(oracle_mod):
    return oracle_mod


def _params(nx, ny, nz, **kw):
    import pllcart3d_amd as P
    return P.make_params(nx, ny, nz, **kw)


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---------------------------------------------------------------- Thomas

def test_pll_tridiag_bitwise_vs_reference_build(O):
    """orc_pll_tridiag == the reference's own pll_tridiag.c, bit for bit
    (tests/golden/pll_tridiag_ref.npz, made by make_tridiag_fixture.py)."""
    L = O.load()
    g = np.load(os.path.join(GOLD, "pll_tridiag_ref.npz"))
    for t in range(int(g["ncases"])):
        a, b, c, r = (np.ascontiguousarray(g["in_%s_%d" % (k, t)]) for k in "abcr")
        ns, n = a.shape
        b2, c2, r2 = b.copy(), c.copy(), r.copy()
        L.orc_pll_tridiag(ns, n, dp(a), dp(b2), dp(c2), dp(r2))
        assert np.array_equal(r2, g["out_r_%d" % t]), t
        assert np.array_equal(b2, g["out_b_%d" % t]), t
        assert np.array_equal(c2, g["out_c_%d" % t]), t


def test_tridiag_1rhs_dense_and_pll_identity(O):
    """tridiag_1rhs solves the system (dense solve) and is arithmetically the
    same as pll_tridiag at one rank (bitwise equal solutions)."""
    L = O.load()
    rng = np.random.default_rng(5)
    for n in (1, 2, 3, 17, 128):
        a = -rng.uniform(0, 1, n); c = -rng.uniform(0, 1, n); b = 2.1 + rng.uniform(0, 1, n)
        r = rng.normal(size=n)
        x = r.copy()
        L.orc_tridiag_1rhs(n, dp(a), dp(b), dp(c), dp(x))
        A = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
        assert np.allclose(A @ x, r, rtol=0, atol=1e-13)
        a2, b2, c2, r2 = a.copy(), b.copy(), c.copy(), r.copy()
        L.orc_pll_tridiag(1, n, dp(a2), dp(b2), dp(c2), dp(r2))
        assert np.array_equal(x, r2)


# ------------------------------------------------------------------- DCT

def test_redft_fftw_known_answers(O):
    """REDFT10/REDFT01 against FFTW's own output (fftw_dct_kat.npz)."""
    L = O.load()
    kat = np.load(os.path.join(GOLD, "fftw_dct_kat.npz"))
    for key in kat.files:
        t, n = key.split("_")[1:]
        n = int(n)
        x = np.linspace(0, n - 1, n)
        y = np.empty(n)
        (L.orc_redft10 if t == "2" else L.orc_redft01)(n, dp(x), dp(y))
        assert rel_l2(y, kat[key]) < 1e-13, key


def test_dct2_matches_scipy_and_inverts(O):
    L = O.load()
    rng = np.random.default_rng(7)
    for nx, ny, nz in [(3, 8, 8), (2, 12, 5), (4, 16, 32)]:
        x = rng.normal(size=(nx, ny, nz))
        y = np.empty_like(x)
        L.orc_fdct2(nx, ny, nz, dp(x), dp(y))
        ref = scipy.fft.dctn(x, type=2, axes=(1, 2))
        assert rel_l2(y, ref) < 1e-14
        z = np.empty_like(x)
        L.orc_idct2(nx, ny, nz, dp(y), dp(z))
        assert rel_l2(z, x) < 1e-14   # REDFT01(REDFT10) = 4 ny nz * identity


# ------------------------------------------------------------------ WENO

def _weno(O, d1, kk, sign):
    d1 = np.ascontiguousarray(d1, dtype=np.float64)
    return O.load().orc_weno(dp(d1), kk, sign)


def test_weno_exact_for_quadratics(O):
    """Every 3-point candidate is exact for a quadratic, so the weighted
    derivative is f'(x_m) for either upwind direction."""
    h = 0.05
    x = np.arange(20) * h
    f = 0.3 - 1.7 * x + 2.9 * x ** 2
    d1 = np.diff(f) / h
    for m in range(5, 14):
        exact = -1.7 + 2 * 2.9 * x[m]
        assert abs(_weno(O, d1, m - 3, 1) - exact) < 1e-11
        assert abs(_weno(O, d1, m + 2, -1) - exact) < 1e-11


def test_weno_fifth_order_on_smooth_data(O):
    """Linear weights (0.1, 0.6, 0.3) give 5th order; a wrong weight drops
    to 3rd.  Error ratio between h and h/2 on sin(x)."""
    errs = []
    for n in (40, 80, 160):
        h = 2 * np.pi / n
        x = np.arange(n + 10) * h
        d1 = np.diff(np.sin(x)) / h
        m = 5 + n // 4
        errs.append(max(abs(_weno(O, d1, m - 3, 1) - np.cos(x[m])),
                        abs(_weno(O, d1, m + 2, -1) - np.cos(x[m]))))
    assert errs[0] / errs[1] > 20 and errs[1] / errs[2] > 20


def test_weno_is_essentially_non_oscillatory(O):
    """Across a jump the weight of every stencil that crosses it vanishes:
    the result equals the smooth-side candidate to ~eps."""
    h = 0.1
    x = np.arange(12) * h
    f = np.where(x > 0.75, 10.0, 0.0) + x ** 2   # jump between x[7] and x[8]
    d1 = np.diff(f) / h
    m = 5   # sign +1 uses d1[2..6]: the jump is in d1[7], outside -> smooth
    assert abs(_weno(O, d1, m - 3, 1) - 2 * x[m]) < 1e-9
    m = 7   # d1[4..8] contains the jump at d1[7] in v4: the left candidate wins
    left = (2 * d1[4] - 7 * d1[5] + 11 * d1[6]) / 6.0
    assert abs(_weno(O, d1, m - 3, 1) - left) < 1e-6 * abs(left)


# ------------------------------------------------------- composed operators

def _quadratic_state(O, nx, ny, nz, U, V, W, g):
    prm = _params(nx, ny, nz, Vratio=1.0, Pe=100.0)
    o = O.Oracle(prm)
    o.init()
    X1, Y1, Z1 = prm.X1, prm.Y1, prm.Z1
    dx, dy, dz = (prm.X2 - X1) / nx, (prm.Y2 - Y1) / ny, (prm.Z2 - Z1) / nz
    xc = X1 + (np.arange(nx + 6) - 3 + 0.5) * dx
    yc = Y1 + (np.arange(ny + 2) - 1 + 0.5) * dy
    zc = Z1 + (np.arange(nz + 2) - 1 + 0.5) * dz
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    b = (0.2, -0.4, 0.6)
    cval = 0.5 + b[0] * X + b[1] * Y + b[2] * Z + g[0] * X ** 2 + g[1] * Y ** 2 + g[2] * Z ** 2
    for name, val in (("C_OLD", cval), ("U_NEW", U), ("V_NEW", V), ("W_NEW", W),
                      ("C_N", 0.0), ("C_NM1", 0.0)):
        o.field(name)[...] = val
    grads = (b[0] + 2 * g[0] * X, b[1] + 2 * g[1] * Y, b[2] + 2 * g[2] * Z)
    return o, (dx, dy, dz), grads


@pytest.mark.parametrize("shape", [(12, 8, 8), (12, 8, 6)])
@pytest.mark.parametrize("signs", [(1, 1, 1), (-1, -1, -1), (1, -1, 1), (-1, 1, -1)])
def test_hc_op_quadratic_closed_form(O, shape, signs):
    """Hc_op on a quadratic c and uniform velocity: exact derivative except
    at the ENO end cells, where a second difference read one index outside
    the range the pass wrote is used (pllcart3d.c:1568/1596/1626).  What it
    reads there: x pass 0 (fresh scratch); y pass at js-1: 0, at je+1: the x
    pass's c_xx/2 if that index is an x unknown; z pass at ks-1: 0, at
    ke+1: the y pass's c_yy/2 if an y unknown, else the x pass's, else 0."""
    nx, ny, nz = shape
    U, V, W = 0.7 * signs[0], 0.9 * signs[1], 0.8 * signs[2]
    g = (0.3, -0.7, 1.1)
    o, (dx, dy, dz), (gx, gy, gz) = _quadratic_state(O, nx, ny, nz, U, V, W, g)
    o.op("CONCENTRATION", 1e-4, 0, U)
    N = o.field("C_N")
    is_, ie, js, je, ks, ke = 3, nx + 2, 1, ny, 1, nz
    hi_y = g[0] if is_ <= ny + 1 <= ie else 0.0
    hi_z = g[1] if js <= nz + 1 <= je else (g[0] if is_ <= nz + 1 <= ie else 0.0)

    def deriv(exact, m, lo, hi, s, h, gam, hi_stale):
        d = exact.copy()
        if s > 0:
            d[lo] = exact[lo] - gam * h              # D2[lo-1] = 0 wins the minabs
        else:
            if abs(gam) > abs(hi_stale):             # D2[hi+1] = stale wins
                d[hi] = exact[hi] + gam * h - hi_stale * h
        return d

    cx = deriv(gx[:, 1, 1], None, is_, ie, signs[0], dx, g[0], 0.0)
    cy = deriv(gy[3, :, 1], None, js, je, signs[1], dy, g[1], hi_y)
    cz = deriv(gz[3, 1, :], None, ks, ke, signs[2], dz, g[2], hi_z)
    I, J, K = np.meshgrid(np.arange(is_, ie + 1), np.arange(js, je + 1), np.arange(ks, ke + 1), indexing="ij")
    expect = -(W * cz[K] + U * cx[I] + V * cy[J])
    got = N[is_:ie + 1, js:je + 1, ks:ke + 1]
    assert np.abs(got - expect).max() < 1e-11 * np.abs(expect).max()


def test_viscosity_on_linear_c(O):
    """calculate_viscosity: exp(M c) at centres; at edges the 4-point mean of
    a linear c equals c at the edge, so mu_edge = exp(M c(edge))."""
    nx, ny, nz = 6, 5, 4
    prm = _params(nx, ny, nz, Vratio=2.5)
    o = O.Oracle(prm)
    o.init()
    dx, dy, dz = 4.0 / nx, 1.0 / ny, 1.0 / nz
    xc = (np.arange(nx + 6) - 3 + 0.5) * dx
    yc = -0.5 + (np.arange(ny + 2) - 1 + 0.5) * dy
    zc = -0.5 + (np.arange(nz + 2) - 1 + 0.5) * dz
    lin = lambda x, y, z: 0.1 + 0.05 * x + 0.3 * y - 0.2 * z
    X, Y, Z = np.meshgrid(xc, yc, zc, indexing="ij")
    o.field("C_NEW")[...] = lin(X, Y, Z)
    o.op("VISCOSITY")
    M = 2.5
    cell, xy, xz, yz = (o.field(n) for n in ("MUN_CELL", "MUN_XY", "MUN_XZ", "MUN_YZ"))
    s = np.s_[2:nx + 4, 0:ny + 2, 0:nz + 2]
    assert np.allclose(cell[s], np.exp(M * lin(X, Y, Z))[s], rtol=1e-14, atol=0)
    s = np.s_[3:nx + 4, 1:ny + 2, 1:nz + 1]
    assert np.allclose(xy[s], np.exp(M * lin(X - dx / 2, Y - dy / 2, Z))[s], rtol=1e-13, atol=0)
    s = np.s_[3:nx + 4, 1:ny + 1, 1:nz + 2]
    assert np.allclose(xz[s], np.exp(M * lin(X - dx / 2, Y, Z - dz / 2))[s], rtol=1e-13, atol=0)
    s = np.s_[3:nx + 3, 1:ny + 2, 1:nz + 2]
    assert np.allclose(yz[s], np.exp(M * lin(X, Y - dy / 2, Z - dz / 2))[s], rtol=1e-13, atol=0)


def _laplace_ghosts(phi, nx, ny, nz):
    """phi with the ghost values the Poisson operator implies: Neumann in y,
    z and at the inlet; at the outlet Neumann for every (y,z) mode but the
    mean, whose ghost value is 0 (pllcart3d.c:2059-2073)."""
    p = phi.copy()
    s = p[3:nx + 3]
    p[3:nx + 3, 0, :] = p[3:nx + 3, 1, :]
    p[3:nx + 3, ny + 1, :] = p[3:nx + 3, ny, :]
    p[3:nx + 3, :, 0] = p[3:nx + 3, :, 1]
    p[3:nx + 3, :, nz + 1] = p[3:nx + 3, :, nz]
    p[2, 1:ny + 1, 1:nz + 1] = p[3, 1:ny + 1, 1:nz + 1]
    last = p[nx + 2, 1:ny + 1, 1:nz + 1]
    p[nx + 3, 1:ny + 1, 1:nz + 1] = last - last.mean()
    del s
    return p


@pytest.mark.parametrize("shape", [(8, 6, 4), (10, 8, 8)])
def test_poisson_residual(O, shape):
    """phisolver_dct solves the 7-point Laplacian with its boundary rules:
    L phi = Re/(dt*alpha) * div on every interior cell."""
    nx, ny, nz = shape
    prm = _params(nx, ny, nz)
    o = O.Oracle(prm)
    o.init()
    rng = np.random.default_rng(11)
    div = np.zeros((nx + 6, ny + 2, nz + 2))
    div[3:nx + 3, 1:ny + 1, 1:nz + 1] = rng.normal(size=(nx, ny, nz))
    o.field("DIV")[...] = div
    dt, rk = 1e-3, 1
    o.op("PHISOLVER", dt, rk)
    phi = _laplace_ghosts(o.field("PHI"), nx, ny, nz)
    dx, dy, dz = 4.0 / nx, 1.0 / ny, 1.0 / nz
    c = np.s_[3:nx + 3, 1:ny + 1, 1:nz + 1]
    L = ((phi[2:nx + 2, 1:ny + 1, 1:nz + 1] - 2 * phi[c] + phi[4:nx + 4, 1:ny + 1, 1:nz + 1]) / dx ** 2 +
         (phi[3:nx + 3, 0:ny, 1:nz + 1] - 2 * phi[c] + phi[3:nx + 3, 2:ny + 2, 1:nz + 1]) / dy ** 2 +
         (phi[3:nx + 3, 1:ny + 1, 0:nz] - 2 * phi[c] + phi[3:nx + 3, 1:ny + 1, 2:nz + 2]) / dz ** 2)
    alpha = 5.0 / 12.0 - 17.0 / 60.0
    f = 1.0 / (dt * alpha) * div[c]
    assert rel_l2(L, f) < 1e-11


def test_divergence_and_cfl_closed_forms(O):
    nx, ny, nz = 6, 5, 4
    prm = _params(nx, ny, nz, CFLmax=0.3)
    o = O.Oracle(prm)
    o.init()
    dx, dy, dz = 4.0 / nx, 1.0 / ny, 1.0 / nz
    i = np.arange(nx + 6)[:, None, None]; j = np.arange(ny + 2)[None, :, None]; k = np.arange(nz + 2)[None, None, :]
    o.field("U_NEW")[...] = 0.3 * i * dx + 0 * j + 0 * k
    o.field("V_NEW")[...] = -0.5 * j * dy + 0 * i + 0 * k
    o.field("W_NEW")[...] = 1.25 * k * dz + 0 * i + 0 * j
    o.op("DIVERGENCE")
    d = o.field("DIV")[3:nx + 3, 1:ny + 1, 1:nz + 1]
    assert np.allclose(d, 0.3 - 0.5 + 1.25, rtol=1e-13)
    for name, v in (("U_OLD", 0.4), ("V_OLD", -0.2), ("W_OLD", 0.1)):
        o.field(name)[...] = v
    cfl = o.op("FIND_CFL")
    assert abs(cfl - 0.3 / (0.4 / dx + 0.2 / dy + 0.1 / dz)) < 1e-15


def test_bulk_velo_first_of_ties(O):
    nx, ny, nz = 6, 3, 3
    o = O.Oracle(_params(nx, ny, nz))
    o.init()
    u = o.field("U_OLD")
    u[...] = 0.0
    u[nx + 2, 2, 1] = -0.75       # first (j,k) with |u| = 0.75 in scan order
    u[nx + 2, 3, 3] = 0.75
    u[nx + 2, 1, 2] = 0.5
    assert o.op("BULK_VELO") == -0.75


def test_outflow_deficit_trapezoid(O):
    """velo_correction with phi = 0: the outlet face gains the inflow minus
    outflow per unit area, the fluxes being 2D trapezoids over interior
    nodes (pllcart3d.c:2171-2223)."""
    nx, ny, nz = 8, 10, 6
    o = O.Oracle(_params(nx, ny, nz))
    o.init()
    yc = -0.5 + (np.arange(ny + 2) - 1 + 0.5) / ny
    zc = -0.5 + (np.arange(nz + 2) - 1 + 0.5) / nz
    prof = 1.5 * (1 - 4 * yc[1:ny + 1] ** 2)
    qin = np.trapezoid(np.trapezoid(np.tile(prof[:, None], (1, nz)), zc[1:nz + 1], axis=1), yc[1:ny + 1])
    u = o.field("U_NEW")
    before = u[nx + 3, 1:ny + 1, 1:nz + 1].copy()
    assert np.all(before == 0.0)
    o.op("VELO_CORRECTION", 1e-3, 0)
    assert np.allclose(u[nx + 3, 1:ny + 1, 1:nz + 1], qin / 1.0, rtol=1e-14)


def test_concentration_adi_factored_residual(O):
    """With zero velocity, concentration's increment d solves the factored
    system (I - f A_y)(I - f A_z)(I - f A_x) d = dt*alpha*(A c_old) with
    Neumann y/z rows, d = 0 at the inlet ghost and at the outflow face."""
    nx, ny, nz = 10, 6, 5
    prm = _params(nx, ny, nz, Pe=50.0)
    o = O.Oracle(prm)
    o.init()
    rng = np.random.default_rng(3)
    c = o.field("C_OLD")
    c[...] = rng.uniform(0, 1, c.shape)
    for n in ("U_NEW", "V_NEW", "W_NEW", "C_N", "C_NM1"):
        o.field(n)[...] = 0.0
    dt, rk = 2e-3, 2
    g, r = 3.0 / 4.0, -5.0 / 12.0
    cold = c.copy()
    o.op("CONCENTRATION", dt, rk, 0.0)
    dx, dy, dz, Pe = 4.0 / nx, 1.0 / ny, 1.0 / nz, 50.0
    ax, ay, az = 1 / (Pe * dx * dx), 1 / (Pe * dy * dy), 1 / (Pe * dz * dz)
    ci = np.s_[3:nx + 3, 1:ny + 1, 1:nz + 1]
    lap = (ax * (cold[2:nx + 2, 1:ny + 1, 1:nz + 1] - 2 * cold[ci] + cold[4:nx + 4, 1:ny + 1, 1:nz + 1]) +
           ay * (cold[3:nx + 3, 0:ny, 1:nz + 1] - 2 * cold[ci] + cold[3:nx + 3, 2:ny + 2, 1:nz + 1]) +
           az * (cold[3:nx + 3, 1:ny + 1, 0:nz] - 2 * cold[ci] + cold[3:nx + 3, 1:ny + 1, 2:nz + 2]))
    S = dt * (g + r) * lap
    d = o.field("C_NEW")[ci] - cold[ci]
    f = 0.5 * dt * (g + r)

    def app(d, axis, coef, neumann):
        p = np.pad(d, [(1, 1) if a == axis else (0, 0) for a in range(3)])
        sl = lambda a, b: tuple(slice(a, b) if x == axis else slice(None) for x in range(3))
        if neumann:
            p[sl(0, 1)] = p[sl(1, 2)]
            p[sl(-1, None)] = p[sl(-2, -1)]
        n = d.shape[axis]
        return d - f * coef * (p[sl(0, n)] - 2 * p[sl(1, n + 1)] + p[sl(2, n + 2)])

    res = app(app(app(d, 0, ax, False), 2, az, True), 1, ay, True)
    assert rel_l2(res, S) < 1e-12


def test_projection_leaves_interior_divergence_free(O):
    """After whole steps, div(u) vanishes on every cell but the last x-plane
    (whose outlet face is not corrected, pllcart3d.c:2162)."""
    import pllcart3d_amd as P
    prm = _params(16, 8, 8, Vratio=2.0, Pe=200.0, Inf_thick=0.1)
    o = O.Oracle(prm)
    o.init(P.ic_fill(prm, P.IC_NOISE, 5, 0.05))
    o.run(3)
    u, v, w = o.field("U_OLD"), o.field("V_OLD"), o.field("W_OLD")
    nx, ny, nz = 16, 8, 8
    d = ((u[4:nx + 4, 1:ny + 1, 1:nz + 1] - u[3:nx + 3, 1:ny + 1, 1:nz + 1]) * nx / 4.0 +
         (v[3:nx + 3, 2:ny + 2, 1:nz + 1] - v[3:nx + 3, 1:ny + 1, 1:nz + 1]) * ny +
         (w[3:nx + 3, 1:ny + 1, 2:nz + 2] - w[3:nx + 3, 1:ny + 1, 1:nz + 1]) * nz)
    scale = np.abs(u[3:nx + 3]).max() * nx / 4.0
    assert np.abs(d[:-1]).max() < 1e-11 * scale
    assert np.abs(d[-1]).max() > 1e-6 * scale


def test_uniform_state_is_steady(O):
    """c uniform and no flow: nothing changes over whole steps."""
    prm = _params(8, 6, 6)
    o = O.Oracle(prm)
    c0 = np.full((14, 8, 8), 0.37)
    o.init(c0)
    o.field("U_NEW")[...] = 0.0
    o.field("U_OLD")[...] = 0.0
    o.run(2)
    assert np.all(o.field("C_OLD") == 0.37)
    assert np.all(o.field("V_OLD") == 0.0)


# ------------------------------------------------ non-monotonic viscosity

def _nmvp_consts(M, Mum):
    alpha = np.exp(M)
    mum = Mum * alpha
    return alpha, mum, np.arcsin(alpha / mum), np.pi - np.arcsin(1 / mum)


def test_nmvp_law_end_values_and_peak(O):
    """calculate_viscosity_nmvp: mu(0) = 1, mu(1) = e^M, and on a ramp of c
    the profile rises to the single peak Mum e^M and falls to e^M; the edge
    means of a field constant in y and z reduce to the face values."""
    nx, ny, nz, M, Mum = 240, 3, 2, 2.5, 1.5
    prm = _params(nx, ny, nz, Vratio=M, Nmvp=1, Mum=Mum, Lambda=0.0)
    o = O.Oracle(prm)
    o.init()
    assert 0.1 < o.cm < 0.9
    names = ("MUN_CELL", "MUN_XY", "MUN_XZ", "MUN_YZ")
    for c0, mu0 in ((0.0, 1.0), (1.0, np.exp(M))):
        o.field("C_NEW")[...] = c0
        o.op("VISCOSITY")
        for n in names:
            f = o.field(n)[3:nx + 3, 1:ny + 1, 1:nz + 1]
            assert np.allclose(f, mu0, rtol=1e-14, atol=0), (c0, n)
    ramp = (np.arange(nx + 6) - 2.0) / (nx + 1)          # c = 0 at i = 2, 1 at i = nx+3
    o.field("C_NEW")[...] = ramp[:, None, None]
    o.op("VISCOSITY")
    cell = o.field("MUN_CELL")[2:nx + 4, 1, 1]
    assert abs(cell[0] - 1.0) < 1e-14 and abs(cell[-1] - np.exp(M)) < 1e-13 * np.exp(M)
    top = int(np.argmax(cell))
    assert 0 < top < len(cell) - 1
    assert np.all(np.diff(cell[:top + 1]) > 0) and np.all(np.diff(cell[top:]) < 0)
    assert Mum * np.exp(M) * (1 - 1e-4) < cell[top] <= Mum * np.exp(M) * (1 + 1e-15)
    s = np.s_[3:nx + 3, 1:ny + 1, 1:nz + 1]
    assert np.allclose(o.field("MUN_YZ")[s], o.field("MUN_CELL")[s], rtol=1e-15, atol=0)
    assert np.allclose(o.field("MUN_XY")[s], o.field("MUN_XZ")[s], rtol=1e-15, atol=0)


@pytest.mark.parametrize("M,Mum,Lam", [(2.5, 1.5, 0.0), (3.0, 1.2, 0.2), (1.5, 2.0, -0.1)])
def test_slope_parameter_matches_analytic_root(O, M, Mum, Lam):
    """slope_parameter's bisection on a fine linear profile c = 0..1 finds
    the root of lambda(cm) = (mu'(0) + mu'(1)) / (e^M + 1) = Lambda, with
    mu' taken analytically for the law psi = psi0 + (psi1-psi0) eta and
    eta = (1+a)c/(1+ac); the law peaks at mu = Mum e^M at c = 1 - cm."""
    from scipy.optimize import brentq
    alpha, mum, p0, p1 = _nmvp_consts(M, Mum)
    etam = (np.pi / 2 - p0) / (p1 - p0)

    def a_of(cm):
        return ((1 - cm) - etam) / ((1 - cm) * (etam - 1))

    def dmu(c, cm):
        a = a_of(cm)
        eta = (1 + a) * c / (1 + a * c)
        return mum * np.cos(p0 + (p1 - p0) * eta) * (p1 - p0) * (1 + a) / (1 + a * c) ** 2

    lam = lambda cm: (dmu(0.0, cm) + dmu(1.0, cm)) / (alpha + 1) - Lam
    root = brentq(lam, 0.1 + 1e-9, 0.9 - 1e-9, xtol=1e-14)
    n = 20001
    c1d = np.linspace(0.0, 1.0, n)
    L = O.load()
    cm = L.orc_slope_parameter(dp(c1d), n, M, Mum, Lam)
    assert abs(cm - root) < 2e-3, (cm, root)
    # peak of that orientation at c = 1 - cm (bisection midpoint accuracy)
    cc = np.linspace(0, 1, 200001)
    a = a_of(cm)
    mu = mum * np.sin(p0 + (p1 - p0) * (1 + a) * cc / (1 + a * cc))
    assert abs(cc[np.argmax(mu)] - (1 - cm)) < 1e-4
    assert abs(mu.max() - mum) < 1e-9 * mum


def test_slope_parameter_without_root_fails(O):
    """A Lambda that lambda(cm) never reaches in [0.1, 0.9]: the bisection
    gives up after 500 iterations, where the reference writes "Lambda did not
    converge" and exits (pllcart3d.c:812-817); the restatement returns -1
    and the oracle's init raises."""
    import pllcart3d_amd as P
    n = 2001
    c1d = np.linspace(0.0, 1.0, n)
    L = O.load()
    assert L.orc_slope_parameter(dp(c1d), n, 2.5, 1.5, 1e3) == -1.0
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1, Nmvp=1, Mum=1.5, Lambda=1e3)
    orc = O.Oracle(prm)
    with pytest.raises(RuntimeError, match="did not converge"):
        orc.init(P.ic_fill(prm, P.IC_NOISE, 3, 0.05))


# ------------------------------------------------------ wave perturbation

def test_spline_is_the_natural_cubic_spline(O):
    """spline: scipy's natural cubic spline inside [x1, x_{N-1}], y[0]
    below x1 (also on [x0, x1)) and y[N-1] beyond the last point; linear data
    is reproduced exactly."""
    from scipy.interpolate import CubicSpline
    L = O.load()
    N, dx = 40, 0.1
    x = (np.arange(N) - 2.5) * dx
    rng = np.random.default_rng(9)
    y = np.sin(2.0 * x) + 0.1 * rng.normal(size=N)
    xx = np.concatenate([x - 0.037, x + 0.061, [x[0] + 0.5 * dx, x[-1] + 1e-3, x[0] - 1.0]])
    # the routine evaluates N probes per call: feed them in chunks of N
    out = []
    for c0 in range(0, len(xx), N):
        probe = np.full(N, x[5])
        part = xx[c0:c0 + N]
        probe[:len(part)] = part
        res = np.zeros(N)
        L.orc_spline(N, dx, dp(x), dp(probe), dp(y), dp(res))
        out.append(res[:len(part)])
    got = np.concatenate(out)
    ref = CubicSpline(x, y, bc_type="natural")(xx)
    inside = (xx >= x[1]) & (xx <= x[-1])
    assert np.allclose(got[inside], ref[inside], rtol=0, atol=1e-12)
    assert np.all(got[(xx < x[1])] == y[0])
    assert np.all(got[(xx > x[-1])] == y[-1])
    lin = 0.3 + 2.0 * x
    res = np.zeros(N)
    probe = x - 0.037
    L.orc_spline(N, dx, dp(x), dp(probe), dp(lin), dp(res))
    ok = probe >= x[1]
    assert np.allclose(res[ok], 0.3 + 2.0 * probe[ok], rtol=0, atol=1e-13)


@pytest.mark.parametrize("subharm", [0, 1])
def test_shift_conc_moves_each_line_by_meshdx(O, subharm):
    """shift_conc: every (j, k) line of c_old, ghosts included, becomes its
    natural spline sampled at x - meshdx(z_k), meshdx = A cos(2 pi z / L)
    (+ the half-wavenumber subharmonic)."""
    from scipy.interpolate import CubicSpline
    nx, ny, nz = 20, 4, 6
    prm = _params(nx, ny, nz)
    o = O.Oracle(prm)
    o.init()
    rng = np.random.default_rng(4)
    c0 = rng.uniform(0, 1, size=o.shape)
    o.field("C_OLD")[...] = c0
    A, lam = 0.13, 0.8
    o.shift_conc(A, lam, subharm)
    got = o.field("C_OLD")
    dx, dz = 4.0 / nx, 1.0 / nz
    xc = (np.arange(nx + 6) - 3 + 0.5) * dx
    zc = -0.5 + (np.arange(nz + 2) - 1 + 0.5) * dz
    for k in range(nz + 2):
        md = A * np.cos(2 * np.pi / lam * zc[k])
        if subharm:
            md = A * np.cos(2 * np.pi / lam * (zc[k] + lam / 2)) + A * np.cos(np.pi / lam * (zc[k] + lam / 2))
        xx = xc - md
        for j in range(ny + 2):
            ref = CubicSpline(xc, c0[:, j, k], bc_type="natural")(xx)
            ref[xx > xc[-1]] = c0[-1, j, k]
            ref[xx < xc[1]] = c0[0, j, k]
            assert np.allclose(got[:, j, k], ref, rtol=0, atol=1e-12), (j, k)
