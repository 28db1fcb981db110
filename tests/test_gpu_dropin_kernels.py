"""The drop-in kernels on host data, on the GPU through the C-ABI:
pll_tridiag (pll_tridiag.c:152-158), tridiag_1rhs (pllcart3d.c:3511-3536)
and calculate_weno (pllcart3d.c:3455-3507).

pll_tridiag is held bit for bit to the reference's own pll_tridiag.c
(tests/golden/pll_tridiag_ref.npz, written by make_tridiag_fixture.py from
the reference build) and to the oracle at the x-line shape of the
256x128x128 workload; tridiag_1rhs and calculate_weno bit for bit to the
oracle, whose functions are pinned to the reference build.
"""
import ctypes as C
import os

import numpy as np
import pytest

import pllcart3d_amd as P

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
dpt = C.POINTER(C.c_double)


class Tds(C.Structure):
    # structures_3d.h:2-10
    _fields_ = [("acof", dpt), ("bcof", dpt), ("ccof", dpt), ("rhs", dpt)]


def _p(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(dpt)


def _lib():
    L = P.load()
    L.pll_tridiag.argtypes = [C.POINTER(Tds), C.c_int, C.c_int]
    L.pll_tridiag.restype = None
    L.tridiag_1rhs.argtypes = [C.c_int, dpt, dpt, dpt, dpt]
    L.tridiag_1rhs.restype = C.c_int
    L.calculate_weno.argtypes = [dpt, C.c_int, C.c_int]
    L.calculate_weno.restype = C.c_double
    return L


def gpu_pll_tridiag(a, b, c, r):
    """pll_tridiag on rows of 2D arrays (one system per row), in place."""
    L = _lib()
    ns, n = a.shape
    arr = (Tds * ns)()
    for s in range(ns):
        arr[s] = Tds(_p(a[s]), _p(b[s]), _p(c[s]), _p(r[s]))
    before = P.last_error()
    L.pll_tridiag(arr, ns, n)
    assert P.last_error() == before, P.last_error()


def test_pll_tridiag_bitwise_vs_reference_build():
    """Every case of the fixture written by the reference's pll_tridiag.c:
    b, c and r after the call equal the reference's in every bit."""
    g = np.load(os.path.join(GOLD, "pll_tridiag_ref.npz"))
    for t in range(int(g["ncases"])):
        a, b, c, r = (np.ascontiguousarray(g["in_%s_%d" % (k, t)]).copy() for k in "abcr")
        print("[pll_tridiag] fixture case %d: %d systems of %d" % (t, a.shape[0], a.shape[1]), flush=True)
        gpu_pll_tridiag(a, b, c, r)
        assert np.array_equal(r, g["out_r_%d" % t]), t
        assert np.array_equal(b, g["out_b_%d" % t]), t
        assert np.array_equal(c, g["out_c_%d" % t]), t


@pytest.mark.parametrize("ns,n", [(16384, 256), (1000, 3), (7, 1), (333, 129)])
def test_pll_tridiag_bitwise_vs_oracle(oracle_mod, ns, n):
    """pll_tridiag at the x-line shape of the 256x128x128 workload (16384
    systems of 256) and ragged counts: equal to the oracle in every bit,
    twice in a row (the staging buffers are kept between calls)."""
    L = oracle_mod.load()
    rng = np.random.default_rng(ns + n)
    for rep in range(2):
        a = -rng.uniform(0, 1, (ns, n))
        c = -rng.uniform(0, 1, (ns, n))
        b = 2.05 + rng.uniform(0, 1, (ns, n))
        r = rng.normal(size=(ns, n))
        ob, oc, orr = b.copy(), c.copy(), r.copy()
        L.orc_pll_tridiag(ns, n, _p(a), _p(ob), _p(oc), _p(orr))
        print("[pll_tridiag] %d x %d call %d" % (ns, n, rep + 1), flush=True)
        gpu_pll_tridiag(a, b, c, r)
        assert np.array_equal(r, orr) and np.array_equal(b, ob) and np.array_equal(c, oc), (ns, n, rep)


def test_tridiag_1rhs_bitwise_vs_oracle(oracle_mod):
    """tridiag_1rhs: solution in rhs equal to the oracle's bit for bit; a, b
    and c are left as they were (pllcart3d.c:3511-3536)."""
    L = _lib()
    O = oracle_mod.load()
    rng = np.random.default_rng(11)
    for n in (1, 2, 3, 17, 128, 255, 4096):
        a = -rng.uniform(0, 1, n)
        c = -rng.uniform(0, 1, n)
        b = 2.1 + rng.uniform(0, 1, n)
        r = rng.normal(size=n)
        a0, b0, c0 = a.copy(), b.copy(), c.copy()
        x = r.copy()
        O.orc_tridiag_1rhs(n, _p(a0.copy()), _p(b0.copy()), _p(c0.copy()), _p(x))
        print("[tridiag_1rhs] n=%d" % n, flush=True)
        assert L.tridiag_1rhs(n, _p(a), _p(b), _p(c), _p(r)) == 0
        assert np.array_equal(r, x), n
        assert np.array_equal(a, a0) and np.array_equal(b, b0) and np.array_equal(c, c0), n


def test_calculate_weno_bitwise_vs_oracle(oracle_mod):
    """calculate_weno on random, smooth and discontinuous D1 arrays, both
    upwind signs: the device value equals the oracle's in every bit."""
    L = _lib()
    O = oracle_mod.load()
    rng = np.random.default_rng(12)
    x = np.linspace(0, 1, 40)
    arrays = [rng.normal(size=40), np.sin(3 * x), np.where(x > 0.5, 1.0, 0.0), np.zeros(40),
              1e-8 * rng.normal(size=40)]
    n = 0
    for d1 in arrays:
        d1 = np.ascontiguousarray(d1)
        for kk in range(4, 36):
            for sign in (1, -1):
                want = O.orc_weno(_p(d1), kk, sign)
                got = L.calculate_weno(_p(d1), kk, sign)
                assert got == want or (np.isnan(got) and np.isnan(want)), (kk, sign, got, want)
                n += 1
    print("[calculate_weno] %d evaluations equal" % n, flush=True)
