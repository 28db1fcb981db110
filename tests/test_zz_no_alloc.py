"""pll_tridiag_noalloc (no allocation: lines staged in the library's static
device array) against pll_tridiag through the C-ABI: equal in every bit,
on two calls in a row, the second smaller; and its size limit."""
import ctypes as C

import numpy as np
import pytest

import pllcart3d_amd as P

dpt = C.POINTER(C.c_double)
MAX_ROWS = 1 << 20


class Tds(C.Structure):
    # structures_3d.h:2-10
    _fields_ = [("acof", dpt), ("bcof", dpt), ("ccof", dpt), ("rhs", dpt)]


def _lib():
    L = P.load()
    L.pll_tridiag.argtypes = [C.POINTER(Tds), C.c_int, C.c_int]
    L.pll_tridiag.restype = None
    L.pll_tridiag_noalloc.argtypes = [C.POINTER(Tds), C.c_int, C.c_int]
    L.pll_tridiag_noalloc.restype = C.c_int
    return L


def _tds(a, b, c, r):
    arr = (Tds * a.shape[0])()
    for s in range(a.shape[0]):
        arr[s] = Tds(*(x[s].ctypes.data_as(dpt) for x in (a, b, c, r)))
    return arr


def _system(rng, ns, n):
    a = -rng.uniform(0, 1, (ns, n))
    c = -rng.uniform(0, 1, (ns, n))
    b = 2.05 + rng.uniform(0, 1, (ns, n))
    r = rng.normal(size=(ns, n))
    return a, b, c, r


def test_size_limit_is_rejected_without_a_gpu_call():
    """A system longer than the static array's quarter is refused before any
    device work, with an error naming the limit."""
    L = _lib()
    n = MAX_ROWS + 1
    arr = (Tds * 1)()
    assert L.pll_tridiag_noalloc(arr, 1, n) == -1
    assert "PLLC_NOALLOC_MAX_ROWS" in P.last_error()


@pytest.mark.gpu
def test_noalloc_equals_pll_tridiag_twice():
    """Two calls in a row (the second with fewer and shorter systems, so the
    array holds the first call's values beyond the second's): b, c and r equal
    pll_tridiag's bit for bit.  The first call's 5000 systems of 256 rows take
    two passes through the array (4096 systems per pass)."""
    L = _lib()
    rng = np.random.default_rng(21)
    for ns, n in ((5000, 256), (37, 100)):
        a, b, c, r = _system(rng, ns, n)
        ref = [x.copy() for x in (a, b, c, r)]
        L.pll_tridiag(_tds(*ref), ns, n)
        print("[noalloc] %d systems of %d" % (ns, n), flush=True)
        assert L.pll_tridiag_noalloc(_tds(a, b, c, r), ns, n) == 0, P.last_error()
        for got, want, name in zip((b, c, r), ref[1:], "bcr"):
            assert np.array_equal(got, want), (ns, n, name)
