"""The drop-in DCT entry points (plan_dct_3d / fdct_3d / idct_3d,
fftw_3d.h:12-106) on the matrix-core contraction, against scipy's DCT-II /
DCT-III (the FFTW REDFT10 / REDFT01 definitions, unnormalised) and their
round trip with the 1/(4 n1 n2) normalisation of idct_3d."""
import ctypes as C
import os

import numpy as np
import pytest
import scipy.fft

import pllcart3d_amd as P

pytestmark = pytest.mark.gpu

dp = C.POINTER(C.c_double)
dppp = C.POINTER(C.POINTER(dp))


def _lib():
    L = C.CDLL(os.path.join(os.path.dirname(P.__file__), "libpllcart3d_amd.so"))
    L.alloc3d_dble.restype = dppp
    L.alloc3d_dble.argtypes = [C.c_int, C.c_int, C.c_int]
    L.plan_dct_3d.argtypes = [C.c_int, C.c_int, C.c_int]
    L.fdct_3d.argtypes = [C.c_int, C.c_int, C.c_int, dppp, dppp]
    L.idct_3d.argtypes = [C.c_int, C.c_int, C.c_int, dppp, dppp]
    L.pllc_last_error.restype = C.c_char_p
    return L


def _arr(L, n1, n2, nt):
    a = L.alloc3d_dble(n1, n2, nt)
    view = np.ctypeslib.as_array(a[0][0], shape=(n1, n2, nt))   # one contiguous block
    return a, view


@pytest.mark.parametrize("n1,n2,nt", [(12, 10, 7), (64, 48, 5), (128, 128, 3), (33, 130, 2)])
def test_fdct_idct_match_scipy(n1, n2, nt):
    L = _lib()
    rng = np.random.default_rng(n1 * 1000 + n2)
    a, av = _arr(L, n1, n2, nt)
    b, bv = _arr(L, n1, n2, nt)
    c, cv = _arr(L, n1, n2, nt)
    x = rng.normal(size=(n1, n2, nt))
    av[...] = x
    print("[dct] %dx%dx%d" % (n1, n2, nt), flush=True)
    L.plan_dct_3d(n1, n2, nt)
    L.fdct_3d(n1, n2, nt, a, b)
    assert L.pllc_last_error() in (b"", None)
    ref = scipy.fft.dctn(x, type=2, axes=(0, 1))
    e = np.linalg.norm(bv - ref) / np.linalg.norm(ref)
    assert e < 1e-14, e
    L.idct_3d(n1, n2, nt, b, c)
    ref3 = scipy.fft.dctn(bv, type=3, axes=(0, 1)) / (4.0 * n1 * n2)
    assert np.linalg.norm(cv - ref3) / np.linalg.norm(ref3) < 1e-14
    assert np.linalg.norm(cv - x) / np.linalg.norm(x) < 1e-13


@pytest.mark.parametrize("shape", [(8, 64, 64), (12, 32, 48), (8, 16, 64)])
@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_small_plane_transform_equals_two_launch_path(monkeypatch, shape, precision):
    """Planes of at most 64 x 64 (multiples of 16) run both dense
    contractions of the Poisson solve in one launch (k_dct_planes: square
    planes and the any-shape form); it makes the MFMA sequence of the two
    k_gemm launches (PLLC_DCT=gemm), so three steps from a state that varies
    in z leave every field equal in every bit."""
    nx, ny, nz = shape
    prm = P.make_params(nx, ny, nz, Vratio=3.0, Pe=1000.0, Inf_thick=0.05)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.02)
    out = {}
    for mode in ("gemm", "default"):
        if mode == "gemm":
            monkeypatch.setenv("PLLC_DCT", "gemm")
        else:
            monkeypatch.delenv("PLLC_DCT", raising=False)
        print("[dct_small] %s %s %s" % (shape, precision, mode), flush=True)
        s = P.Sim(prm, precision)
        s.init(c0)
        s.run(3)
        out[mode] = {f: s.download(f) for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "V_PG", "W_PG")}
        s.close()
    for f, a in out["gemm"].items():
        assert np.array_equal(a, out["default"][f]), (shape, precision, f)
