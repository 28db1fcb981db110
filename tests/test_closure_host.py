"""Host-side pieces of the learned closure, without a GPU: the guard's test
(out-of-range c, NaN, mass injected by the correction), the mixing-zone
length (closed form on an erf profile) and the box filter (cell means,
face fluxes)."""
import ctypes as C
import math

import numpy as np
import pytest

import pllcart3d_amd as P
from pllcart3d_amd import closure as CL


def _guard(stats, dt=1e-3, tol_bound=1e-2, tol_mass=1e-4):
    s = np.ascontiguousarray(stats, dtype=np.float64)
    return P.load().pllc_closure_guard_ok(s.ctypes.data_as(C.POINTER(C.c_double)), dt, tol_bound, tol_mass)


def test_guard_accepts_bounded_conservative_states():
    assert _guard([0.0, 1.0, 5000.0, 0.0]) == 1
    assert _guard([-0.005, 1.009, 5000.0, 100.0]) == 1      # |dt corr| = 0.1 <= 1e-4 * 5000


@pytest.mark.parametrize("stats", [[-0.02, 1.0, 5000.0, 0.0],      # c below -tol_bound
                                   [0.0, 1.5, 5000.0, 0.0],        # c above 1 + tol_bound
                                   [0.0, 1.0, 5000.0, 600.0],      # dt * 600 > 1e-4 * 5000
                                   [0.0, 1.0, 5000.0, -600.0],
                                   [float("nan"), 1.0, 5000.0, 0.0],
                                   [0.0, 1.0, float("nan"), 0.0],
                                   [0.0, float("inf"), 5000.0, 0.0]])
def test_guard_rejects_out_of_range_c_nan_and_mass_injection(stats):
    """Injected out-of-range, non-finite or mass-injecting states fail."""
    assert _guard(stats) == 0


def _profile_field(p, prof):
    c = np.zeros((p.nx + 6, p.ny + 2, p.nz + 2))
    c[3:3 + p.nx] = prof[:, None, None]
    return c


def test_mixing_length_of_an_erf_profile():
    """c = 0.5 (1 + erf((x - x0)/w)) averaged over y and z: the 0.05 and 0.95
    points are x0 -/+ w erfinv(0.9), so the length is 2 w erfinv(0.9) (the
    linear interpolation between cell centres errs by O(dx^2))."""
    from scipy.special import erf, erfinv
    p = P.make_params(2048, 4, 4)
    dx = (p.X2 - p.X1) / p.nx
    xc = p.X1 + (np.arange(p.nx) + 0.5) * dx
    w = 0.3
    prof = 0.5 * (1 + erf((xc - 1.7) / w))
    L = CL.mixing_length(_profile_field(p, prof), p)
    assert abs(L - 2 * w * erfinv(0.9)) < 1e-5
    # the displacing fluid entering at X1 (decreasing profile): same length
    assert abs(CL.mixing_length(_profile_field(p, 1 - prof), p) - L) < 1e-12


def test_box_filter_keeps_cell_means_and_face_fluxes():
    rng = np.random.default_rng(2)
    fp = P.make_params(8, 6, 4)
    f = rng.normal(size=(fp.nx + 6, fp.ny + 2, fp.nz + 2))
    cm = CL.filter_cells(f, fp)
    assert cm.shape == (4, 3, 2)
    assert np.isclose(cm[1, 2, 0], f[3 + 2:3 + 4, 1 + 4:1 + 6, 1:3].mean())
    # u faces: coarse face ic is fine face 2 ic (array index 3 + 2 ic), the
    # flux through it (mean x area) kept
    uf = CL.filter_faces(f, fp, 0)
    assert uf.shape == (5, 3, 2)
    assert np.isclose(uf[2, 1, 1], f[3 + 4, 1 + 2:1 + 4, 1 + 2:1 + 4].mean())
    assert np.isclose(uf[4, 0, 0], f[3 + 8, 1:3, 1:3].mean())     # the outlet face
    vf = CL.filter_faces(f, fp, 1)
    assert vf.shape == (4, 4, 2) and np.isclose(vf[1, 3, 0], f[3 + 2:3 + 4, 1 + 6, 1:3].mean())
    wf = CL.filter_faces(f, fp, 2)
    assert wf.shape == (4, 3, 3) and np.isclose(wf[0, 0, 2], f[3:5, 1:3, 1 + 4].mean())


def test_parameter_layout_round_trip():
    Pd = CL.init_params(3)
    Pd["b3"] = np.array([0.25])
    flat = CL.flat_from_dict(Pd)
    assert flat.size == 3 * 27 * 256 + 48
    back = CL.dict_from_flat(flat)
    for k in Pd:
        assert np.array_equal(back[k], Pd[k]), k
    assert math.isclose(flat[3 * 27 * 256 + 32], 0.25)
