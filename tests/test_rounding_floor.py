"""The oracle's rounding floors (tests/rounding_floor.py), which set the
GPU parity bars of the velocity components and the pressure gradient.

The FFT-ordered Poisson solve is first pinned as a solve of the same
system (equal to the oracle's phi to rounding); then the floors of the
cheaper cases are measured again and must agree with the FLOORS table
within a factor of two either way."""
import os

import numpy as np
import pytest

import pllcart3d_amd as P
from tests import rounding_floor as RF


@pytest.fixture(scope="module")
def orc_threads(oracle_mod):
    oracle_mod.set_threads(min(8, len(os.sched_getaffinity(0))))
    return oracle_mod


def test_fft_poisson_solves_the_same_system(orc_threads):
    """poisson_fft and the oracle's phisolver_dct on the same divergence:
    phi equal to 1e-13 relative L2 (so the "fft" floor measures rounding
    only); a wrong eigenvalue, boundary row or normalisation fails."""
    O = orc_threads
    prm = P.make_params(12, 40, 24, Vratio=3.0, Pe=1000.0, Inf_thick=0.3)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.05)
    a = O.Oracle(prm)
    a.init(c0)
    a.run(2)
    dt = min(a.op("FIND_CFL"), a.timestep)
    vb = a.op("BULK_VELO")
    for op in RF.SUBSTAGE[:7]:
        a.op(op, dt, 1, vb)
    div = a.field("DIV").copy()
    assert np.abs(div[3:-3, 1:-1, 1:-1]).max() > 0
    a.op("PHISOLVER", dt, 1, vb)
    want = a.field("PHI").copy()
    a.field("PHI")[...] = 0.0
    RF.poisson_fft(a, prm, dt, 1)
    got = a.field("PHI")
    e = RF.rel_l2(got[3:-3, 1:-1, 1:-1], want[3:-3, 1:-1, 1:-1])
    print("fft vs dense phi %.2e" % e)
    assert e < 1e-13


def _check_table(case, got):
    for q, v in got.items():
        ref = RF.FLOORS[case][q]
        print("%s %-8s measured %.2e table %.2e" % (case, q, v, ref))
        assert ref / 2 <= v <= 2 * ref, (case, q, v, ref)


def test_floor_table_24x128x128(orc_threads):
    _check_table("modes_24x128x128", RF.measure_case(orc_threads, "modes_24x128x128"))


def test_floor_noise_64x32x32_component_peak(orc_threads):
    """The largest component floor of the 64x32x32 case (w, at the first
    step of the FFT-ordered solve) and the bar it sets."""
    prm, c0, _, _ = RF.case_setup("noise_64x32x32")
    w = RF.measure(orc_threads, prm, c0, 3, "fft")
    ref = RF.FLOORS["noise_64x32x32"]["W_OLD"]
    print("noise_64x32x32 W_OLD measured %.2e table %.2e" % (w["W_OLD"], ref))
    assert ref / 2 <= w["W_OLD"] <= 2 * ref
    assert RF.bar("noise_64x32x32", "W_OLD") == pytest.approx(10 * ref)
    assert RF.bar("noise_64x32x32", "c") == 1e-11
