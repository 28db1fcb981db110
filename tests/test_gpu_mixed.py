"""Mixed-precision path (FP32 fields, stencils and line solves; FP64 CFL,
flux and bulk reductions and the (0,0) Poisson pencil) vs the FP64 oracle.

The bar, written here: after 10 steps the concentration is within 1e-5
relative L2 of the FP64 solution (north_star's bar) and the velocity (u, v,
w as one vector field) within 1e-5.  phi, which north_star gives no bar, is
held to 1e-3 as a sanity bound (measured up to 1.9e-4, at 24x256x256).
Every field's own error is printed so its size is on record; the spanwise
component w is ~1e-4 of u at these states, so its own relative error
carries the velocity's FP32 rounding amplified by |u|/|w| and is held to
1e-2.
"""
import numpy as np
import pytest

import pllcart3d_amd as P
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {"C_OLD": 1e-5, "U_OLD+V_OLD+W_OLD": 1e-5, "PHI": 1e-3, "U_OLD": 1e-5, "V_OLD": 1e-4, "W_OLD": 1e-2}


def _err(sim, orc, name):
    parts = name.split("+")
    a = np.concatenate([sim.download(p).ravel() for p in parts])
    b = np.concatenate([orc.field(p).ravel() for p in parts])
    return rel_l2(a, b)


@pytest.mark.parametrize("shape", [(32, 16, 16), (64, 32, 32)])
def test_mixed_tracks_fp64_oracle(oracle_mod, shape):
    prm = P.make_params(*shape, Vratio=2.0, Pe=500.0, Inf_thick=0.05)
    c0 = P.ic_fill(prm, P.IC_NOISE, 1, 0.01)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm, "mixed")
    sim.init(c0)
    for n in range(10):
        orc.run(1)
        sim.run(1)
    errs = {f: _err(sim, orc, f) for f in TOL}
    print("[mixed %s] after 10 steps: %s" % (shape, " ".join("%s=%.2e" % kv for kv in errs.items())), flush=True)
    for f, tol in TOL.items():
        assert errs[f] <= tol, (f, errs[f])


def test_mixed_thomas_tracks_fp64_oracle(oracle_mod, monkeypatch):
    """PLLC_LINE_SOLVER=thomas (one thread per line, the reference's
    arithmetic in the line solves and explicit stencils) in FP32 against the
    FP64 oracle after 10 steps, to the bars above."""
    monkeypatch.setenv("PLLC_LINE_SOLVER", "thomas")
    prm = P.make_params(32, 16, 16, Vratio=2.0, Pe=500.0, Inf_thick=0.05)
    c0 = P.ic_fill(prm, P.IC_NOISE, 1, 0.01)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm, "mixed")
    sim.init(c0)
    for n in range(10):
        orc.run(1)
        sim.run(1)
    errs = {f: _err(sim, orc, f) for f in TOL}
    print("[mixed thomas] after 10 steps: %s" % " ".join("%s=%.2e" % kv for kv in errs.items()), flush=True)
    for f, tol in TOL.items():
        assert errs[f] <= tol, (f, errs[f])


def test_mixed_nmvp_tracks_fp64_oracle(oracle_mod):
    """The non-monotonic viscosity law (Nmvp = 1: slope parameter found on
    the host in FP64, the law evaluated in FP32 by the viscosity kernel)
    against the FP64 oracle after 10 steps, to the bars above."""
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1, Nmvp=1, Mum=1.5, Lambda=0.0)
    c0 = P.ic_fill(prm, P.IC_NOISE, 3, 0.05)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm, "mixed")
    sim.init(c0)
    for n in range(10):
        orc.run(1)
        sim.run(1)
    errs = {f: _err(sim, orc, f) for f in TOL}
    print("[mixed nmvp] after 10 steps: %s" % " ".join("%s=%.2e" % kv for kv in errs.items()), flush=True)
    for f, tol in TOL.items():
        assert errs[f] <= tol, (f, errs[f])


@pytest.mark.parametrize("n", [128, 256])
def test_mixed_square_planes_track_fp64_oracle(oracle_mod, n):
    """ny = nz = n (the fast transform in FP32: one plane per block at 128,
    lines at 256) on a 16x16-mode front over 24 x-cells, whose planes vary
    in y and z: c, the velocity vector and phi after 10 steps against the
    FP64 oracle, to the bars above."""
    prm = P.make_params(24, n, n, Vratio=3.0, Pe=1000.0, Inf_thick=0.3)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.05)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm, "mixed")
    sim.init(c0)
    for k in range(10):
        orc.run(1)
        sim.run(1)
    errs = {f: _err(sim, orc, f) for f in ("C_OLD", "U_OLD+V_OLD+W_OLD", "PHI")}
    print("[mixed 24x%dx%d] after 10 steps: %s" % (n, n, " ".join("%s=%.2e" % kv for kv in errs.items())), flush=True)
    for f, e in errs.items():
        assert e <= TOL[f], (f, e)
