"""Mixed-precision path (FP32 fields, stencils and line solves; FP64 CFL,
flux and bulk reductions and the (0,0) Poisson pencil) vs the FP64 oracle.

The bar, written here: after 10 steps the concentration is within 1e-5
relative L2 of the FP64 solution (the required bar) and the velocity (u, v,
w as one vector field) within 1e-5.  phi, for which no bar is required, is
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


@pytest.mark.parametrize("n,dct", [(128, "fft"), (256, "fft"), (128, "mfma")])
def test_mixed_square_planes_track_fp64_oracle(oracle_mod, n, dct, monkeypatch):
    """ny = nz = n (the fast transform in FP32: one plane per block at 128,
    lines at 256; with PLLC_DCT=mfma at 128 the BF16x3 contractions on the
    matrix cores) on a 16x16-mode front over 24 x-cells, whose planes vary
    in y and z: c, the velocity vector and phi after 10 steps against the
    FP64 oracle, to the bars above."""
    if dct == "mfma":
        monkeypatch.setenv("PLLC_DCT", "mfma")
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
    print("[mixed 24x%dx%d %s] after 10 steps: %s" % (n, n, dct, " ".join("%s=%.2e" % kv for kv in errs.items())),
          flush=True)
    for f, e in errs.items():
        assert e <= TOL[f], (f, e)


def test_bf16x3_poisson_solve_vs_fp64_oracle(oracle_mod, monkeypatch):
    """One phisolver_dct from a shared state: phi of the mixed path with the
    BF16x3 matrix-core transform (PLLC_DCT=mfma) and with the FP32 fast
    transform against the FP64 oracle's; both printed, the BF16x3 one held to
    1e-4 relative L2 (its products keep about 16 significant bits)."""
    prm = P.make_params(12, 128, 128, Vratio=3.0, Pe=1000.0, Inf_thick=0.3)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.05)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    orc.run(2)
    dt = min(orc.op("FIND_CFL"), orc.timestep)
    vb = orc.op("BULK_VELO")
    for op in ("CONCENTRATION", "VISCOSITY", "XMOM", "YMOM", "ZMOM", "VELO_BOUNDARY", "DIVERGENCE"):
        orc.op(op, dt, 1, vb)
    state = {f: orc.field(f).copy() for f in P.FIELDS}
    orc.op("PHISOLVER", dt, 1, vb)
    want = orc.field("PHI")[3:-3, 1:-1, 1:-1]
    errs = {}
    for dct in ("fft", "mfma"):
        if dct == "mfma":
            monkeypatch.setenv("PLLC_DCT", "mfma")
        sim = P.Sim(prm, "mixed")
        sim.init(c0)
        for f, a in state.items():
            sim.upload(f, a)
        sim.op("PHISOLVER", dt, 1, vb)
        errs[dct] = rel_l2(sim.download("PHI")[3:-3, 1:-1, 1:-1], want)
        sim.close()
    print("[poisson 12x128x128] phi vs FP64: fft %.2e bf16x3 %.2e" % (errs["fft"], errs["mfma"]), flush=True)
    assert errs["mfma"] <= 1e-4 and errs["fft"] <= 1e-4
