"""GPU plain-FP64 path vs the FP64 CPU oracle, through the C-ABI.

Bar: <= 1e-11 relative L2 per field (the plain path repeats the
reference's arithmetic with -ffp-contract=off; the remaining differences
come from exp() and from the DCT summation order).
"""
import numpy as np
import pytest

import pllcart3d_amd as P
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-11

STATE = [f for f in P.FIELDS]
SMALL = dict(nx=24, ny=12, nz=10, Vratio=3.0, Pe=300.0, Inf_thick=0.1)


def _pair(oracle_mod, recipe=P.IC_NOISE, seed=3, amp=0.05, **kw):
    cfg = dict(SMALL)
    cfg.update(kw)
    prm = P.make_params(**cfg)
    c0 = P.ic_fill(prm, recipe, seed, amp)
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm)
    sim.init(c0)
    return prm, orc, sim


def _sync_state(orc, sim):
    for f in STATE:
        sim.upload(f, orc.field(f).copy())


def _compare(orc, sim, fields, tol=TOL, what=""):
    worst = {}
    for f in fields:
        e = rel_l2(sim.download(f), orc.field(f))
        worst[f] = e
        assert e <= tol, "%s %s rel L2 %.3e > %.1e" % (what, f, e, tol)
    return worst


def test_init_matches(oracle_mod):
    prm, orc, sim = _pair(oracle_mod)
    print("[init] 24x12x10", flush=True)
    _compare(orc, sim, STATE, what="init")
    assert sim.timestep == orc.timestep


@pytest.mark.parametrize("nsteps_before", [0, 3])
def test_each_operator_from_shared_state(oracle_mod, nsteps_before):
    """Every operator of the substage body, one at a time from the same state."""
    prm, orc, sim = _pair(oracle_mod)
    if nsteps_before:
        orc.run(nsteps_before)
    _sync_state(orc, sim)
    dt = orc.op("FIND_CFL")
    assert abs(sim.op("FIND_CFL") - dt) <= 1e-15 * abs(dt)
    dt = min(dt, orc.timestep)
    print("[ops] before=%d dt=%g" % (nsteps_before, dt), flush=True)
    written = {
        "CONCENTRATION": ["C_NEW", "C_N"],
        "VISCOSITY": ["MUN_CELL", "MUN_XY", "MUN_XZ", "MUN_YZ"],
        "XMOM": ["U_NEW", "U_N"], "YMOM": ["V_NEW", "V_N"], "ZMOM": ["W_NEW", "W_N"],
        "VELO_BOUNDARY": ["U_NEW", "V_NEW", "W_NEW"], "DIVERGENCE": ["DIV"], "PHISOLVER": ["PHI"],
        "PHI_GHOSTS": ["PHI"], "VELO_CORRECTION": ["U_NEW", "V_NEW", "W_NEW"],
        "VELOGHOST": ["U_NEW", "V_NEW", "W_NEW"],
        "PRESSURE_GRADIENT": ["U_PG", "V_PG", "W_PG"], "UPDATE_OLD": STATE,
    }
    for rk in range(3):
        vb_o = orc.op("BULK_VELO", rk=rk)
        vb_g = sim.op("BULK_VELO", rk=rk)
        assert vb_o == vb_g
        for op in written:
            orc.op(op, dt, rk, vb_o)
            sim.op(op, dt, rk, vb_o)
            _compare(orc, sim, written[op], what="rk%d %s" % (rk, op))
            # keep the two states identical so each operator is judged alone
            _sync_state(orc, sim)


def test_fused_steps_track_oracle(oracle_mod):
    """Whole steps of the fused path vs the oracle, compared after every step."""
    prm, orc, sim = _pair(oracle_mod)
    for n in range(12):
        to, dto = orc.run(1)
        tg, dtg = sim.run(1)
        assert abs(dtg - dto) <= 1e-12 * dto
        w = _compare(orc, sim, ["C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "V_PG", "W_PG"],
                     what="step %d" % (n + 1))
        print("[steps] %d worst %.2e" % (n + 1, max(w.values())), flush=True)


def test_nmvp_steps_track_oracle(oracle_mod):
    """Non-monotonic viscosity law (Nmvp = 1): the slope parameter found at
    set-up and whole fused steps."""
    prm, orc, sim = _pair(oracle_mod, Nmvp=1, Mum=1.5, Lambda=0.0)
    print("[nmvp] cm oracle %.15f gpu %.15f" % (orc.cm, sim.cm), flush=True)
    assert sim.cm == orc.cm and 0.1 < sim.cm < 0.9
    _compare(orc, sim, ["MUO_CELL", "MUO_XY", "MUO_XZ", "MUO_YZ"], what="nmvp init")
    for n in range(6):
        orc.run(1)
        sim.run(1)
        w = _compare(orc, sim, ["C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "MUO_CELL", "MUO_YZ"],
                     what="nmvp step %d" % (n + 1))
        print("[nmvp] %d worst %.2e" % (n + 1, max(w.values())), flush=True)
