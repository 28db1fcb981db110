"""GPU plain-FP64 path vs the FP64 CPU oracle, through the C-ABI.

Bar: <= 1e-11 relative L2 after every step in c, in the velocity (u, v, w
taken together as one vector field) and in phi.  Each velocity component
and the pressure gradient (one vector field) are held to ten times the
rounding floor the oracle itself shows on the same case and steps, measured
in the test (tests/rounding_floor.py: the oracle against itself with
FFT-ordered transforms, and with every state value moved by one ulp after
each step), never below 1e-11.  The plain path repeats the reference's
arithmetic with -ffp-contract=off; the remaining differences come from
exp()/sin() on the device vs the host libm, the partitioned ADI solver's
elimination order and the transforms' order of operations.
"""
import numpy as np
import pytest

import pllcart3d_amd as P
from tests import rounding_floor as RF
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


def _bars(oracle_mod, prm, nsteps, recipe=P.IC_NOISE, seed=3, amp=0.05):
    """The case's bars: 10x the oracle's own rounding floor over nsteps."""
    bars, floors = RF.live_bars(oracle_mod, prm, P.ic_fill(prm, recipe, seed, amp), nsteps)
    print("[floors] %s" % " ".join("%s %.1e" % kv for kv in floors.items()), flush=True)
    return bars


def _sync_state(orc, sim):
    for f in STATE:
        sim.upload(f, orc.field(f).copy())


def _compare(orc, sim, fields, tol=TOL, what=""):
    """fields: names, or '+'-joined names compared as one stacked vector"""
    worst = {}
    for f in fields:
        parts = f.split("+")
        a = np.concatenate([sim.download(p).ravel() for p in parts])
        b = np.concatenate([orc.field(p).ravel() for p in parts])
        e = rel_l2(a, b)
        worst[f] = e
        assert e <= tol, "%s %s rel L2 %.3e > %.1e" % (what, f, e, tol)
    return worst


def _compare_step(orc, sim, what, bars):
    """c, velocity, phi to 1e-11; pg and each component to the case's bars."""
    e = {}
    for q, parts in RF.QUANTITIES.items():
        a = np.concatenate([sim.download(p).ravel() for p in parts])
        b = np.concatenate([orc.field(p).ravel() for p in parts])
        e[q] = rel_l2(a, b)
    print("[%s] %s" % (what, " ".join("%s %.1e" % kv for kv in e.items())), flush=True)
    for q, v in e.items():
        assert v <= bars[q], "%s %s rel L2 %.3e > %.1e" % (what, q, v, bars[q])


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
            # keep the two states identical so each operator is compared alone
            _sync_state(orc, sim)


@pytest.mark.parametrize("solver", ["part", "thomas"])
def test_fused_steps_track_oracle(oracle_mod, solver, monkeypatch):
    """Whole steps of the fused path vs the oracle, compared after every
    step, with the partitioned and the one-thread-per-line ADI solvers."""
    if solver == "thomas":
        monkeypatch.setenv("PLLC_LINE_SOLVER", "thomas")
    prm, orc, sim = _pair(oracle_mod)
    bars = _bars(oracle_mod, prm, 12)
    for n in range(12):
        to, dto = orc.run(1)
        tg, dtg = sim.run(1)
        assert abs(dtg - dto) <= 1e-12 * dto
        _compare_step(orc, sim, "%s step %d" % (solver, n + 1), bars)


def test_body_force_steps_track_oracle(oracle_mod):
    """Fx, Fy != 0: the body force -F/Re c of u and v (the only place c
    enters the momentum sources); whole steps vs the oracle."""
    prm, orc, sim = _pair(oracle_mod, nx=24, ny=12, nz=10, Fx=0.7, Fy=-0.4)
    bars = _bars(oracle_mod, prm, 3)
    for k in range(3):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "body force step %d" % (k + 1), bars)


def test_wide_planes_track_oracle(oracle_mod):
    """ny nz = 67600 > 65536: the bulk velocity takes its two-pass form
    (per-block candidates, then one block), and the transforms the dense
    contractions at a plane size that is no multiple of their tiles; whole
    steps vs the oracle, a 16x16-mode front over 12 x-cells (w nonzero)."""
    kw = dict(recipe=P.IC_MULTIMODE, seed=5, amp=0.05)
    prm, orc, sim = _pair(oracle_mod, nx=12, ny=260, nz=260, Inf_thick=0.3, Pe=1000.0, **kw)
    bars = _bars(oracle_mod, prm, 2, **kw)
    for k in range(2):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "12x260x260 step %d" % (k + 1), bars)


def test_nmvp_steps_track_oracle(oracle_mod):
    """Non-monotonic viscosity law (Nmvp = 1): the slope parameter found at
    set-up and whole fused steps."""
    prm, orc, sim = _pair(oracle_mod, Nmvp=1, Mum=1.5, Lambda=0.0)
    bars = _bars(oracle_mod, prm, 6)
    print("[nmvp] cm oracle %.15f gpu %.15f" % (orc.cm, sim.cm), flush=True)
    assert sim.cm == orc.cm and 0.1 < sim.cm < 0.9
    _compare(orc, sim, ["MUO_CELL", "MUO_XY", "MUO_XZ", "MUO_YZ"], what="nmvp init")
    for n in range(6):
        orc.run(1)
        sim.run(1)
        _compare(orc, sim, ["MUO_CELL", "MUO_XY", "MUO_XZ", "MUO_YZ"], what="nmvp step %d" % (n + 1))
        _compare_step(orc, sim, "nmvp step %d" % (n + 1), bars)


@pytest.mark.parametrize("shape", [(64, 40, 36), (96, 24, 50)])
def test_multi_chunk_lines_track_oracle(oracle_mod, shape):
    """Lines cut into several chunks of unequal length (x: 4-6 chunks,
    y and z: 2-4) by the partitioned ADI solver."""
    nx, ny, nz = shape
    prm, orc, sim = _pair(oracle_mod, nx=nx, ny=ny, nz=nz)
    bars = _bars(oracle_mod, prm, 3)
    for n in range(3):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "%s step %d" % ("x".join(map(str, shape)), n + 1), bars)


@pytest.mark.parametrize("subharm", [0, 1])
def test_shift_conc_matches_oracle(oracle_mod, subharm):
    """The wave perturbation (shift_conc + spline) after a few steps, then
    more steps from the shifted state."""
    prm, orc, sim = _pair(oracle_mod)
    bars = _bars(oracle_mod, prm, 4)
    orc.run(2)
    sim.run(2)
    orc.shift_conc(0.1, 0.7, subharm)
    sim.shift_conc(0.1, 0.7, subharm)
    e = rel_l2(sim.download("C_OLD"), orc.field("C_OLD"))
    print("[shift] subharm %d c_old rel L2 %.2e" % (subharm, e), flush=True)
    assert e <= 1e-14
    for n in range(2):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "shift step %d" % (n + 1), bars)


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_graph_replays_equal_eager_steps(precision, monkeypatch):
    """Pairs of steps replayed from the captured HIP graph equal eager steps
    bit for bit, replay after replay, with downloads (eager copies and
    synchronisations) between the replays."""
    prm = P.make_params(**SMALL)
    c0 = P.ic_fill(prm, P.IC_NOISE, 5, 0.05)
    monkeypatch.setenv("PLLC_GRAPH", "0")
    eager = P.Sim(prm, precision=precision)
    monkeypatch.setenv("PLLC_GRAPH", "1")
    graph = P.Sim(prm, precision=precision)
    eager.init(c0)
    graph.init(c0)
    for rep in range(4):
        eager.run(2)
        graph.run(2)
        for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "MUO_CELL"):
            a, b = graph.download(f), eager.download(f)
            assert np.array_equal(a, b), (precision, rep, f, np.abs(a - b).max())
        print("[graph] %s replay %d equal" % (precision, rep + 1), flush=True)
    # an odd count (one eager step) and then pairs again from the other parity
    eager.run(3)
    graph.run(3)
    for f in ("C_OLD", "U_OLD", "PHI"):
        assert np.array_equal(graph.download(f), eager.download(f)), (precision, "odd", f)


@pytest.mark.parametrize("graph", ["1", "0"])
@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_side_stream_momentum_equals_one_stream(precision, graph, monkeypatch):
    """On grids of 600k cells and more the fused step runs the three
    momentum explicit parts on a side stream beside the concentration update
    and the viscosity; with and without it (PLLC_FORK=0), every field is
    equal bit for bit after 200 graph-replayed or 5 eager steps (96x96x96,
    885k cells, 16x16-mode front)."""
    prm = P.make_params(96, 96, 96, Vratio=3.0, Pe=2000.0, Inf_thick=0.05)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 2, 0.01)
    monkeypatch.setenv("PLLC_GRAPH", graph)
    monkeypatch.setenv("PLLC_FORK", "0")
    one = P.Sim(prm, precision=precision)
    monkeypatch.delenv("PLLC_FORK")
    two = P.Sim(prm, precision=precision)
    one.init(c0)
    two.init(c0)
    nsteps = 200 if graph == "1" else 5
    one.run(nsteps)
    two.run(nsteps)
    for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "W_PG", "MUO_CELL"):
        a, b = two.download(f), one.download(f)
        assert np.array_equal(a, b), (precision, graph, nsteps, f, np.abs(a - b).max())
    one.close()
    two.close()


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_ensemble_members_equal_standalone_runs(precision):
    """Members of different grids, Vratio and Pe advanced together equal the
    same simulations run alone, bit for bit."""
    from pllcart3d_amd.configs import make_params
    specs = [make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1),
             make_params(32, 16, 12, Vratio=1.5, Pe=800.0, Inf_thick=0.08),
             make_params(40, 12, 20, Vratio=2.5, Pe=250.0, Inf_thick=0.1)]
    ics = [P.ic_fill(p, P.IC_NOISE, 7 + i, 0.03) for i, p in enumerate(specs)]
    ens = P.Ensemble(specs, precision)
    alone = []
    for p, c0, m in zip(specs, ics, ens.members):
        m.init(c0)
        s = P.Sim(p, precision)
        s.init(c0)
        alone.append(s)
    ens.run(5)
    for s in alone:
        s.run(5)
    for i, (m, s) in enumerate(zip(ens.members, alone)):
        for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG"):
            assert np.array_equal(m.download(f), s.download(f)), (precision, i, f)
    print("[ensemble] %s 3 members x 5 steps equal" % precision, flush=True)
    ens.close()


def test_autotuned_shapes_track_oracle(oracle_mod):
    """The autotuner picks the sweeps' lines per block by timing; simulations
    created afterwards with those shapes still follow the oracle."""
    shape = dict(nx=64, ny=40, nz=36)
    prm, _, tmp = _pair(oracle_mod, **shape)
    lpb = tmp.autotune(1)
    print("[autotune] lines per block x/y/z", lpb, flush=True)
    assert all(v in (8, 16, 32, 64) for v in lpb)
    tmp.close()
    prm, orc, sim = _pair(oracle_mod, **shape)
    bars = _bars(oracle_mod, prm, 2)
    for n in range(2):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "tuned step %d" % (n + 1), bars)


@pytest.mark.parametrize("shape", [(37, 9, 13), (16, 3, 2), (20, 5, 33)])
def test_ragged_and_thin_grids_track_oracle(oracle_mod, shape):
    """Sizes that are not multiples of any tile, lines too short for the
    partitioned solver (nz = 2: the w lines have one row, so those sweeps
    fall back to one thread per line), odd ny."""
    nx, ny, nz = shape
    prm, orc, sim = _pair(oracle_mod, nx=nx, ny=ny, nz=nz)
    bars = _bars(oracle_mod, prm, 3)
    for n in range(3):
        orc.run(1)
        sim.run(1)
        _compare_step(orc, sim, "%s step %d" % ("x".join(map(str, shape)), n + 1), bars)


OPS_SEQ = ["CONCENTRATION", "VISCOSITY", "XMOM", "YMOM", "ZMOM", "VELO_BOUNDARY", "DIVERGENCE", "PHISOLVER",
           "PHI_GHOSTS", "VELO_CORRECTION", "VELOGHOST", "PRESSURE_GRADIENT", "UPDATE_OLD"]


@pytest.mark.parametrize("precision", ["plain", "mixed"])
@pytest.mark.parametrize("shape", [(24, 12, 10), (8, 128, 128), (8, 256, 256)])
def test_fused_step_equals_operator_sequence(precision, shape):
    """The fused step (pointer swaps, batched u/v/w sweeps, the ghost copies
    written by the x-sweeps, the correction kernel and the last inverse
    transform, velo_correction and pressure_gradient in one kernel) against
    the reference's call sequence one operator at a time: equal in every
    bit of every field incl. ghosts, in both precisions, with the dense
    (24x12x10) and the fast (8x128x128 one plane per block, 8x256x256 lines)
    transforms."""
    prm = P.make_params(*shape, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    c0 = P.ic_fill(prm, P.IC_NOISE, 3, 0.05)
    a = P.Sim(prm, precision)
    a.init(c0)
    b = P.Sim(prm, precision)
    b.init(c0)
    a.run(2)
    for _ in range(2):
        dt = min(b.op("FIND_CFL"), b.timestep)
        for rk in range(3):
            vb = b.op("BULK_VELO", rk=rk)
            for op in OPS_SEQ:
                b.op(op, dt, rk, vb)
    print("[fused = ops] %s %s" % (precision, "x".join(map(str, shape))), flush=True)
    for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "V_PG", "W_PG", "MUO_CELL", "MUO_XY"):
        x, y = a.download(f), b.download(f)
        assert np.array_equal(x, y), (precision, f, np.abs(x - y).max())
