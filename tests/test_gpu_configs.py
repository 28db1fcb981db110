"""The plain FP64 path at the run configurations' own sizes and states, vs
the FP64 oracle (OpenMP on the host's cores), through the C-ABI.

Bar (the required parity): after every step, relative L2 <= 1e-11 in c, in the
velocity (u, v, w as one vector field) and in phi.  Each velocity
component on its own and the pressure gradient (one vector field) are held
to ten times the rounding floor the oracle itself shows for the same case
(tests/rounding_floor.py: the oracle against itself with FFT-ordered
(y, z) transforms, and with every state value moved by one ulp after each
step), never below 1e-11.

  * front_256x128x128: 256x128x128, R=3, Pe=2000, front width 0.02 displaced
    by 16x16 seeded modes (seed 2), the benchmark's workload; STEPS_256
    steps (default 30; PLLC_CFG_STEPS overrides, 1000 is the workload's
    full run).
  * noise_64x32x32: 64x32x32, R=2, Pe=500, width 0.05, +-0.01 noise (seed
    1), all of its 200 steps.
  * front_512x256x256: 512x256x256, R=3, Pe=4000, the same front recipe,
    its first 3 steps.
  * planes that vary in z: ny = nz = 128 and 256 with 24 x-cells and a
    16x16-mode front, so the (y, z) transforms and the z-sweeps see 3D data.
"""
import os
import time

import numpy as np
import pytest

import pllcart3d_amd as P
from pllcart3d_amd.configs import CONFIGS, config_params, ensemble_members
from tests.conftest import rel_l2
from tests.rounding_floor import bar

pytestmark = pytest.mark.gpu

STEPS_256 = int(os.environ.get("PLLC_CFG_STEPS", "30"))
VEC = {"c": ["C_OLD"], "velocity": ["U_OLD", "V_OLD", "W_OLD"], "phi": ["PHI"],
       "pg": ["U_PG", "V_PG", "W_PG"]}
COMPONENTS = ["U_OLD", "V_OLD", "W_OLD"]


@pytest.fixture(scope="module")
def threads(oracle_mod):
    n = min(16, len(os.sched_getaffinity(0)))
    oracle_mod.set_threads(n)
    return n


def _errors(sim, orc):
    e = {}
    for k, parts in VEC.items():
        a = np.concatenate([sim.download(p).ravel() for p in parts])
        b = np.concatenate([orc.field(p).ravel() for p in parts])
        e[k] = rel_l2(a, b)
    for f in COMPONENTS:
        e[f] = rel_l2(sim.download(f), orc.field(f))
    return e


def _check(e, what, case):
    line = " ".join("%s %.2e" % kv for kv in e.items())
    print("[%s] %s" % (what, line), flush=True)
    for q, v in e.items():
        assert v <= bar(case, q), (what, q, v, bar(case, q))


def _config_pair(oracle_mod, name):
    c = CONFIGS[name]
    prm = config_params(name)
    c0 = P.ic_fill(prm, c["ic"], c["seed"], c["amp"])
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm)
    sim.init(c0)
    return prm, orc, sim


def _track(orc, sim, nsteps, tag, case, every=1):
    worst = {}
    t0 = time.time()
    for n in range(nsteps):
        _, dto = orc.run(1)
        _, dtg = sim.run(1)
        assert abs(dtg - dto) <= 1e-12 * dto, (n, dtg, dto)
        if (n + 1) % every == 0 or n + 1 == nsteps:
            e = _errors(sim, orc)
            for k, v in e.items():
                worst[k] = max(worst.get(k, 0.0), v)
            _check(e, "%s step %d (%.0f s)" % (tag, n + 1, time.time() - t0), case)
    print("[%s] worst over %d steps: %s" % (tag, nsteps, " ".join("%s %.2e" % kv for kv in worst.items())),
          flush=True)
    return worst


def test_front_256x128x128_tracks_oracle(oracle_mod, threads):
    """The benchmark's workload, plain FP64, compared after every step."""
    print("[front_256x128x128] %d steps, oracle on %d threads" % (STEPS_256, threads), flush=True)
    prm, orc, sim = _config_pair(oracle_mod, "front_256x128x128")
    e = _errors(sim, orc)
    assert max(e.values()) == 0.0, e
    # the state is three-dimensional: w is not rounding noise around zero
    w = orc.field("W_OLD")
    print("[front_256x128x128] |w| max %.3e |u| max %.3e" % (np.abs(w).max(), np.abs(orc.field("U_OLD")).max()))
    _track(orc, sim, STEPS_256, "front_256x128x128", "front_256x128x128")
    assert np.abs(orc.field("W_OLD")).max() > 1e-3


def test_mixed_front_256x128x128_tracks_fp64_oracle(oracle_mod, threads):
    """The benchmark's workload in mixed precision against the FP64 oracle
    for 20 steps: the concentration within 1e-5 relative L2 (the required
    bar) and the velocity vector within 1e-5 after every 5th step; each
    step's dt (an FP64 reduction over the FP32 velocities) within 1e-5 of
    the oracle's."""
    c = CONFIGS["front_256x128x128"]
    prm = config_params("front_256x128x128")
    c0 = P.ic_fill(prm, c["ic"], c["seed"], c["amp"])
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm, "mixed")
    sim.init(c0)
    t0 = time.time()
    worst = {}
    for n in range(20):
        _, dto = orc.run(1)
        _, dtg = sim.run(1)
        assert abs(dtg - dto) <= 1e-5 * dto, (n, dtg, dto)
        if (n + 1) % 5 == 0:
            e = _errors(sim, orc)
            for k, v in e.items():
                worst[k] = max(worst.get(k, 0.0), v)
            print("[mixed front_256x128x128] step %d (%.0f s): %s" % (n + 1, time.time() - t0,
                  " ".join("%s %.2e" % kv for kv in e.items())), flush=True)
            assert e["c"] <= 1e-5 and e["velocity"] <= 1e-5, (n + 1, e)
    print("[mixed front_256x128x128] worst over 20 steps: %s" % " ".join("%s %.2e" % kv for kv in worst.items()),
          flush=True)
    sim.close()


def test_front_512x256x256_first_steps_track_oracle(oracle_mod, threads):
    """The 512x256x256 workload (33.5 M cells; its 256x256 planes take the
    line-pass transforms), plain FP64, its first 3 steps against the oracle
    (about 8 s per oracle step on 16 threads), compared after every step."""
    print("[front_512x256x256] 3 steps, oracle on %d threads" % threads, flush=True)
    prm, orc, sim = _config_pair(oracle_mod, "front_512x256x256")
    _track(orc, sim, 3, "front_512x256x256", "front_512x256x256")
    sim.close()


def test_noise_64x32x32_all_200_steps(oracle_mod, threads):
    """The 64x32x32 case in plain FP64 for all of its 200 steps, compared
    after every step."""
    prm, orc, sim = _config_pair(oracle_mod, "noise_64x32x32")
    _track(orc, sim, CONFIGS["noise_64x32x32"]["steps"], "noise_64x32x32", "noise_64x32x32")


@pytest.mark.parametrize("n", [128, 256])
@pytest.mark.parametrize("dct", ["fast", "gemm"])
def test_z_varying_square_planes_track_oracle(oracle_mod, threads, dct, n, monkeypatch):
    """ny = nz = n with a 16x16-mode front over 24 x-cells (front width 0.3,
    so several cells lie inside it and every (y, z) plane varies in both
    directions): the fast plane transform (n = 128 one block per plane,
    n = 256 line passes) or, with PLLC_DCT=gemm, the dense contractions on
    the matrix cores; whole steps against the oracle."""
    if dct == "gemm":
        monkeypatch.setenv("PLLC_DCT", "gemm")
    prm = P.make_params(24, n, n, Vratio=3.0, Pe=1000.0, Inf_thick=0.3)
    c0 = P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.05)
    core = c0[3:-3, 1:-1, 1:-1]
    assert np.abs(np.diff(core, axis=2)).max() > 1e-3 and np.abs(np.diff(core, axis=1)).max() > 1e-3
    orc = oracle_mod.Oracle(prm)
    orc.init(c0)
    sim = P.Sim(prm)
    sim.init(c0)
    _track(orc, sim, 4, "%s 24x%dx%d" % (dct, n, n), "modes_24x%dx%d" % (n, n))
    w = orc.field("W_OLD")[3:-3, 1:-1, 1:-1]
    assert np.abs(w).max() > 1e-4


@pytest.mark.parametrize("precision", ["plain", "mixed"])
def test_ensemble_members_at_workload_shapes_equal_standalone(precision):
    """Three members of the 64-member ensemble, one of each grid shape
    (64x32x32, 128x48x48, 192x64x64) with their drawn R and Pe, advanced
    together, equal the same simulations run alone in every bit."""
    members, seen = [], set()
    for p, ic, seed, amp in ensemble_members(64):
        key = (p.nx, p.ny, p.nz)
        if key not in seen:
            seen.add(key)
            members.append((p, ic, seed, amp))
    assert len(members) == 3
    ens = P.Ensemble([m[0] for m in members], precision)
    alone = []
    for (p, ic, seed, amp), m in zip(members, ens.members):
        c0 = P.ic_fill(p, ic, seed, amp)
        m.init(c0)
        s = P.Sim(p, precision)
        s.init(c0)
        alone.append(s)
    ens.run(6)
    for s in alone:
        s.run(6)
    for i, (m, s) in enumerate(zip(ens.members, alone)):
        for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG"):
            assert np.array_equal(m.download(f), s.download(f)), (precision, i, f)
    print("[ensemble] %s members %s x 6 steps equal" % (precision, sorted(seen)), flush=True)
    ens.close()


def test_ensemble_batches_equal_standalone():
    """Mixed precision: members of equal grids are stepped as batches (one
    launch per kernel for all of them, members taken from a grid dimension).
    The first 9 members of the 64-member workload (three to four per grid
    shape, each with its own drawn R, Pe and seed) advanced together equal
    the same simulations run alone in every bit: 6 steps (graph-replayed
    pairs), then 1 (eager), then an upload into one member (whose next step
    copies instead of swapping, so every member of its batch does) and 4
    more steps."""
    specs = ensemble_members(9)
    ens = P.Ensemble([m[0] for m in specs], "mixed")
    shapes = {}
    for p, *_ in specs:
        shapes[(p.nx, p.ny, p.nz)] = shapes.get((p.nx, p.ny, p.nz), 0) + 1
    nb = ens.batches()
    print("[ensemble batches] %d batches over %s" % (nb, shapes), flush=True)
    assert nb == sum(1 for v in shapes.values() if v >= 2)
    alone = []
    for (p, ic, seed, amp), m in zip(specs, ens.members):
        c0 = P.ic_fill(p, ic, seed, amp)
        m.init(c0)
        s = P.Sim(p, "mixed")
        s.init(c0)
        alone.append(s)

    def check(tag):
        for i, (m, s) in enumerate(zip(ens.members, alone)):
            for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "W_PG"):
                assert np.array_equal(m.download(f), s.download(f)), (tag, i, f)
        print("[ensemble batches] %s: every member equal" % tag, flush=True)

    ens.run(6)
    for s in alone:
        s.run(6)
    check("6 steps")
    ens.run(1)
    for s in alone:
        s.run(1)
    check("7 steps")
    v = alone[0].download("V_OLD") * 0.5
    ens.members[0].upload("V_OLD", v)
    alone[0].upload("V_OLD", v)
    ens.run(4)
    for s in alone:
        s.run(4)
    check("upload + 4 steps")
    ens.close()


def test_ensemble_batches_of_other_planes_equal_standalone():
    """Batches whose (y, z) planes take the other forms of the batched
    transform: 16 x 32 planes (any ny, nz) and 16 x 16 planes (square, one
    wave per block), each member with its own R, Pe and seed, equal the
    same simulations run alone in every bit after 5 steps."""
    grids = [(24, 16, 32)] * 3 + [(20, 16, 16)] * 2
    specs = []
    for i, (nx, ny, nz) in enumerate(grids):
        prm = P.make_params(nx, ny, nz, Vratio=1.5 + 0.5 * i, Pe=300.0 + 200.0 * i, Inf_thick=0.05)
        specs.append((prm, P.IC_NOISE, 700 + i, 0.01))
    ens = P.Ensemble([m[0] for m in specs], "mixed")
    assert ens.batches() == 2
    alone = []
    for (p, ic, seed, amp), m in zip(specs, ens.members):
        c0 = P.ic_fill(p, ic, seed, amp)
        m.init(c0)
        s = P.Sim(p, "mixed")
        s.init(c0)
        alone.append(s)
    ens.run(5)
    for s in alone:
        s.run(5)
    for i, (m, s) in enumerate(zip(ens.members, alone)):
        for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG", "W_PG"):
            assert np.array_equal(m.download(f), s.download(f)), (grids[i], i, f)
    print("[ensemble batches] 16x32 and 16x16 planes: every member equal", flush=True)


def test_ensemble_64_members_400_steps_equal_standalone():
    """The 64-member workload in full: members on the grids 64x32x32,
    128x48x48 and 192x64x64 with their drawn R, Pe and seeds, advanced 400
    steps as an ensemble in mixed precision (equal grids in batches), equal
    every member run alone for 400 steps in every bit."""
    specs = ensemble_members(64)
    t0 = time.time()
    ens = P.Ensemble([m[0] for m in specs], "mixed")
    assert ens.batches() == 3
    c0s = []
    for (p, ic, seed, amp), m in zip(specs, ens.members):
        c0 = P.ic_fill(p, ic, seed, amp)
        m.init(c0)
        c0s.append(c0)
    ens.run(400)
    print("[ensemble 64] 400 steps in %.1f s" % (time.time() - t0), flush=True)
    for i, ((p, *_), m, c0) in enumerate(zip(specs, ens.members, c0s)):
        s = P.Sim(p, "mixed")
        s.init(c0)
        s.run(400)
        for f in ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI"):
            assert np.array_equal(m.download(f), s.download(f)), ((p.nx, p.ny, p.nz), i, f)
        s.close()
    print("[ensemble 64] every member equal after 400 steps (%.1f s)" % (time.time() - t0), flush=True)
    ens.close()


def test_ensemble_batch_kernel_timing_advances_that_batch_only():
    """pllc_ensemble_batch_info reports each batch's members and grid;
    pllc_ensemble_time_kernel times one kernel class over steps of one
    batch's members alone: they advance exactly as standalone runs do, and
    the other batch's members keep their state."""
    grids = [(24, 16, 16)] * 3 + [(20, 16, 32)] * 2
    specs = []
    for i, (nx, ny, nz) in enumerate(grids):
        prm = P.make_params(nx, ny, nz, Vratio=2.0 + 0.3 * i, Pe=500.0 + 100.0 * i, Inf_thick=0.05)
        specs.append((prm, P.IC_NOISE, 900 + i, 0.01))
    ens = P.Ensemble([m[0] for m in specs], "mixed")
    assert ens.batches() == 2
    info = sorted(ens.batch_info(b) for b in range(2))
    assert info == [(2, (20, 16, 32)), (3, (24, 16, 16))], info
    c0s = []
    for (p, ic, seed, amp), m in zip(specs, ens.members):
        c0 = P.ic_fill(p, ic, seed, amp)
        m.init(c0)
        c0s.append(c0)
    b = [ens.batch_info(k)[1] for k in range(2)].index((24, 16, 16))
    avg, n = ens.time_kernel(b, "dct", 2)
    assert n == 12 and avg > 0.0, (avg, n)
    for i, ((p, *_), m, c0) in enumerate(zip(specs, ens.members, c0s)):
        s = P.Sim(p, "mixed")
        s.init(c0)
        if grids[i] == (24, 16, 16):
            s.run(2)
        for f in ("C_OLD", "U_OLD", "PHI"):
            assert np.array_equal(m.download(f), s.download(f)), (grids[i], i, f)
        s.close()
    ens.close()
