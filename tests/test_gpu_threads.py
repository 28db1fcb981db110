"""Two handles driven from two host threads at once (ctypes releases the
GIL for the library's calls) give the same fields, bit for bit, as the same
two runs one after the other."""
import threading

import numpy as np
import pytest

import pllcart3d_amd as P

pytestmark = pytest.mark.gpu

FIELDS = ("C_OLD", "U_OLD", "V_OLD", "W_OLD", "PHI", "U_PG")


def _run(shape, precision, steps, seed, out, key):
    prm = P.make_params(*shape, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    sim = P.Sim(prm, precision)
    sim.init(P.ic_fill(prm, P.IC_NOISE, seed, 0.05))
    sim.run(steps)
    out[key] = {f: sim.download(f) for f in FIELDS}
    sim.close()


def test_handles_on_two_threads_equal_sequential_runs():
    jobs = [((32, 16, 12), "plain", 4, 3), ((24, 20, 16), "mixed", 4, 5)]
    seq = {}
    for n, (shape, prec, steps, seed) in enumerate(jobs):
        _run(shape, prec, steps, seed, seq, n)
    par = {}
    errors = []

    def work(n):
        try:
            _run(*jobs[n], par, n)
        except Exception as e:   # reported below, in the test's thread
            errors.append(e)

    ts = [threading.Thread(target=work, args=(n,)) for n in range(len(jobs))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not errors, errors
    print("[threads] %d handles on %d threads" % (len(jobs), len(ts)), flush=True)
    for n in range(len(jobs)):
        for f in FIELDS:
            assert np.array_equal(seq[n][f], par[n][f]), (n, f)
