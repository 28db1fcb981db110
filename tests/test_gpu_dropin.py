"""The drop-in surface end to end: run3_amd (the reference's main call
sequence through the reference-named entry points of libpllcart3d_amd.so,
param.input read by input()) vs the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

import pllcart3d_amd as P
from tests.conftest import rel_l2

pytestmark = pytest.mark.gpu

# param.input values in the reference's order (param.input:1-68), Nmvp = 0
PARAMS = [("Nx_cell_global", 24), ("Ny_cell_global", 12), ("Nz_cell_global", 10), ("Nx_ghost", 3),
          ("Ny_ghost", 1), ("Nz_ghost", 1), ("Initial", 1), ("Ugrid", 1), ("Niter_print", 1000),
          ("Out_mid", 100.0), ("X1", 0.0), ("X2", 4.0), ("Y1", -0.5), ("Y2", 0.5), ("Z1", -0.5), ("Z2", 0.5),
          ("Timemax", 100.0), ("Timestep", 2e-3), ("Re", 1.0), ("Pe", 300.0), ("Vratio", 3.0), ("Fx", 0.0),
          ("Fy", 0.0), ("Inf_pos", 1.0), ("Inf_thick", 0.1), ("CFLmax", 0.3), ("Out_inter", 100.0),
          ("Wavpert", 0), ("Ampli", 0.1), ("Lamb", 3.0), ("Subharm", 0), ("Nmvp", 0), ("Mum", 1.5),
          ("Lambda", 0.0)]


DUMP = ["C_NEW", "U_NEW", "V_NEW", "W_NEW", "PHI", "U_PG", "V_PG", "W_PG"]


def _check_dump(dump, orc, tag, pg_tol=1e-11):
    """c, velocity (u, v, w as one vector field), phi and the pressure
    gradient (one vector field) <= 1e-11 relative L2 (the pressure gradient
    to pg_tol).  With the erf front the flow is two-dimensional: w and w_pg
    are rounding noise around 0 in both solvers, so they are held to the
    velocity's scale instead."""
    groups = {"c": [0], "velocity": [1, 2, 3], "phi": [4], "pg": [5, 6, 7]}
    for name, qs in groups.items():
        a = np.concatenate([dump[q].ravel() for q in qs])
        b = np.concatenate([orc.field(DUMP[q]).ravel() for q in qs])
        e = rel_l2(a, b)
        print("[%s] %s rel L2 %.2e" % (tag, name, e), flush=True)
        assert e <= (pg_tol if name == "pg" else 1e-11), (name, e)
    # w against the velocity's scale, w_pg against the pressure gradient's
    # (and its tolerance)
    uscale = np.abs(orc.field("U_NEW")).max()
    pscale = np.abs(orc.field("U_PG")).max()
    assert np.abs(dump[3] - orc.field(DUMP[3])).max() <= 1e-11 * uscale, DUMP[3]
    assert np.abs(dump[7] - orc.field(DUMP[7])).max() <= pg_tol * pscale, DUMP[7]


def write_param_input(path, params=PARAMS):
    with open(path, "w") as f:
        for name, val in params:
            f.write("//%s\n%r\n" % (name, val))


def test_driver_matches_oracle(tmp_path, oracle_mod):
    write_param_input(tmp_path / "param.input")
    nsteps = 4
    exe = os.path.join(os.path.dirname(P.__file__), "run3_amd")
    print("[dropin] run3_amd %d steps 24x12x10" % nsteps, flush=True)
    r = subprocess.run([exe, str(nsteps), "dump.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "timetaken" in (tmp_path / "Outscreen").read_text()
    shape = (30, 14, 12)
    dump = np.fromfile(tmp_path / "dump.bin").reshape(8, *shape)
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    orc = oracle_mod.Oracle(prm)
    orc.init()
    orc.run(nsteps)
    _check_dump(dump, orc, "dropin")


def test_driver_128_planes_matches_oracle(tmp_path, oracle_mod):
    """8x128x128 through the drop-in: phisolver_dct's (y, z) transforms run
    as the fast plane transform.  The pressure gradient is held to ten times
    the oracle's own rounding floor for this case (tests/rounding_floor.py,
    measured here): its 1/(Re dy^2)-scaled second differences of grad(phi)
    multiply phi's rounding."""
    dims = {"Nx_cell_global": 8, "Ny_cell_global": 128, "Nz_cell_global": 128}
    write_param_input(tmp_path / "param.input", [(n, dims.get(n, v)) for n, v in PARAMS])
    nsteps = 2
    exe = os.path.join(os.path.dirname(P.__file__), "run3_amd")
    print("[dropin] run3_amd %d steps 8x128x128" % nsteps, flush=True)
    r = subprocess.run([exe, str(nsteps), "dump.bin"], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    dump = np.fromfile(tmp_path / "dump.bin").reshape(8, 14, 130, 130)
    prm = P.make_params(8, 128, 128, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    orc = oracle_mod.Oracle(prm)
    orc.init()
    orc.run(nsteps)
    from tests import rounding_floor as RF
    bars, floors = RF.live_bars(oracle_mod, prm, P.ic_fill(prm, P.IC_ERF), nsteps)
    print("[dropin 8x128x128] pg floor %.1e bar %.1e" % (floors["pg"], bars["pg"]), flush=True)
    _check_dump(dump, orc, "dropin 8x128x128", pg_tol=bars["pg"])


def test_driver_nmvp_matches_oracle(tmp_path, oracle_mod):
    """run3_amd with Nmvp = 1 (the reference's shipped param.input setting):
    slope_parameter + calculate_viscosity_nmvp through the drop-in."""
    _run(tmp_path, 3, {"Nmvp": 1, "Mum": 1.5, "Lambda": 0.0}, "dump.bin")
    dump = np.fromfile(tmp_path / "dump.bin").reshape(8, 30, 14, 12)
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1, Nmvp=1, Mum=1.5, Lambda=0.0)
    orc = oracle_mod.Oracle(prm)
    orc.init()
    orc.run(3)
    assert ("cm = %f" % orc.cm) in (tmp_path / "Outscreen").read_text()
    _check_dump(dump, orc, "dropin nmvp")


def _run(tmp, nsteps, overrides, dump=None):
    vals = dict(PARAMS)
    vals.update(overrides)
    with open(tmp / "param.input", "w") as f:
        for name, _ in PARAMS:
            f.write("//%s\n%r\n" % (name, vals[name]))
    exe = os.path.join(os.path.dirname(P.__file__), "run3_amd")
    args = [exe, str(nsteps)] + ([dump] if dump else [])
    r = subprocess.run(args, cwd=tmp, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    return r


def test_driver_outputs_and_restart(tmp_path, oracle_mod):
    """caf files, conc_average, find_max and the midfield/intermediate
    restart.  midfield (pllcart3d.c:3121-3228) saves c_old, c's N_nm1 and
    u/v/w old, new, N_nm1 and pg; a restarted run starts with every other
    array zero (c_new among them, whose never-written points - outlet and
    ghost rows - then reach c_old and the viscosity), so 3 steps + restart +
    2 steps is checked against the oracle restarted the same way, not
    against 5 straight steps."""
    shape = (30, 14, 12)
    a, b, c = tmp_path / "a", tmp_path / "b", tmp_path / "c"
    for d in (a, b, c):
        d.mkdir()
    io = {"Out_inter": 0.0035, "Out_mid": 0.005, "Niter_print": 1}
    print("[dropin] outputs + restart", flush=True)
    _run(a, 3, io)
    for f in ("middata", "midfield000.bin"):
        (b / f).write_bytes((a / f).read_bytes())
    _run(b, 2, dict(io, Initial=0), "dump.bin")
    _run(c, 5, io, "dump.bin")
    db = np.fromfile(b / "dump.bin").reshape(8, *shape)
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    rst = oracle_mod.Oracle(prm)
    rst.init()
    rst.run(3)
    for f in ("C_NEW", "C_N", "U_N", "V_N", "W_N", "PHI", "DIV"):
        rst.field(f)[...] = 0.0
    rst.run(2)
    _check_dump(db, rst, "restart")

    # caf0002 holds the state after step 4 (writes at t=0, 0.004, 0.008)
    orc = oracle_mod.Oracle(prm)
    orc.init()
    orc.run(1)
    cyz = orc.field("C_NEW")[3:27, 1:13, 1:11].mean(axis=(1, 2))
    rec = np.fromfile(c / "cavgyz.bin", dtype=np.float32)
    assert abs(rec[0] - 0.002) < 1e-6
    assert np.allclose(rec[1:25], cyz, rtol=1e-6, atol=1e-7)
    orc.run(3)
    u, v, w, cc = (orc.field(n) for n in ("U_NEW", "V_NEW", "W_NEW", "C_NEW"))
    ref = np.stack([0.5 * (u[3:27, 1:13, 1:11] + u[4:28, 1:13, 1:11]),
                    0.5 * (v[3:27, 1:13, 1:11] + v[3:27, 2:14, 1:11]),
                    0.5 * (w[3:27, 1:13, 1:11] + w[3:27, 1:13, 2:12]),
                    cc[3:27, 1:13, 1:11]], axis=-1).astype(np.float32)
    caf = np.fromfile(c / "caf0002R000.bin", dtype=np.float32).reshape(24, 12, 10, 4)
    assert np.allclose(caf, ref, rtol=1e-6, atol=1e-7)
    assert (c / "caf0000R000.bin").exists() and (c / "caf0001R000.bin").exists()
    out = (c / "Outscreen").read_text()
    assert out.count("max\n") == 5 and "timetaken" in out



def test_spline_entry_point_matches_oracle(oracle_mod):
    """spline() of the drop-in on host arrays (one line, evaluated on the
    device) vs the oracle's restatement: bitwise."""
    import ctypes as C
    L = C.CDLL(os.path.join(os.path.dirname(P.__file__), "libpllcart3d_amd.so"))
    O = oracle_mod.load()
    dpt = C.POINTER(C.c_double)

    class Grid(C.Structure):   # structures_3d.h Grid_data (include/pllcart3d.h)
        _fields_ = [(n, C.c_int) for n in ("istart", "iend", "jstart", "jend", "kstart", "kend",
                                           "nx_cell", "ny_cell", "nz_cell", "nx_max", "ny_max", "nz_max",
                                           "nmax_xyz", "nx_global_offset")] + \
                   [(n, C.c_double) for n in ("dx", "dy", "dz")] + [(n, dpt) for n in ("X", "Y", "Z", "Xc", "Yc", "Zc")]

    N, dx = 36, 0.125
    g = Grid()
    g.nx_max, g.dx = N, dx
    x = (np.arange(N) - 2.5) * dx
    rng = np.random.default_rng(11)
    y = np.cos(1.7 * x) + 0.05 * rng.normal(size=N)
    for shift in (0.04, -0.09, 0.0):
        xx = x - shift
        got = np.zeros(N)
        L.spline(oracle_mod.dp(x), oracle_mod.dp(xx), oracle_mod.dp(y), oracle_mod.dp(got), C.byref(g))
        ref = np.zeros(N)
        O.orc_spline(N, dx, oracle_mod.dp(x), oracle_mod.dp(xx), oracle_mod.dp(y), oracle_mod.dp(ref))
        assert np.array_equal(got, ref), (shift, np.abs(got - ref).max())
