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


def find_max_record(u, v, w, cn, co, cx):
    """find_max (pllcart3d.c:3539-3603) restated on whole arrays: per field
    the value of largest magnitude, the last index in i-j-k order winning a
    tie (the `<=` comparison from a running maximum 0); the first cell in
    i-j-k order with a NaN in c_old ('a'), c_new ('b') or N ('c')."""
    out = []
    for f in (u, v, w, cn):
        flat = f.ravel()
        with np.errstate(invalid="ignore"):
            m = np.nanmax(np.abs(flat))   # a NaN never passes the comparison
            idx = int(np.nonzero(np.abs(flat) == m)[0][-1])
        out.append((np.unravel_index(idx, f.shape), flat[idx]))
    bad = np.isnan(co) | np.isnan(cn) | np.isnan(cx)
    nan = None
    if bad.any():
        p = int(np.nonzero(bad.ravel())[0][0])
        tag = "a" if np.isnan(co.ravel()[p]) else ("b" if np.isnan(cn.ravel()[p]) else "c")
        nan = (tag, np.unravel_index(p, co.shape))
    return out, nan


def _max_blocks(text):
    """The (i, j, k, value) rows of every `max` block of an Outscreen."""
    lines = text.splitlines()
    blocks = []
    for n, ln in enumerate(lines):
        if ln == "max":
            rows = []
            for r in lines[n + 1:n + 5]:
                i, j, k, val = r.split()
                rows.append(((int(i), int(j), int(k)), float(val)))
            blocks.append(rows)
    return blocks


MIDFIELD_RECORDS = ["C_OLD", "C_NM1", "U_OLD", "U_NEW", "U_NM1", "U_PG", "V_OLD", "V_NEW", "V_NM1", "V_PG",
                    "W_OLD", "W_NEW", "W_NM1", "W_PG"]


def write_restart(d, orc, tstep, tval):
    """middata and midfield000.bin of the oracle's state in the reference's
    formats (midfield, pllcart3d.c:3121-3178: per (i, j) the 14 rows of
    nz_max values in MIDFIELD_RECORDS order)."""
    with open(d / "middata", "w") as f:
        f.write("%d\n%d\n%f\n%f\n%f\n%f\n" % (tstep, 1, tval, 100.0, 100.0, 100.0))
    rec = np.stack([orc.field(n) for n in MIDFIELD_RECORDS], axis=2)   # [i][j][14][k]
    rec.tofile(d / "midfield000.bin")


def test_driver_diagnostic_values(tmp_path, oracle_mod):
    """find_max's printed indices and values after every step, and
    find_wmax's records (written with Initial = 0 and Wavpert = 1), from a
    restart of a z-varying state (a 16x16-mode front after 2 oracle steps),
    against the oracle restarted the same way: find_max restated on the
    oracle's arrays, find_wmax = max(0, largest interior w) as float32."""
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    rst = oracle_mod.Oracle(prm)
    rst.init(P.ic_fill(prm, P.IC_MULTIMODE, 4, 0.05))
    t, _ = rst.run(2)
    t = float("%f" % t)   # middata keeps tval to 6 decimals
    write_restart(tmp_path, rst, 2, t)
    nsteps = 3
    _run(tmp_path, nsteps, {"Initial": 0, "Wavpert": 1, "Niter_print": 1, "Out_mid": 100.0,
                            "Out_inter": 100.0})
    blocks = _max_blocks((tmp_path / "Outscreen").read_text())
    wrec = np.fromfile(tmp_path / "wmax.bin", dtype=np.float32).reshape(-1, 2)
    assert len(blocks) == nsteps and len(wrec) == nsteps, (len(blocks), len(wrec))
    # the restarted run starts with every array midfield does not save zero
    for f in ("C_NEW", "C_N", "U_N", "V_N", "W_N", "PHI", "DIV"):
        rst.field(f)[...] = 0.0
    for s in range(nsteps):
        t, _ = rst.run(1, t)
        rec, nan = find_max_record(*(rst.field(f) for f in ("U_NEW", "V_NEW", "W_NEW", "C_NEW", "C_OLD", "C_N")))
        assert nan is None
        fields = [rst.field(f) for f in ("U_NEW", "V_NEW", "W_NEW", "C_NEW")]
        for (ix, val), (gix, gval), f in zip(rec, blocks[s], fields):
            # the index is the oracle's unless the two differ only by a near
            # tie (values within 1e-12 of each other, which rounding may
            # order either way)
            if tuple(int(x) for x in ix) != gix:
                assert abs(abs(f[gix]) - abs(val)) <= 1e-12 * abs(val), (s, ix, gix, f[gix], val)
            assert abs(gval - val) <= 5e-7 + 1e-12 * abs(val), (s, gval, val)
        w = rst.field("W_NEW")[3:-3, 1:-1, 1:-1]
        wm = max(0.0, float(w.max()))
        print("[diag] step %d max rows %s wmax %.6e (file %.6e) t %.6f" % (s + 1, blocks[s], wm, wrec[s, 0], t),
              flush=True)
        assert wm > 1e-6
        assert abs(float(wrec[s, 0]) - wm) <= 1.2e-7 * wm + 1e-12, (s, wrec[s, 0], wm)
        assert abs(float(wrec[s, 1]) - t) <= 1.2e-7 * t, (s, wrec[s, 1], t)


def test_driver_find_max_nan_guard(tmp_path, oracle_mod):
    """A restart file with a NaN in one cell of c_old: after the first step
    the NaN has spread along the cell's line solves; find_max names the
    first NaN cell in i-j-k order ('a': c_old) as the oracle's state does,
    and the driver exits with status 1 (pllcart3d.c:3568-3588)."""
    a, b = tmp_path / "a", tmp_path / "b"
    a.mkdir()
    b.mkdir()
    _run(a, 2, {"Out_mid": 0.003})
    (b / "middata").write_bytes((a / "middata").read_bytes())
    # midfield (pllcart3d.c:3121-3178) writes, for every (i, j), 14 rows of
    # nz_max values, c_old first
    raw = np.fromfile(a / "midfield000.bin").reshape(30, 14, 14, 12).copy()
    raw[10, 5, 0, 4] = np.nan
    raw.tofile(b / "midfield000.bin")
    vals = dict(PARAMS)
    vals.update({"Initial": 0, "Niter_print": 1, "Out_mid": 100.0})
    with open(b / "param.input", "w") as f:
        for name, _ in PARAMS:
            f.write("//%s\n%r\n" % (name, vals[name]))
    exe = os.path.join(os.path.dirname(P.__file__), "run3_amd")
    r = subprocess.run([exe, "2"], cwd=b, capture_output=True, text=True, timeout=120)
    text = (b / "Outscreen").read_text()
    line = [ln for ln in text.splitlines() if ln.startswith("nan in")]
    print("[nan guard] rc %d: %s" % (r.returncode, line), flush=True)
    assert r.returncode == 1 and len(line) == 1
    tag, i, j, k = line[0].split()[2:6]
    assert tag == "a"
    # the oracle from the same restart state, one step
    prm = P.make_params(24, 12, 10, Vratio=3.0, Pe=300.0, Inf_thick=0.1)
    rst = oracle_mod.Oracle(prm)
    rst.init()
    t, _ = rst.run(2)
    t = float("%f" % t)   # middata keeps tval to 6 decimals
    for f in ("C_NEW", "C_N", "U_N", "V_N", "W_N", "PHI", "DIV"):
        rst.field(f)[...] = 0.0
    rst.field("C_OLD")[10, 5, 4] = np.nan
    with np.errstate(invalid="ignore"):
        rst.run(1, t)
        _, nan = find_max_record(*(rst.field(f) for f in ("U_NEW", "V_NEW", "W_NEW", "C_NEW", "C_OLD", "C_N")))
    assert nan is not None and nan[0] == "a"
    assert tuple(int(x) for x in nan[1]) == (int(i), int(j), int(k)), (nan, line)
