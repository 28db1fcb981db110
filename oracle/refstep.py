"""Runs the reference's own time-step functions (oracle/_ref/ref_step).

TEST INFRASTRUCTURE ONLY: used by tests/ and tests/golden/ to pin the
oracle against the reference itself.  ref_step is built from pllcart3d.c
where it lies (oracle/Makefile, target `ref`) and exists only where the
reference tree does; callers skip when it is absent.

ref_step runs every function of pllcart3d.c except main and phisolver_dct
(both need FFTW); a request names a list of ops, run in order on one state.
"""
import os
import struct
import subprocess
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF_STEP = os.path.join(HERE, "_ref", "ref_step")

# the 29 state fields, in the oracle's field-id order
FIELDS = ["C_OLD", "C_NEW", "C_N", "C_NM1",
          "U_OLD", "U_NEW", "U_PG", "U_N", "U_NM1",
          "V_OLD", "V_NEW", "V_PG", "V_N", "V_NM1",
          "W_OLD", "W_NEW", "W_PG", "W_N", "W_NM1",
          "MUO_CELL", "MUO_XY", "MUO_XZ", "MUO_YZ",
          "MUN_CELL", "MUN_XY", "MUN_XZ", "MUN_YZ",
          "PHI", "DIV"]
OPS = {"CONCENTRATION": 0, "VISCOSITY": 1, "XMOM": 2, "YMOM": 3, "ZMOM": 4,
       "VELO_BOUNDARY": 5, "DIVERGENCE": 6, "PHI_GHOSTS": 8, "VELO_CORRECTION": 9,
       "VELOGHOST": 10, "PRESSURE_GRADIENT": 11, "UPDATE_OLD": 12, "BULK_VELO": 13,
       "FIND_CFL": 14, "INIT": 100, "A_ALL": 101, "SLOPE": 102, "SHIFT_CONC": 103,
       "CONC_AVERAGE": 104, "FIND_MAX": 105, "FIND_WMAX": 106, "WRITE_TO_FILE": 107,
       "MIDFIELD": 108, "INTERMEDIATE": 109, "TIMESTEP_CAP": 111, "A_NEW": 112,
       "INIT_KEEP_C": 113}
MAGIC = 0x50524546


def available():
    return os.path.exists(REF_STEP)


def shape_of(p):
    return (p.nx + 6, p.ny + 2, p.nz + 2)


def run(p, ops, state=None, cm=0.0, wave=(0, 0.0, 0.0, 0), workdir=None, files=()):
    """Run `ops` (list of (name, rk, dt, vbulk)) on a state.

    p: a params structure (fields nx..Lambda as pllc_params); state: dict
    field -> array of shape_of(p) (missing fields are zero); cm: the
    non-monotonic law's slope parameter for VISCOSITY with Nmvp = 1;
    wave: (Wavpert, Ampli, Lamb, Subharm) for SHIFT_CONC.  Returns (list
    of return values, dict field -> array, dict file name -> bytes for the
    names in `files` that the run wrote into its working directory).
    """
    shp = shape_of(p)
    wavpert, ampli, lamb, subharm = wave
    hdr = struct.pack("<8i", MAGIC, p.nx, p.ny, p.nz, p.Nmvp, len(ops), wavpert, subharm)
    dbl = struct.pack("<20d", p.X1, p.X2, p.Y1, p.Y2, p.Z1, p.Z2, p.Re, p.Pe, p.Vratio, p.Fx,
                      p.Fy, p.Inf_pos, p.Inf_thick, p.CFLmax, p.Timestep, p.Mum, p.Lambda,
                      ampli, lamb, cm)
    body = b"".join(struct.pack("<2i2d", OPS[name], rk, dt, vb) for name, rk, dt, vb in ops)
    if state is None:
        st = struct.pack("<i", 0)
    else:
        buf = np.zeros((len(FIELDS),) + shp)
        for i, f in enumerate(FIELDS):
            if f in state:
                a = np.asarray(state[f], dtype=np.float64)
                assert a.shape == shp, (f, a.shape, shp)
                buf[i] = a
        st = struct.pack("<i", 1) + buf.tobytes()
    with tempfile.TemporaryDirectory(dir=workdir) as d:
        req, rep = os.path.join(d, "req.bin"), os.path.join(d, "rep.bin")
        with open(req, "wb") as f:
            f.write(hdr + dbl + body + st)
        r = subprocess.run([REF_STEP, req, rep], cwd=d, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=600)
        if r.returncode != 0:
            raise RuntimeError("ref_step failed (%d): %s" % (r.returncode, r.stdout[-2000:]))
        raw = np.fromfile(rep, dtype=np.float64)
        out_files = {}
        for name in files:
            fp = os.path.join(d, name)
            if os.path.exists(fp):
                with open(fp, "rb") as f:
                    out_files[name] = f.read()
    rets = list(raw[:len(ops)])
    flds = raw[len(ops):].reshape((len(FIELDS),) + shp)
    return rets, {f: flds[i].copy() for i, f in enumerate(FIELDS)}, out_files


def parse_input(param_text):
    """Run the reference's input() on a param.input text; returns its values."""
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "param.input"), "w") as f:
            f.write(param_text)
        out = os.path.join(d, "vals.txt")
        r = subprocess.run([REF_STEP, "--input", out], cwd=d, stdout=subprocess.PIPE,
                           stderr=subprocess.STDOUT, text=True, timeout=60)
        if r.returncode != 0:
            raise RuntimeError("ref_step --input failed: " + r.stdout[-2000:])
        with open(out) as f:
            return f.read().split()
