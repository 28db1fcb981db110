"""ctypes wrapper of the FP64 CPU oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the CPU
baseline of bench.py, always to check results, never as the product.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
REF_TRIDIAG = os.path.join(HERE, "_ref", "ref_tridiag")

FIELDS = ["C_OLD", "C_NEW", "C_N", "C_NM1",
          "U_OLD", "U_NEW", "U_PG", "U_N", "U_NM1",
          "V_OLD", "V_NEW", "V_PG", "V_N", "V_NM1",
          "W_OLD", "W_NEW", "W_PG", "W_N", "W_NM1",
          "MUO_CELL", "MUO_XY", "MUO_XZ", "MUO_YZ",
          "MUN_CELL", "MUN_XY", "MUN_XZ", "MUN_YZ",
          "PHI", "DIV"]
FIELD_ID = {n: i for i, n in enumerate(FIELDS)}
OPS = ["CONCENTRATION", "VISCOSITY", "XMOM", "YMOM", "ZMOM", "VELO_BOUNDARY", "DIVERGENCE",
       "PHISOLVER", "PHI_GHOSTS", "VELO_CORRECTION", "VELOGHOST", "PRESSURE_GRADIENT",
       "UPDATE_OLD", "BULK_VELO", "FIND_CFL"]
OP_ID = {n: i for i, n in enumerate(OPS)}

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        dp = C.POINTER(C.c_double)
        L.orc_create.restype = C.c_void_p
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_dims.restype = C.c_size_t
        L.orc_dims.argtypes = [C.c_void_p, C.POINTER(C.c_int)]
        L.orc_field.restype = dp
        L.orc_field.argtypes = [C.c_void_p, C.c_int]
        L.orc_timestep.restype = C.c_double
        L.orc_timestep.argtypes = [C.c_void_p]
        L.orc_cm.restype = C.c_double
        L.orc_cm.argtypes = [C.c_void_p]
        L.orc_slope_parameter.restype = C.c_double
        L.orc_slope_parameter.argtypes = [dp, C.c_int, C.c_double, C.c_double, C.c_double]
        L.orc_init.argtypes = [C.c_void_p, dp]
        L.orc_rebuild_A.argtypes = [C.c_void_p]
        L.orc_set_cm.argtypes = [C.c_void_p, C.c_double]
        L.orc_op.restype = C.c_double
        L.orc_op.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, C.c_double]
        L.orc_run.restype = C.c_double
        L.orc_run.argtypes = [C.c_void_p, C.c_int, dp]
        L.orc_set_threads.argtypes = [C.c_int]
        L.orc_shift_conc.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_int]
        L.orc_spline.argtypes = [C.c_int, C.c_double, dp, dp, dp, dp]
        L.orc_weno.restype = C.c_double
        L.orc_weno.argtypes = [dp, C.c_int, C.c_int]
        L.orc_tridiag_1rhs.argtypes = [C.c_int, dp, dp, dp, dp]
        L.orc_pll_tridiag.argtypes = [C.c_int, C.c_int, dp, dp, dp, dp]
        L.orc_redft10.argtypes = [C.c_int, dp, dp]
        L.orc_redft01.argtypes = [C.c_int, dp, dp]
        L.orc_fdct2.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
        L.orc_idct2.argtypes = [C.c_int, C.c_int, C.c_int, dp, dp]
        _lib = L
    return _lib


def dp(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Oracle:
    """One oracle simulation; params is a pllc_params-compatible Structure."""

    def __init__(self, params):
        self.L = load()
        self.nmvp = params.Nmvp
        self.h = self.L.orc_create(C.byref(params))
        d = (C.c_int * 3)()
        self.n = self.L.orc_dims(self.h, d)
        self.shape = (d[0], d[1], d[2])

    def __del__(self):
        try:
            if self.h:
                self.L.orc_destroy(self.h)
        except Exception:
            pass

    def field(self, name):
        """Live view of an oracle field (writes go into the oracle)."""
        p = self.L.orc_field(self.h, FIELD_ID[name])
        return np.ctypeslib.as_array(p, shape=self.shape)

    def init(self, c_full=None):
        self.L.orc_init(self.h, None if c_full is None else dp(np.ascontiguousarray(c_full)))
        # Nmvp = 1: the reference exits when the slope parameter's bisection
        # does not converge in 500 iterations (pllcart3d.c:812-817); the C
        # restatement returns cm = -1 then
        if self.nmvp == 1 and not self.L.orc_cm(self.h) > 0.0:
            raise RuntimeError("oracle init: slope parameter did not converge in 500 iterations")

    def rebuild_A(self):
        """Coefficient tables of u, v, w from MUO_* (after writing fields)."""
        self.L.orc_rebuild_A(self.h)

    def set_cm(self, cm):
        """Slope parameter of the Nmvp = 1 law for a state written directly."""
        self.L.orc_set_cm(self.h, cm)

    def op(self, name, dt=0.0, rk=0, vbulk=0.0):
        return self.L.orc_op(self.h, OP_ID[name], dt, rk, vbulk)

    def run(self, nsteps, tval=0.0):
        t = C.c_double(tval)
        dt = self.L.orc_run(self.h, nsteps, C.byref(t))
        return t.value, dt

    def shift_conc(self, ampli, lamb, subharm=0):
        self.L.orc_shift_conc(self.h, ampli, lamb, subharm)

    @property
    def timestep(self):
        return self.L.orc_timestep(self.h)

    @property
    def cm(self):
        return self.L.orc_cm(self.h)


def set_threads(n):
    load().orc_set_threads(n)
