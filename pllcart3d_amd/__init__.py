"""pllcart3d_amd -- the pllcart3d time step (RK3-CN, HJ-WENO5 advection,
variable-viscosity ADI/CN momentum, DCT Poisson projection) on one MI355X.

The product is the C-ABI library libpllcart3d_amd.so (include/pllcart3d.h)
and the C driver run3_amd; this package is the Python handle over the same
library for tests and benchmarks.
"""
import ctypes as C

import numpy as np

from ._lib import (FIELDS, FIELD_ID, IC_ERF, IC_MULTIMODE, IC_NOISE, OP_ID, OPS, PRECISION, Params,
                   last_error, load)
from .configs import CONFIGS, make_params

__all__ = ["Sim", "Ensemble", "Params", "make_params", "CONFIGS", "ic_fill", "FIELDS", "FIELD_ID", "OPS",
           "OP_ID", "IC_ERF", "IC_NOISE", "IC_MULTIMODE", "load", "last_error"]


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def ic_fill(params, recipe=IC_ERF, seed=0, amp=0.0):
    """Initial concentration on every cell (nx+6, ny+2, nz+2), host only."""
    L = load()
    out = np.zeros((params.nx + 6, params.ny + 2, params.nz + 2))
    if L.pllc_ic_fill(C.byref(params), recipe, seed, amp, _dp(out)) != 0:
        raise RuntimeError(last_error())
    return out


class Sim:
    """One simulation owned by the library (device state on the GPU).

    precision: "plain" (FP64 throughout) or "mixed" (FP32 fields and
    stencils, FP64 reductions and (0,0) Poisson pencil).  Fields cross the
    boundary as float64 either way.
    """

    def __init__(self, params, precision="plain"):
        self.L = load()
        self.params = params
        self.precision = precision
        self.h = self.L.pllc_sim_create2(C.byref(params), PRECISION[precision])
        if not self.h:
            raise RuntimeError("pllc_sim_create: " + last_error())
        d = (C.c_int * 3)()
        self.n = self.L.pllc_sim_dims(self.h, d)
        self.shape = (d[0], d[1], d[2])

    def close(self):
        if self.h:
            self.L.pllc_sim_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != 0:
            raise RuntimeError(what + ": " + last_error())

    def init(self, c_full=None):
        p = None
        if c_full is not None:
            c_full = np.ascontiguousarray(c_full, dtype=np.float64)
            assert c_full.shape == self.shape
            p = _dp(c_full)
        self._chk(self.L.pllc_sim_init(self.h, p), "pllc_sim_init")

    def upload(self, field, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float64)
        assert arr.shape == self.shape
        self._chk(self.L.pllc_sim_upload(self.h, FIELD_ID[field], _dp(arr)), "pllc_sim_upload")

    def download(self, field):
        out = np.empty(self.shape)
        self._chk(self.L.pllc_sim_download(self.h, FIELD_ID[field], _dp(out)), "pllc_sim_download")
        return out

    def op(self, name, dt=0.0, rk=0, vbulk=0.0):
        v = self.L.pllc_sim_op(self.h, OP_ID[name], dt, rk, vbulk)
        if v != v:
            raise RuntimeError("pllc_sim_op(%s): %s" % (name, last_error()))
        return v

    def run(self, nsteps):
        t = C.c_double(0.0)
        dt = C.c_double(0.0)
        self._chk(self.L.pllc_sim_run(self.h, nsteps, C.byref(t), C.byref(dt)), "pllc_sim_run")
        return t.value, dt.value

    @property
    def timestep(self):
        return self.L.pllc_sim_timestep(self.h)

    def autotune(self, nsteps=3):
        """Pick the sweeps' lines per block by timing steps (advances the
        state); later simulations of this grid and precision use the result."""
        out = (C.c_int * 3)()
        self._chk(self.L.pllc_sim_autotune(self.h, nsteps, out), "pllc_sim_autotune")
        return tuple(out)

    def synchronize(self):
        self._chk(self.L.pllc_sim_synchronize(self.h), "pllc_sim_synchronize")

    def shift_conc(self, ampli, lamb, subharm=0):
        """The wave perturbation (shift_conc) on c_old."""
        self._chk(self.L.pllc_sim_shift_conc(self.h, ampli, lamb, subharm), "pllc_sim_shift_conc")

    @property
    def cm(self):
        """Slope parameter of the non-monotonic viscosity law (Nmvp = 1)."""
        return self.L.pllc_sim_slope_parameter(self.h)

    def time_steps(self, warmup, nsteps):
        ms = C.c_double(0.0)
        self._chk(self.L.pllc_sim_time_steps(self.h, warmup, nsteps, C.byref(ms)), "pllc_sim_time_steps")
        return ms.value

    def time_kernel(self, name, nsteps):
        avg = C.c_double(0.0)
        n = C.c_int(0)
        self._chk(self.L.pllc_sim_time_kernel(self.h, name.encode(), nsteps, C.byref(avg), C.byref(n)),
                  "pllc_sim_time_kernel")
        return avg.value, n.value


class _Member(Sim):
    """A member of an Ensemble: the ensemble owns (and frees) the handle."""

    def __init__(self, ens, h, params, precision):
        self.L = ens.L
        self.params = params
        self.precision = precision
        self.h = h
        self._ens = ens
        d = (C.c_int * 3)()
        self.n = self.L.pllc_sim_dims(self.h, d)
        self.shape = (d[0], d[1], d[2])

    def close(self):
        self.h = None


class Ensemble:
    """Independent members (own grid, Vratio, Pe, initial state) advanced
    together.  In mixed precision the members of equal grids form batches,
    each advanced by one launch per kernel of the step; the other members
    replay their own captured two-step graphs on their own streams as the
    branches of one graph, so their kernels overlap."""

    def __init__(self, params_list, precision="plain"):
        self.L = load()
        arr = (Params * len(params_list))(*params_list)
        self.h = self.L.pllc_ensemble_create(arr, len(params_list), PRECISION[precision])
        if not self.h:
            raise RuntimeError("pllc_ensemble_create: " + last_error())
        self.members = [_Member(self, self.L.pllc_ensemble_member(self.h, i), p, precision)
                        for i, p in enumerate(params_list)]

    def batches(self):
        """The number of batches of members stepped together."""
        return self.L.pllc_ensemble_batches(self.h)

    def batch_info(self, b):
        """(members, (nx, ny, nz)) of batch b."""
        n = C.c_int(0)
        d = (C.c_int * 3)()
        if self.L.pllc_ensemble_batch_info(self.h, b, C.byref(n), d) != 0:
            raise RuntimeError("pllc_ensemble_batch_info: " + last_error())
        return n.value, (d[0], d[1], d[2])

    def time_kernel(self, b, name, nsteps):
        """Average ms and count of the launches of kernel class `name` over
        nsteps steps of batch b's members alone (they advance; the other
        members do not)."""
        avg = C.c_double(0.0)
        n = C.c_int(0)
        if self.L.pllc_ensemble_time_kernel(self.h, b, name.encode(), nsteps, C.byref(avg), C.byref(n)) != 0:
            raise RuntimeError("pllc_ensemble_time_kernel: " + last_error())
        return avg.value, n.value

    def run(self, nsteps):
        if self.L.pllc_ensemble_run(self.h, nsteps) != 0:
            raise RuntimeError("pllc_ensemble_run: " + last_error())

    def time_steps(self, warmup, nsteps):
        ms = C.c_double(0.0)
        if self.L.pllc_ensemble_time_steps(self.h, warmup, nsteps, C.byref(ms)) != 0:
            raise RuntimeError("pllc_ensemble_time_steps: " + last_error())
        return ms.value

    def close(self):
        if self.h:
            for m in self.members:
                m.h = None
            self.L.pllc_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
