"""The C-ABI library loads without a GPU and exports every function that
include/pllcart3d.h declares (no compute calls here)."""
import ctypes as C
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pllcart3d.h")


def declared_functions():
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", " ", txt, flags=re.S)
    txt = re.sub(r"//[^\n]*", " ", txt)
    txt = re.sub(r"#[^\n]*", " ", txt)
    names = re.findall(r"\b([A-Za-z_]\w*)\s*\([^;{}()]*(?:\([^()]*\)[^;{}()]*)*\)\s*;", txt)
    return sorted(set(n for n in names if n not in ("if", "while", "return", "sizeof")))


def test_header_declares_the_reference_entry_points():
    names = set(declared_functions())
    # global_function_decl.h:2-45 (time-step and set-up entry points), pll_tridiag.c, fftw_3d.h
    ref = ["input", "grid_setup", "simulation_setup", "mesh_generation", "initial_condition",
           "concentration", "calculate_viscosity", "xmom", "ymom", "zmom", "A_conc", "A_umom", "A_vmom",
           "A_wmom", "divergence", "phisolver_dct", "velo_correction", "update_veloghost_points",
           "update_ghost_points", "update_velo_boundary", "update_phi_ghosts", "pressure_gradient",
           "find_cfl", "bulk_velo", "update_old_conc", "update_old_velo", "update_old_visc",
           "calculate_weno", "tridiag_1rhs", "pll_tridiag", "plan_dct_3d", "fdct_3d", "idct_3d",
           "alloc3d_dble", "alloc_velo", "alloc_conc", "alloc_visc", "alloc_tds", "free_tds", "free3d_dble"]
    missing = [n for n in ref if n not in names]
    assert not missing, missing


def test_library_exports_every_declared_symbol():
    from pllcart3d_amd import _lib
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    for g in ("Nx_cell_global", "Re", "Pe", "Vratio", "RK1", "RK2", "Rank", "Size", "Timestep"):
        assert hasattr(L, g), g
    rk1 = (C.c_double * 3).in_dll(L, "RK1")
    assert list(rk1) == [8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0]
    assert C.c_int.in_dll(L, "Size").value == 1


def test_last_error_is_a_string_without_gpu():
    import pllcart3d_amd as P
    assert isinstance(P.last_error(), str)
    # a host-only call with bad arguments reports through pllc_last_error
    prm = P.make_params(0, 4, 4)
    out = np.zeros(10)
    assert P.load().pllc_ic_fill(C.byref(prm), 0, 0, 0.0, out.ctypes.data_as(C.POINTER(C.c_double))) != 0
    assert "pllc_ic_fill" in P.last_error()
