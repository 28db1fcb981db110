"""ctypes bindings of libpllcart3d_amd.so (declarations in include/pllcart3d.h).

Loading the library needs no GPU; every compute call does.  The library is
the product: there is no fallback path, a missing build raises.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpllcart3d_amd.so")
DRIVER_PATH = os.path.join(HERE, "run3_amd")

PARAM_FIELDS = ["X1", "X2", "Y1", "Y2", "Z1", "Z2", "Re", "Pe", "Vratio", "Fx", "Fy",
                "Inf_pos", "Inf_thick", "CFLmax", "Timestep"]


class Params(C.Structure):
    """pllc_params: one run's set-up (param.input values of the time step)."""
    _fields_ = ([("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int)] + [(n, C.c_double) for n in PARAM_FIELDS] +
                [("Nmvp", C.c_int), ("Mum", C.c_double), ("Lambda", C.c_double)])


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

IC_ERF, IC_NOISE, IC_MULTIMODE = 0, 1, 2
PRECISION = {"plain": 0, "mixed": 1}

_lib = None


def load():
    """Load the in-tree library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libpllcart3d_amd.so is not built; run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    dp = C.POINTER(C.c_double)
    vp = C.c_void_p
    L.pllc_last_error.restype = C.c_char_p
    L.pllc_sim_create.restype = vp
    L.pllc_sim_create.argtypes = [C.POINTER(Params)]
    L.pllc_sim_create2.restype = vp
    L.pllc_sim_create2.argtypes = [C.POINTER(Params), C.c_int]
    L.pllc_sim_precision.argtypes = [vp]
    L.pllc_sim_destroy.argtypes = [vp]
    L.pllc_sim_dims.restype = C.c_size_t
    L.pllc_sim_dims.argtypes = [vp, C.POINTER(C.c_int)]
    L.pllc_sim_init.argtypes = [vp, dp]
    L.pllc_sim_upload.argtypes = [vp, C.c_int, dp]
    L.pllc_sim_download.argtypes = [vp, C.c_int, dp]
    L.pllc_sim_op.restype = C.c_double
    L.pllc_sim_op.argtypes = [vp, C.c_int, C.c_double, C.c_int, C.c_double]
    L.pllc_sim_run.argtypes = [vp, C.c_int, dp, dp]
    L.pllc_sim_timestep.restype = C.c_double
    L.pllc_sim_timestep.argtypes = [vp]
    L.pllc_sim_shift_conc.argtypes = [vp, C.c_double, C.c_double, C.c_int]
    L.pllc_sim_autotune.argtypes = [vp, C.c_int, C.POINTER(C.c_int)]
    L.pllc_ensemble_create.restype = vp
    L.pllc_ensemble_create.argtypes = [C.POINTER(Params), C.c_int, C.c_int]
    L.pllc_ensemble_destroy.argtypes = [vp]
    L.pllc_ensemble_size.argtypes = [vp]
    L.pllc_ensemble_batches.argtypes = [vp]
    L.pllc_ensemble_member.restype = vp
    L.pllc_ensemble_member.argtypes = [vp, C.c_int]
    L.pllc_ensemble_run.argtypes = [vp, C.c_int]
    L.pllc_ensemble_time_steps.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
    L.pllc_ensemble_batch_info.argtypes = [vp, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int)]
    L.pllc_ensemble_time_kernel.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, C.POINTER(C.c_double),
                                            C.POINTER(C.c_int)]
    L.pllc_sim_slope_parameter.restype = C.c_double
    L.pllc_sim_slope_parameter.argtypes = [vp]
    L.pllc_sim_synchronize.argtypes = [vp]
    L.pllc_ic_fill.argtypes = [C.POINTER(Params), C.c_int, C.c_ulonglong, C.c_double, dp]
    L.pllc_sim_time_steps.argtypes = [vp, C.c_int, C.c_int, dp]
    L.pllc_sim_time_kernel.argtypes = [vp, C.c_char_p, C.c_int, dp, C.POINTER(C.c_int)]
    L.pllc_kernel_names.restype = C.c_char_p
    fp = C.POINTER(C.c_float)
    ip = C.POINTER(C.c_int)
    L.pllc_cnn_create.restype = vp
    L.pllc_cnn_create.argtypes = [C.c_int, C.c_int, C.c_int]
    L.pllc_cnn_destroy.argtypes = [vp]
    L.pllc_cnn_nparams.argtypes = []
    L.pllc_cnn_set_params.argtypes = [vp, dp]
    L.pllc_cnn_get_params.argtypes = [vp, dp]
    L.pllc_cnn_set_norm.argtypes = [vp, dp, dp, C.c_double]
    L.pllc_cnn_forward.argtypes = [vp, fp, dp]
    L.pllc_cnn_add_sample.argtypes = [vp, fp, fp, C.POINTER(C.c_ubyte)]
    L.pllc_cnn_num_samples.argtypes = [vp]
    L.pllc_cnn_loss_grad.argtypes = [vp, C.c_int, dp, dp]
    L.pllc_cnn_train.argtypes = [vp, C.c_int, ip, C.c_int, C.c_double, dp]
    L.pllc_sim_set_closure.argtypes = [vp, vp, C.c_int]
    L.pllc_sim_run_guarded.argtypes = [vp, C.c_int, C.c_int, C.c_double, C.c_double, ip, dp]
    L.pllc_closure_guard_ok.argtypes = [dp, C.c_double, C.c_double, C.c_double]
    _lib = L
    return L


def last_error():
    return load().pllc_last_error().decode()
