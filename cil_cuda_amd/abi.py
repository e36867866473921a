"""ctypes mirror of include/cil_hip.h and the loader of the in-tree gfx950 library.

The structs below follow the C declarations field by field; `check_layout()` compares
their sizes with the library's own `cil_hip_sizeof`.  Loading fails loudly when the HIP
library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

MAX_SPECIES = 4
MAX_ELEMS = 2
MAX_TYPES = MAX_SPECIES * MAX_ELEMS
MAX_WALLS = 4
ABI_VERSION = 1

# cil_status
OK, E_INVALID, E_NO_DEVICE, E_CAPACITY, E_OUTSIDE, E_NONFINITE, E_HIP, E_BOND, E_IO = 0, -1, -2, -3, -4, -5, -6, -7, -8
(GSD_UINT8, GSD_UINT16, GSD_UINT32, GSD_UINT64, GSD_INT8, GSD_INT16, GSD_INT32, GSD_INT64, GSD_FLOAT,
 GSD_DOUBLE) = range(1, 11)
# cil_law
LAW_NONE, LAW_LJ, LAW_SC, LAW_HZ, LAW_EA = 0, 1, 2, 3, 4
# cil_spring
SPRING_FENE, SPRING_HARMONIC = 0, 1
# cil_noise
NOISE_NONE, NOISE_CELL_WIENER, NOISE_FDT = 0, 1, 2
# cil_rng
RNG_COUNTER, RNG_EXTERNAL, RNG_DSFMT = 0, 1, 2
# cil_division
DIV_NONE, DIV_THETA_CYCLE, DIV_LINEAR_GROWTH, DIV_BINARY = 0, 1, 2, 3
# cil_integrator
INT_EULER, INT_HEUN = 0, 1
# cil_precision
PREC_FP64, PREC_FP32 = 0, 1
# CellState (Cell.cuh:26-30)
CS_DIVIDING, CS_CRAWLING, CS_UNDEFINED = 0, 1, 2

D2 = C.c_double * MAX_ELEMS


class Gaussian(C.Structure):
    _fields_ = [("mu", C.c_double), ("sigma", C.c_double)]


class Species(C.Structure):
    _fields_ = [
        ("nelem", C.c_int),
        ("sigma", D2), ("zeta", D2), ("zeta_p", D2), ("zeta_p_beta", D2), ("motility", D2),
        ("eps_core", D2), ("eps_well", D2), ("sig_well", D2),
        ("cross_adhesion", C.c_double),
        ("spring", C.c_int),
        ("kappa", C.c_double), ("kappa_push", C.c_double), ("spring_ratio", C.c_double),
        ("r2max", C.c_double), ("div_r2max", C.c_double),
        ("rss", C.c_double),
        ("cil", C.c_int), ("quad_drag", C.c_int),
        ("k_harm", C.c_double),
        ("d_noise", C.c_double), ("d_noise_radial", C.c_double), ("d_noise_angular", C.c_double),
        ("kT", C.c_double),
        ("cycle_p0", C.c_double), ("cycle_tau_r", C.c_double), ("cycle_tau_v", C.c_double),
        ("cycle_tau_p", C.c_double), ("cycle_v_s", C.c_double), ("cycle_v_1", C.c_double),
        ("cycle_beta_vs", C.c_double),
        ("cycle_rmaxed_inv", C.c_double), ("cycle_rho_ratio", C.c_double), ("cycle_p_shift", C.c_double),
        ("cycle_p_shift_tau_inv", C.c_double), ("apop_pressure", C.c_double),
        ("division_tau", Gaussian), ("migration_tau", Gaussian),
        ("growth_rate", C.c_double), ("l_div", C.c_double), ("l_birth", C.c_double), ("sigma_jitter", C.c_double),
        ("no_division", C.c_int), ("division_swelling", C.c_int),
        ("division_r0", C.c_double), ("division_r2min", C.c_double), ("division_r2min_ratio", C.c_double),
        ("growthk0", D2), ("growthk1", D2), ("growthk2", D2),
    ]


class Wall(C.Structure):
    _fields_ = [("kind", C.c_int), ("axis", C.c_int), ("x0", C.c_double), ("rcut", C.c_double),
                ("epsilon", C.c_double), ("positive", C.c_int)]


class Params(C.Structure):
    _fields_ = [
        ("abi_version", C.c_int), ("dim", C.c_int), ("nspecies", C.c_int),
        ("species", Species * MAX_SPECIES),
        ("law", (C.c_int * MAX_TYPES) * MAX_TYPES),
        ("rcut2", (C.c_double * MAX_TYPES) * MAX_TYPES),
        ("force_shifted", C.c_int),
        ("dt", C.c_double),
        ("noise", C.c_int), ("active_noise", C.c_int), ("rng", C.c_int),
        ("seed", C.c_uint64),
        ("division", C.c_int),
        ("osmotic_pi", C.c_double),
        ("nwalls", C.c_int), ("walls", Wall * MAX_WALLS),
        ("drain", C.c_int), ("precision", C.c_int), ("pressure", C.c_int),
        ("max_disp", C.c_double),
        ("integrator", C.c_int),
        ("rng_offset", C.c_uint64),
    ]


class Run(C.Structure):
    _fields_ = [("nspecies", C.c_int), ("gts", C.c_int), ("frames", C.c_int), ("ncells", C.c_int * MAX_SPECIES),
                ("species", (C.c_char * 32) * MAX_SPECIES), ("elements", ((C.c_char * 32) * MAX_ELEMS) * MAX_SPECIES),
                ("kbt", C.c_double), ("seed", C.c_long), ("sigma_max", (C.c_double * MAX_ELEMS) * MAX_SPECIES),
                ("f_ext", (C.c_double * 3) * MAX_SPECIES), ("div_app_record", C.c_int)]


class Box(C.Structure):
    _fields_ = [("length", C.c_double * 3), ("periodic", C.c_int * 3)]


PD = C.POINTER(C.c_double)
PI = C.POINTER(C.c_int)


class SpeciesView(C.Structure):
    _fields_ = [
        ("ncells", C.c_int), ("capacity", C.c_int),
        ("pos", PD), ("frc", PD), ("vel", PD), ("sigma", PD), ("sigma_t", PD), ("pressure", PD),
        ("state", PI), ("t0", PI), ("tswitch", PI), ("tpressure", PI),
        ("theta", PD), ("rho", PD), ("ang_vel", PD), ("rest_len", PD),
    ]


VIEW_DOUBLE_FIELDS = ("pos", "frc", "vel", "sigma", "sigma_t", "pressure", "theta", "rho", "ang_vel", "rest_len")
VIEW_INT_FIELDS = ("state", "t0", "tswitch", "tpressure")

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libcil_hip.so")
_lib = None


class SurrogateStats(C.Structure):
    _fields_ = [
        ("macro_steps", C.c_long), ("plain_steps", C.c_long), ("judged", C.c_long), ("failed", C.c_long),
        ("last_err_rms", C.c_double), ("last_err_max", C.c_double), ("worst_err_rms", C.c_double),
        ("knn_list_builds", C.c_long), ("knn_fallbacks", C.c_long),
    ]


MLP_NPARAM = 26243
MLP_NFEAT = 72
MLP_OUT_SCALE = 0.002   # length units per unit network output (the drift correction)
SURR_BASE_SCALE = 0.01  # length units per unit of features 66-71 (one-pass drift, noise sum)


class CilError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cil_hip error {code}: {msg}")
        self.code = code


def lib():
    """Load the in-tree gfx950 library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    L.cil_hip_create.argtypes = [C.POINTER(Params), C.POINTER(Box), PI, C.POINTER(vp)]
    L.cil_hip_destroy.argtypes = [vp]
    L.cil_hip_upload.argtypes = [vp, C.POINTER(SpeciesView), C.c_int]
    L.cil_hip_step.argtypes = [vp, C.c_int, PI]
    L.cil_hip_download.argtypes = [vp, C.POINTER(SpeciesView), C.c_int, PI]
    L.cil_hip_forces.argtypes = [vp, PD, PD]
    L.cil_hip_pairs.argtypes = [vp, C.POINTER(C.c_int32), C.c_size_t, C.POINTER(C.c_size_t)]
    L.cil_hip_pairs_noalloc.argtypes = [vp, C.POINTER(C.c_int32), C.c_size_t, C.POINTER(C.c_size_t)]
    L.cil_hip_check_positions.argtypes = [vp, PI]
    L.cil_hip_set_external_noise.argtypes = [vp, PD, C.c_size_t, C.c_int]
    L.cil_hip_num_cells.argtypes = [vp, PI]
    L.cil_hip_timings.argtypes = [vp, PD, C.c_int, PI, C.c_char_p, C.c_size_t]
    L.cil_hip_bin_info.argtypes = [vp, PI, PD]
    L.cil_hip_set_tuning.argtypes = [vp, C.c_int, C.c_int]
    L.cil_hip_set_timing.argtypes = [vp, C.c_int]
    L.cil_hip_set_kernel.argtypes = [vp, C.c_int]
    L.cil_hip_set_neighbour_list.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    L.cil_hip_set_neighbour_list_inner.argtypes = [vp, C.c_double]
    L.cil_hip_autotune.argtypes = [vp, C.c_int, PI, PI, PI, PD]
    PF = C.POINTER(C.c_float)
    PL = C.POINTER(C.c_long)
    L.cil_hip_mlp_nparams.argtypes = []
    L.cil_hip_mlp_set_params.argtypes = [vp, PF, C.c_size_t]
    L.cil_hip_mlp_get_params.argtypes = [vp, PF, C.c_size_t]
    L.cil_hip_mlp_init.argtypes = [vp, C.c_uint64]
    L.cil_hip_mlp_forward.argtypes = [vp, PF, C.c_size_t, PF]
    L.cil_hip_mlp_grad.argtypes = [vp, PF, PF, C.c_size_t, PF, PF]
    L.cil_hip_knn_features.argtypes = [vp, PF, C.c_size_t]
    L.cil_hip_surrogate_collect.argtypes = [vp, C.c_int, C.c_int, C.c_long, PI, PL]
    L.cil_hip_surrogate_set_samples.argtypes = [vp, PF, PF, C.c_long]
    L.cil_hip_surrogate_get_samples.argtypes = [vp, PF, PF, C.c_long, PL]
    L.cil_hip_surrogate_clear_samples.argtypes = [vp]
    L.cil_hip_surrogate_train.argtypes = [vp, C.c_int, C.c_int, C.c_float, C.c_uint64, PF]
    L.cil_hip_surrogate_eval.argtypes = [vp, PI, PD, PD]
    L.cil_hip_surrogate_step.argtypes = [vp, C.c_int, C.c_int, C.c_double, PI, C.POINTER(SurrogateStats)]
    L.cil_hip_surrogate_stats_reset.argtypes = [vp]
    L.cil_hip_batch_create.argtypes = [C.POINTER(vp), C.c_int, C.POINTER(vp)]
    L.cil_hip_batch_destroy.argtypes = [vp]
    L.cil_hip_batch_size.argtypes = [vp]
    L.cil_hip_batch_step.argtypes = [vp, C.c_int, PI]
    L.cil_hip_batch_autotune.argtypes = [vp, C.c_int, PI, PI, PI, PD]
    L.cil_hip_batch_last_error.argtypes = [vp]
    L.cil_hip_batch_last_error.restype = C.c_char_p
    L.cil_hip_last_error.argtypes = [vp]
    L.cil_hip_last_error.restype = C.c_char_p
    L.cil_hip_version.restype = C.c_char_p
    L.cil_hip_config_read.argtypes = [C.c_char_p, C.POINTER(Params), C.POINTER(Box), C.POINTER(Run), C.c_char_p,
                                      C.c_size_t]
    L.cil_hip_config_init.argtypes = [C.c_char_p, C.POINTER(Params), C.POINTER(SpeciesView), C.c_int, C.c_char_p,
                                      C.c_size_t]
    L.cil_hip_dsfmt_fill.argtypes = [C.c_uint32, C.c_uint64, C.c_int, PD, C.c_size_t]
    L.cil_hip_gsd_create.argtypes = [C.c_char_p, C.c_char_p, C.c_char_p, C.c_uint32, C.POINTER(vp)]
    L.cil_hip_gsd_write_chunk.argtypes = [vp, C.c_char_p, C.c_int, C.c_uint64, C.c_uint32, vp]
    L.cil_hip_gsd_end_frame.argtypes = [vp]
    L.cil_hip_gsd_close.argtypes = [vp]
    L.cil_hip_gsd_write_frame.argtypes = [vp, C.POINTER(Params), C.POINTER(Box), C.POINTER(Run), C.POINTER(SpeciesView),
                                          C.c_int, C.c_int]
    L.cil_hip_gsd_write_frame_ctx.argtypes = [vp, vp, C.POINTER(Run), C.c_int]
    L.cil_hip_h5_create.argtypes = [C.c_char_p, C.POINTER(Params), C.POINTER(Box), C.POINTER(Run), C.POINTER(vp)]
    L.cil_hip_h5_write_frame.argtypes = [vp, C.POINTER(Params), C.POINTER(SpeciesView), C.c_int, C.POINTER(Run), C.c_int,
                                         C.c_int, C.c_double]
    L.cil_hip_h5_write_frame_ctx.argtypes = [vp, vp, C.POINTER(Run), C.c_int, C.c_int, C.c_double]
    L.cil_hip_h5_close.argtypes = [vp]
    L.cil_hip_set_division_log.argtypes = [vp, C.c_int]
    L.cil_hip_division_log.argtypes = [vp, C.c_int, PD, C.POINTER(C.c_int), C.c_int, C.POINTER(C.c_int), C.c_int]
    L.cil_hip_sizeof.argtypes = [C.c_int]
    L.cil_hip_sizeof.restype = C.c_size_t
    _lib = L
    return L


def check_layout(L=None):
    L = L or lib()
    want = {0: Params, 1: Species, 2: SpeciesView, 3: Box, 4: SurrogateStats, 5: Run}
    for k, cls in want.items():
        got = L.cil_hip_sizeof(k)
        if got != C.sizeof(cls):
            raise RuntimeError(f"ABI layout mismatch for {cls.__name__}: C {got} vs ctypes {C.sizeof(cls)}")
    return True


EXPORTS = (
    "cil_hip_create", "cil_hip_destroy", "cil_hip_upload", "cil_hip_step", "cil_hip_download",
    "cil_hip_forces", "cil_hip_pairs", "cil_hip_pairs_noalloc", "cil_hip_check_positions", "cil_hip_set_external_noise",
    "cil_hip_num_cells", "cil_hip_timings", "cil_hip_bin_info", "cil_hip_set_tuning", "cil_hip_set_timing", "cil_hip_set_kernel", "cil_hip_autotune",
    "cil_hip_set_neighbour_list", "cil_hip_set_neighbour_list_inner",
    "cil_hip_sizeof", "cil_hip_last_error", "cil_hip_version",
    "cil_hip_config_read", "cil_hip_config_init", "cil_hip_dsfmt_fill", "cil_hip_gsd_create", "cil_hip_gsd_write_chunk",
    "cil_hip_gsd_end_frame", "cil_hip_gsd_close", "cil_hip_gsd_write_frame", "cil_hip_gsd_write_frame_ctx",
    "cil_hip_h5_create", "cil_hip_h5_write_frame", "cil_hip_h5_write_frame_ctx", "cil_hip_h5_close",
    "cil_hip_set_division_log", "cil_hip_division_log",
    "cil_hip_batch_create", "cil_hip_batch_destroy", "cil_hip_batch_size", "cil_hip_batch_step",
    "cil_hip_batch_autotune", "cil_hip_batch_last_error",
    "cil_hip_mlp_nparams", "cil_hip_mlp_set_params", "cil_hip_mlp_get_params", "cil_hip_mlp_init",
    "cil_hip_mlp_forward", "cil_hip_mlp_grad", "cil_hip_knn_features", "cil_hip_surrogate_collect",
    "cil_hip_surrogate_set_samples", "cil_hip_surrogate_get_samples",
    "cil_hip_surrogate_clear_samples", "cil_hip_surrogate_train", "cil_hip_surrogate_eval",
    "cil_hip_surrogate_step", "cil_hip_surrogate_stats_reset",
)
