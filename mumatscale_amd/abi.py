"""ctypes mirror of include/mmca.h (the C ABI of the CA step library).

The structures here have the same field order and types as the C header; a
CPU test checks their sizes against the compiled library.
"""
import ctypes as C

ABI_VERSION = 1

OK = 0
ERR_ARG = 1
ERR_HIP = 2
ERR_COURANT = 3
ERR_NEGATIVE_CE = 4
ERR_GRAIN_CAPACITY = 5
ERR_NUC_THRESHOLD = 6
ERR_NO_DEVICE = 7
ERR_STATE = 8

STATUS_NAMES = {
    OK: "ok",
    ERR_ARG: "bad argument",
    ERR_HIP: "HIP runtime error",
    ERR_COURANT: "Instability by Courant criterion",
    ERR_NEGATIVE_CE: "negative ce (Courant instability)",
    ERR_GRAIN_CAPACITY: "grain capacity exceeded",
    ERR_NUC_THRESHOLD: "nucleation threshold draw does not terminate",
    ERR_NO_DEVICE: "no GPU device",
    ERR_STATE: "call out of order",
}

TEMP_UNIFORM = 0
TEMP_GRADIENT = 1
TEMP_AM_BEAM = 2

PREC_FP64 = 0
PREC_FP32DEV = 1

FIELD_CE = 0
FIELD_CL = 1
FIELD_FS = 2
FIELD_GR = 3
FIELD_DC = 4
FIELD_D = 5
FIELD_TEMP = 6
FIELD_THRESHOLD = 7
FIELD_MOLD = 8
FIELD_OCE = 9

TUNE_ZM_PLANES = 0
TUNE_ZM_ROWS = 1
TUNE_XCD_REMAP = 2
TUNE_GRAPH = 3
TUNE_ZM_LDS = 4
TUNE_CAPTURE_CHUNK = 5
TUNE_TRACK_D = 6
TUNE_BEAM_SEP = 7
TUNE_CAPTURE_BAND = 8

# partitioning branch codes recorded for the adjoint (fs_change_diffuse, growth.c:221-273)
BRANCH_LIQUID = 0
BRANCH_SOLID = 1
BRANCH_GROW = 2
BRANCH_FULL = 3
BRANCH_REMELT = 4


class Params(C.Structure):
    _fields_ = [
        ("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
        ("pad_bndy", C.c_int32 * 3),
        ("sb_coords", C.c_int32 * 3),
        ("gdim", C.c_int32 * 3),
        ("cell_size", C.c_double),
        ("origin", C.c_double * 3),
        ("dt", C.c_double),
        ("courant_limit", C.c_double),
        ("c_init", C.c_double),
        ("liquidus_temp", C.c_double),
        ("solidus_temp", C.c_double),
        ("m_solute", C.c_double * 3),
        ("part_coef", C.c_double),
        ("d_sol", C.c_double), ("d_liq", C.c_double),
        ("fsgrow", C.c_double),
        ("temp_model", C.c_int32),
        ("am_layers", C.c_int32),
        ("am_tracks", C.c_int32),
        ("gnsbx", C.c_int32),
        ("t_init", C.c_double),
        ("cool_rate", C.c_double),
        ("gradient", C.c_double), ("grad_coef", C.c_double),
        ("velocity", C.c_double), ("velo_coef", C.c_double),
        ("laser_power", C.c_double), ("laser_speed", C.c_double),
        ("laser_spot", C.c_double), ("laser_absorp", C.c_double),
        ("laser_hatch", C.c_double), ("sub_init_temp", C.c_double),
        ("th_conductivity", C.c_double), ("density", C.c_double),
        ("specific_heat", C.c_double),
        ("laser_start", C.c_double * 3),
        ("max_total_grains", C.c_int64),
        ("check_negative_ce", C.c_int32),
        ("precision", C.c_int32),
        ("timestep", C.c_uint64),
    ]


class FixedNuc(C.Structure):
    _fields_ = [(n, C.c_double) for n in ("x", "y", "z", "ax", "ay", "az", "threshold")]


class NucVolume(C.Structure):
    _fields_ = [(n, C.c_double) for n in
                ("density", "minx", "miny", "minz", "maxx", "maxy", "maxz", "threshold")]


class NucleationCfg(C.Structure):
    _fields_ = [
        ("random_seed", C.c_uint32),
        ("rank", C.c_int32),
        ("oriented", C.c_int32),
        ("do_surface", C.c_int32),
        ("bulk_density", C.c_double), ("bulk_mean", C.c_double), ("bulk_sigma", C.c_double),
        ("surf_density", C.c_double), ("surf_mean", C.c_double), ("surf_sigma", C.c_double),
        ("n_fixed", C.c_int32),
        ("n_volumes", C.c_int32),
        ("fixed", C.POINTER(FixedNuc)),
        ("volumes", C.POINTER(NucVolume)),
    ]


class Grain(C.Structure):
    _fields_ = [
        ("nuc_x", C.c_double), ("nuc_y", C.c_double), ("nuc_z", C.c_double),
        ("nuc_timestep", C.c_uint64),
        ("nuc_threshold", C.c_double),
        ("rotmat", (C.c_double * 3) * 3),
        ("rotang", C.c_double * 3),
    ]


class Report(C.Structure):
    _fields_ = [
        ("timestep", C.c_uint64),
        ("num_grains", C.c_int32),
        ("new_grains", C.c_int32),
        ("gindex", C.c_int32),
        ("status", C.c_int32),
        ("failed_step", C.c_uint64),
        ("solid_sum", C.c_double),
        ("solute_sum", C.c_double),
        ("solute_ledger", C.c_double),
    ]


def status_string(code):
    return STATUS_NAMES.get(int(code), "unknown status %d" % int(code))


class MmcaError(RuntimeError):
    """Raised where the reference calls error() (debug.h:54-58)."""

    def __init__(self, code, where=""):
        self.code = int(code)
        super().__init__("%s: %s (status %d)" % (where, status_string(code), self.code))
