"""ctypes declarations of the C-ABI in include/cabanapic_amd.h.

The shared library is built in-tree (cabanapic_amd/libcabanapic_amd.so) by
``__graft_entry__.build()``; loading fails loudly when it is missing, so no
caller can silently fall back to another implementation.  The environment
variable CABANAPIC_AMD_LIB names another build of the same library (an
installed copy, or a variant build being compared).
"""
import ctypes
import os

LIB_PATH = os.environ.get("CABANAPIC_AMD_LIB") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "libcabanapic_amd.so")

CPIC_OK = 0
CPIC_ERR_INVALID = 1
CPIC_ERR_UNSUPPORTED = 2
CPIC_ERR_HIP = 3
CPIC_ERR_DEVICE = 4

BOUNDARY_REFLECT = 0
BOUNDARY_PERIODIC = 1

FIELD_SOLVER_EM = 0
FIELD_SOLVER_ES_1D = 1

STAGES = ("interp", "clear", "push", "unload", "advance_b", "advance_e", "energy", "sort")
NUM_STAGES = len(STAGES)

# every symbol the header declares: name -> (restype, argtypes)
_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_D = ctypes.c_double
_PP = ctypes.POINTER(ctypes.c_void_p)


class Config(ctypes.Structure):
    _fields_ = [("real_size", ctypes.c_int32), ("num_ghosts", ctypes.c_int32),
                ("nx", ctypes.c_int64), ("ny", ctypes.c_int64), ("nz", ctypes.c_int64),
                ("max_particles", ctypes.c_int64), ("column_rows", ctypes.c_int64)]


class StepParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in (
        "qdt_2mc", "cdt_dx", "cdt_dy", "cdt_dz", "qsp", "dx", "dy", "dz", "dt",
        "px", "py", "pz", "dt_eps0", "me", "c")] + [
        ("boundary", ctypes.c_int32), ("energy_interval", ctypes.c_int32),
        ("compact_interval", ctypes.c_int32), ("kinetic_interval", ctypes.c_int32)]


class Snapshot(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int64), ("nx", ctypes.c_int64), ("real_size", ctypes.c_int32),
                ("x", ctypes.c_void_p), ("ux", ctypes.c_void_p),
                ("cell", ctypes.POINTER(ctypes.c_int32)), ("ex_line", ctypes.c_void_p),
                ("e_energy", ctypes.c_double), ("b_energy", ctypes.c_double),
                ("kinetic_energy", ctypes.c_double)]


class PlasmaSpec(ctypes.Structure):
    _fields_ = [("seed", ctypes.c_uint64), ("nx", ctypes.c_int64), ("ny", ctypes.c_int64),
                ("nz", ctypes.c_int64), ("nppc", ctypes.c_int64), ("weight", ctypes.c_double),
                ("uth", ctypes.c_double * 3), ("drift", ctypes.c_double * 3),
                ("pert_amp", ctypes.c_double), ("len_x", ctypes.c_double),
                ("pert_mode", ctypes.c_int32), ("counter_streaming", ctypes.c_int32),
                ("profile", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("len_y", ctypes.c_double), ("ring_radius", ctypes.c_double),
                ("ring_width", ctypes.c_double), ("column_charge", ctypes.c_double),
                ("guide_bz", ctypes.c_double)]


SIGNATURES = {
    "cpic_create": (ctypes.c_int, [ctypes.POINTER(Config), _PP]),
    "cpic_destroy": (ctypes.c_int, [_P]),
    "cpic_last_error": (ctypes.c_char_p, []),
    "cpic_synchronize": (ctypes.c_int, [_P]),
    "cpic_get_stream": (ctypes.c_int, [_P, _PP]),
    "cpic_upload_particles": (ctypes.c_int, [_P, _I64] + [_P] * 8),
    "cpic_download_particles": (ctypes.c_int, [_P, _I64] + [_P] * 8),
    "cpic_num_particles": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "cpic_download_particle_tags": (ctypes.c_int, [_P, _I64, _P]),
    "cpic_column_rows": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "cpic_reserve_column_rows": (ctypes.c_int, [_P, _I64]),
    "cpic_field_energies_noalloc": (ctypes.c_int, [_P, _I32, _I64, _I64, _I64, _P, _P]),
    "cpic_adjoint_create": (ctypes.c_int, [_P, _PP]),
    "cpic_adjoint_destroy": (ctypes.c_int, [_P]),
    "cpic_forward": (ctypes.c_int, [_P, ctypes.POINTER(StepParams), _I64, _I32, _P]),
    "cpic_backward": (ctypes.c_int, [_P, _P, _P, _P, _P]),
    "cpic_adjoint_memory": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64),
                                           ctypes.POINTER(_I64)]),
    "cpic_upload_fields": (ctypes.c_int, [_P, _P]),
    "cpic_download_fields": (ctypes.c_int, [_P, _P]),
    "cpic_upload_interpolators": (ctypes.c_int, [_P, _P]),
    "cpic_download_interpolators": (ctypes.c_int, [_P, _P]),
    "cpic_upload_accumulators": (ctypes.c_int, [_P, _P]),
    "cpic_download_accumulators": (ctypes.c_int, [_P, _P]),
    "cpic_load_interpolator_array": (ctypes.c_int, [_P]),
    "cpic_clear_accumulator_array": (ctypes.c_int, [_P]),
    "cpic_push": (ctypes.c_int, [_P, _D, _D, _D, _D, _D, _I32]),
    "cpic_uncenter_particles": (ctypes.c_int, [_P, _D]),
    "cpic_unload_accumulator_array": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "cpic_advance_b": (ctypes.c_int, [_P, _D, _D, _D]),
    "cpic_advance_e": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "cpic_update_ghosts_j": (ctypes.c_int, [_P]),
    "cpic_set_field_solver": (ctypes.c_int, [_P, _I32]),
    "cpic_set_compensated_deposit": (ctypes.c_int, [_P, _I32]),
    "cpic_energies": (ctypes.c_int, [_P, ctypes.POINTER(_D), ctypes.POINTER(_D)]),
    "cpic_kinetic_energy": (ctypes.c_int, [_P, _D, _D, ctypes.POINTER(_D)]),
    "cpic_sort_particles": (ctypes.c_int, [_P]),
    "cpic_step": (ctypes.c_int, [_P, ctypes.POINTER(StepParams), _I64, _P]),
    "cpic_step_async": (ctypes.c_int, [_P, ctypes.POINTER(StepParams), _I64]),
    "cpic_snapshot_begin": (ctypes.c_int, [_P, _I32]),
    "cpic_snapshot_wait": (ctypes.c_int, [_P, _I32, ctypes.POINTER(Snapshot)]),
    "cpic_set_timing": (ctypes.c_int, [_P, _I32]),
    "cpic_set_timing_stages": (ctypes.c_int, [_P, ctypes.c_uint32]),
    "cpic_stage_times": (ctypes.c_int, [_P, _P, _P]),
    "cpic_last_push_movers": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
    "cpic_device_errors": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "cpic_gauss_residual": (ctypes.c_int, [_P] + [_D] * 5 + [_I32, _P, _P]),
    "cpic_init_plasma": (ctypes.c_int, [ctypes.POINTER(PlasmaSpec), _I32, _I64] + [_P] * 8),
    "cpic_init_column_fields": (ctypes.c_int, [ctypes.POINTER(PlasmaSpec), _I32, _P]),
    "cpic_tangent_create": (ctypes.c_int, [ctypes.POINTER(Config), _PP]),
    "cpic_tangent_destroy": (ctypes.c_int, [_P]),
    "cpic_tangent_contexts": (ctypes.c_int, [_P, _PP, _PP]),
    "cpic_tangent_upload_particles": (ctypes.c_int, [_P, _I64] + [_P] * 9),
    "cpic_tangent_download_particles": (ctypes.c_int, [_P, _I64] + [_P] * 9),
    "cpic_tangent_push": (ctypes.c_int, [_P, _D, _D, _D, _D, _D, _I32]),
    "cpic_tangent_step": (ctypes.c_int, [_P, ctypes.POINTER(StepParams), _I64, _P]),
    "cpic_tangent_set_graph": (ctypes.c_int, [_P, _I32]),
    "cpic_tangent_last_push_movers": (ctypes.c_int, [_P, ctypes.POINTER(_I64)]),
}


class CpicError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"cabanapic_amd error {code}: {msg}")
        self.code = code


_lib = None


def load():
    """Load the in-tree HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                "(the HIP path has no fallback)")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(code):
    if code != CPIC_OK:
        raise CpicError(code, load().cpic_last_error().decode(errors="replace"))
    return code
