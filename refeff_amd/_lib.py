"""ctypes binding of librefeff_fms.so (include/refeff_fms.h).

The library is built in-tree by ``refeff_amd/build.py``.  There is no
fallback: if the shared object is missing every entry point raises
``FmsLibraryMissing``.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librefeff_fms.so")
TEST_LIB_PATH = os.path.join(HERE, "librefeff_fms_test.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "refeff_fms.h")

STATUS_NAMES = {
    0: "OK", 1: "INVALID_ARGUMENT", 2: "MISSING_CENTRAL_ATOM", 3: "DUPLICATE_ABSORBER", 4: "NON_FINITE",
    5: "ZERO_RHO", 6: "POTENTIAL_OUT_OF_RANGE", 7: "MISSING_REPRESENTATIVE", 8: "SINGULAR", 9: "BREAKDOWN",
    10: "CENTRAL_MISMATCH", 11: "UNSUPPORTED", 12: "NO_MEMORY", 13: "NO_CONVERGENCE", 14: "DEVICE_ERROR",
    15: "INVALID_SPIN_COUNT", 16: "EMPTY_CLUSTER", 17: "IO_ERROR",
}

SOLVER_AUTO, SOLVER_MIXED, SOLVER_FP64, SOLVER_GGBI, SOLVER_GGRM, SOLVER_GGGM, SOLVER_GGTF = 0, 1, 2, 3, 4, 5, 6
SOLVER_BICGSTAB = 7
PRECOND_NONE, PRECOND_BLOCK_JACOBI, PRECOND_LEARNED = 0, 1, 2


class FmsLibraryMissing(RuntimeError):
    pass


class FmsError(RuntimeError):
    """Non-zero refeff_fms_status (mirrors refeff's FmsError families); ``variant``
    and ``payload`` carry the FmsError variant when the library reported one."""

    def __init__(self, code: int, where: str, detail: str = "", error=None):
        name = STATUS_NAMES.get(code, str(code))
        if error is not None and error.variant and not detail:
            detail = error.message.decode()
        super().__init__(f"{where}: {name}" + (f" ({detail})" if detail else ""))
        self.code = code
        self.name = name
        self.variant = ERROR_VARIANTS.get(error.variant, str(error.variant)) if error is not None else None
        self.payload = (tuple(error.values), error.real_value, error.name.decode()) if error is not None else None


class PlanDesc(C.Structure):
    _fields_ = [
        ("lfms", C.c_int32), ("minv", C.c_int32), ("spin_channels", C.c_int32), ("spin_selector", C.c_int32),
        ("atom_count", C.c_int32), ("positions", C.c_void_p), ("potentials", C.c_void_p),
        ("max_potential", C.c_int32), ("global_lmax", C.c_int32), ("raw_potential_lmax", C.c_void_p),
        ("direct_cutoff", C.c_float), ("mean_square_displacements", C.c_void_p),
        ("convergence_tolerance", C.c_float), ("zero_tolerance", C.c_float),
        ("full_scattering_requested", C.c_int32), ("solver", C.c_int32), ("max_batch", C.c_int32),
        ("state_capacity", C.c_int32), ("calculated_l", C.c_void_p),
    ]


class ErrorDetail(C.Structure):
    """refeff_fms_error: the FmsError variant and payload behind a status."""
    _fields_ = [("status", C.c_int32), ("variant", C.c_int32), ("values", C.c_int64 * 3), ("real_value", C.c_double),
                ("name", C.c_char * 24), ("message", C.c_char * 160)]


ERROR_VARIANTS = {
    0: "None", 1: "InvalidAngularLimit", 2: "EmptyCluster", 3: "NonFiniteCoordinate", 4: "PotentialOutOfRange",
    5: "NonFiniteMeanSquareDisplacement", 6: "InvalidDirectCutoff", 7: "InvalidTolerance",
    8: "InvalidSpinChannelCount", 9: "StateCapacityExceeded", 10: "MissingRepresentativePotential",
    11: "FullScatteringRequiresLu", 12: "NonFiniteWaveNumber", 13: "NonFinitePhaseShift", 14: "SingularMatrix",
    15: "IterativeSolverBreakdown", 16: "IterativeSolverNoConvergence", 17: "RefinementNoConvergence",
    18: "Unsupported", 19: "Device", 20: "InvalidArgument",
}


class PlanInfo(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("state_count", "padded_order", "channel_count", "potential_count",
                                          "potential_start", "potential_end", "rhs_count", "solver", "batch",
                                          "lu_block")]


class Stats(C.Structure):
    _fields_ = [("seconds_total", C.c_double), ("seconds_assemble", C.c_double), ("seconds_factor", C.c_double),
                ("seconds_solve", C.c_double), ("refinement_sweeps", C.c_int32), ("batches", C.c_int32),
                ("max_correction", C.c_double), ("krylov_energies", C.c_int32), ("krylov_iterations", C.c_int32),
                ("lu_energies", C.c_int32), ("fp64_fallback_energies", C.c_int32)]


class Profile(C.Structure):
    _fields_ = [("milliseconds", C.c_double * 9), ("flops", C.c_double * 9), ("bytes", C.c_double * 9),
                ("launches", C.c_int64 * 9)]


_lib = None


def header_functions() -> list[str]:
    """Function names declared in include/refeff_fms.h."""
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(refeff_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FmsLibraryMissing(f"{LIB_PATH} is not built; run `python refeff_amd/build.py`")
    L = C.CDLL(LIB_PATH)
    vp, i32, i32p, f32p, f64p = C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_float), C.POINTER(C.c_double)
    L.refeff_fms_sort_atoms_by_radius.argtypes = [i32, f32p, i32p, f64p]
    L.refeff_fms_yprep_cluster.argtypes = [i32, f32p, i32p, i32, C.c_float, i32, f32p, i32p, i32p, i32p, i32p]
    L.refeff_fms_sort_representative_atoms.argtypes = [i32, i32, i32, f32p, i32p, i32p]
    L.refeff_fms_debye_sigsqr.argtypes = [i32, C.c_double, C.c_double, C.c_double, i32, f32p, i32p, C.c_float, f32p,
                                          i32p]
    L.refeff_fms_state_kets.argtypes = [i32, i32, i32p, i32, i32p, i32, i32p, i32p, i32p, i32p, i32p, i32p]
    L.refeff_fms_plan_create.argtypes = [C.POINTER(PlanDesc), C.POINTER(vp)]
    L.refeff_fms_last_error.argtypes = [C.POINTER(ErrorDetail)]
    L.refeff_fms_plan_energy_error.argtypes = [vp, i32, C.POINTER(ErrorDetail)]
    L.refeff_fms_plan_destroy.argtypes = [vp]
    L.refeff_fms_plan_destroy.restype = None
    L.refeff_fms_plan_release_workspace.argtypes = [vp]
    L.refeff_fms_plan_release_workspace.restype = C.c_int
    L.refeff_fms_plan_query.argtypes = [vp, C.POINTER(PlanInfo)]
    L.refeff_fms_solve_batch.argtypes = [vp, i32, f64p, f64p, f64p, i32p, C.POINTER(Stats)]
    L.refeff_fms_solve_batch_device.argtypes = [vp, i32, vp, vp, vp, i32p, C.POINTER(Stats), vp]
    L.refeff_fms_solve_batch_full.argtypes = [vp, i32, f64p, f64p, f64p, f64p, i32p, C.POINTER(Stats)]
    L.refeff_fms_transition_bmat_averaged.argtypes = [i32, i32, i32, i32, i32, i32p, f64p]
    L.refeff_fms_transition_bmat.argtypes = [i32, i32, i32, i32, f64p, i32, i32, i32, C.c_double, i32p, f64p]
    L.refeff_fms_mkgtr_trace.argtypes = [i32, i32, i32, i32, i32p, f64p, f64p, f64p, f64p]
    L.refeff_fms_run_in_dir.argtypes = [C.c_char_p, i32, i32p, C.c_char_p, C.c_size_t]
    L.refeff_mkgtr_run_in_dir.argtypes = [C.c_char_p, i32p, C.c_char_p, C.c_size_t]
    L.refeff_fms_version.restype = C.c_char_p
    L.refeff_fms_gnn_create.argtypes = [vp, C.c_float, C.c_uint64, C.POINTER(vp)]
    L.refeff_fms_gnn_destroy.argtypes = [vp]
    L.refeff_fms_gnn_destroy.restype = None
    L.refeff_fms_gnn_info.argtypes = [vp, i32p, i32p, i32p]
    L.refeff_fms_gnn_get_params.argtypes = [vp, f32p]
    L.refeff_fms_gnn_set_params.argtypes = [vp, f32p]
    L.refeff_fms_gnn_collect.argtypes = [vp, vp, i32, f64p, f64p, i32p]
    L.refeff_fms_gnn_loss_grad.argtypes = [vp, f64p, f32p]
    L.refeff_fms_gnn_train.argtypes = [vp, i32, C.c_float, f64p]
    L.refeff_fms_gnn_weights.argtypes = [vp, i32, f32p]
    L.refeff_fms_gnn_apply.argtypes = [vp, i32, f64p, f64p]
    L.refeff_fms_plan_set_preconditioner.argtypes = [vp, i32, vp]
    L.refeff_fms_plan_get_tuning.argtypes = [vp, vp]
    L.refeff_fms_plan_set_tuning.argtypes = [vp, vp]
    L.refeff_fms_plan_tune.argtypes = [vp, i32, f64p, f64p, vp]
    L.refeff_fms_plan_set_create.argtypes = [vp, i32, C.c_char_p, C.POINTER(vp)]
    L.refeff_fms_plan_set_destroy.argtypes = [vp]
    L.refeff_fms_plan_set_destroy.restype = None
    L.refeff_fms_plan_set_plan.argtypes = [vp, i32, C.POINTER(vp)]
    L.refeff_fms_plan_set_tune.argtypes = [vp, vp, vp, vp, i32p]
    L.refeff_fms_plan_set_solve.argtypes = [vp, vp, vp, vp, vp, i32p, vp]
    L.refeff_fms_plan_set_mode.argtypes = [vp, i32]
    L.refeff_fms_plan_profile.argtypes = [vp, i32]
    L.refeff_fms_plan_profile_read.argtypes = [vp, C.POINTER(Profile)]
    L.refeff_ff2x_xanes_xmu.argtypes = [vp, f64p, f64p, f64p, f64p, f64p, f64p, f64p, C.c_char_p, C.c_size_t]
    _lib = L
    return L


_test_lib = None


def test_lib():
    """The test build of the library (adds the kernel-level debug entry points)."""
    global _test_lib
    if _test_lib is None:
        if not os.path.exists(TEST_LIB_PATH):
            raise FmsLibraryMissing(f"{TEST_LIB_PATH} is not built; run `python refeff_amd/build.py`")
        _test_lib = C.CDLL(TEST_LIB_PATH)
    return _test_lib


def check(rc: int, where: str, detail: str = "", error=None):
    if rc != 0:
        raise FmsError(rc, where, detail, error)


def last_error() -> ErrorDetail:
    e = ErrorDetail()
    lib().refeff_fms_last_error(C.byref(e))
    return e
