"""ctypes binding of libpdgpu.so (the C ABI declared in include/pdgpu.h).

The library must have been built (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C peridigm_amd``). There is no CPU fallback: if the
library is missing or no GPU is present, loading fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpdgpu.so")

PDG_OK = 0
PDG_E_ARG = -1
PDG_E_HIP = -2
PDG_E_UNSUPPORTED = -3
PDG_E_NONFINITE = -4
PDG_E_NOCONV = -5

INFLUENCE = {"One": 0, "Parabolic Decay": 1, "Gaussian": 2}


class PdgError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pdgpu error {code}: {msg}")
        self.code = code


class pdg_bc(C.Structure):
    _fields_ = [("nodes", C.POINTER(C.c_int)), ("numNodes", C.c_int), ("coord", C.c_int),
                ("c0", C.c_double), ("c1", C.c_double)]


class pdg_qs_params(C.Structure):
    _fields_ = [("maxSolverIterations", C.c_int), ("tolerance", C.c_double),
                ("absoluteTolerance", C.c_int), ("disableHeuristics", C.c_int),
                ("dampedNewtonDiagonalScaleFactor", C.c_double),
                ("dampedNewtonDiagonalShiftFactor", C.c_double),
                ("adaptiveLoadStepping", C.c_int), ("maxLoadStepReductionsInOneStep", C.c_int),
                ("maxTotalLoadStepReductions", C.c_int), ("reduceAllRemainingLoadSteps", C.c_int),
                ("precond", C.c_int), ("linearTolerance", C.c_double),
                ("maxLinearIterations", C.c_int)]


class pdg_qs_step(C.Structure):
    _fields_ = [("timePrevious", C.c_double), ("timeCurrent", C.c_double),
                ("newtonIterations", C.c_int), ("linearIterations", C.c_int),
                ("residualNorm", C.c_double), ("toleranceMultiplier", C.c_double),
                ("status", C.c_int), ("dampedNewton", C.c_int), ("lastAlpha", C.c_double)]


_vp = C.c_void_p
_i = C.c_int
_ll = C.c_longlong
_d = C.c_double

# name -> (restype, argtypes)
SIGNATURES = {
    "pdg_ctx_create": (_i, [_i, C.POINTER(_vp)]),
    "pdg_ctx_destroy": (_i, [_vp]),
    "pdg_last_error": (C.c_char_p, [_vp]),
    "pdg_set_lanes_per_node": (_i, [_vp, _i]),
    "pdg_set_kernel_variant": (_i, [_vp, _i]),
    "pdg_set_tile_nodes": (_i, [_vp, _i]),
    "pdg_tile_config": (_i, [_vp, C.POINTER(_i), C.POINTER(_i)]),
    "pdg_get_tiling": (_i, [_vp, C.POINTER(_i), C.POINTER(_ll)]),
    "pdg_lds_schedule": (_i, [_vp, C.POINTER(_ll), C.POINTER(_ll), C.POINTER(_ll)]),
    "pdg_set_topology": (_i, [_vp, _i, _vp, _vp, _ll]),
    "pdg_set_constants": (_i, [_vp, _vp, _vp, _d, _d, _d, _i, _d]),
    "pdg_set_density": (_i, [_vp, _vp]),
    "pdg_set_critical_stretch": (_i, [_vp, _vp]),
    "pdg_get_sizes": (_i, [_vp, C.POINTER(_i), C.POINTER(_ll)]),
    "pdg_initialize": (_i, [_vp, _vp]),
    "pdg_compute_damage": (_i, [_vp, _vp, _vp, _vp, _vp, C.POINTER(_ll)]),
    "pdg_compute_force": (_i, [_vp, _d, _vp, _vp, _vp, _vp]),
    "pdg_state_upload": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pdg_state_download": (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "pdg_set_bcs": (_i, [_vp, C.POINTER(pdg_bc), _i]),
    "pdg_set_material_model": (_i, [_vp, _i, _i, _d]),
    "pdg_set_critical_stretch_value": (_i, [_vp, _d]),
    "pdg_set_critical_stretch_function": (_i, [_vp, _i, _vp, _vp]),
    "pdg_qs_params_default": (_i, [C.POINTER(pdg_qs_params)]),
    "pdg_qs_force_dilatation": (_i, [_vp, _vp, _vp, _vp]),
    "pdg_qs_run": (_i, [_vp, C.POINTER(pdg_bc), _i, _vp, _i, C.POINTER(pdg_qs_params), _vp,
                        C.POINTER(pdg_qs_step), _i, C.POINTER(_i)]),
    "pdg_qs_run_record": (_i, [_vp, C.POINTER(pdg_bc), _i, _vp, _i, C.POINTER(pdg_qs_params), _vp,
                               C.POINTER(pdg_qs_step), _i, C.POINTER(_i), _vp]),
    "pdg_set_thermal_expansion": (_i, [_vp, _d]),
    "pdg_set_temperature_change": (_i, [_vp, _vp, _vp]),
    "pdg_set_temperature_time": (_i, [_vp, _d]),
    "pdg_compute_partial_stress": (_i, [_vp, _vp, _vp, _vp]),
    "pdg_verlet_run": (_i, [_vp, _i, _d, _d, _i]),
    "pdg_verlet_run_persistent": (_i, [_vp, _i, _d, _d, _i, C.POINTER(_i)]),
    "pdg_critical_time_step": (_i, [_vp, _d, C.POINTER(_d)]),
    "pdg_set_contact": (_i, [_vp, _d, _i, _d, _d, _d]),
    "pdg_contact_search": (_i, [_vp, _vp, _vp, C.POINTER(_ll)]),
    "pdg_contact_compute_force": (_i, [_vp, _d, _vp, _vp, _vp, _ll, _vp]),
    "pdg_contact_force_density": (_i, [_vp, _vp]),
    "pdg_critical_time_step_noalloc": (_i, [_vp, _d, _vp]),
    "pdg_last_timing": (_i, [_vp, C.POINTER(_d), C.POINTER(_d), C.POINTER(_d)]),
    "pdg_set_per_pass_timing": (_i, [_vp, _i]),
    "pdg_near_critical_count": (_i, [_vp, C.POINTER(_ll)]),
    "pdg_qs_set_fixed": (_i, [_vp, _vp, _i]),
    "pdg_qs_residual": (_i, [_vp, _vp, _vp, C.POINTER(_d)]),
    "pdg_qs_tangent_apply": (_i, [_vp, _vp, _vp, _vp]),
    "pdg_qs_block_diagonal": (_i, [_vp, _vp, _vp]),
    "pdg_qs_solve": (_i, [_vp, _vp, _vp, _vp, _i, _d, _i, C.POINTER(_i), C.POINTER(_d)]),
    "pdg_qs_newton": (_i, [_vp, _vp, _d, _i, _i, _d, _i, C.POINTER(_i), C.POINTER(_i),
                           C.POINTER(_d)]),
    "pdg_qs_fallback_count": (_i, [_vp, C.POINTER(_i)]),
    "pdg_qs_set_absolute_tolerance": (_i, [_vp, _i]),
    "pdg_qs_tolerance_multiplier": (_i, [_vp, C.POINTER(_d)]),
    "pdg_qs_set_graphs": (_i, [_vp, _i]),
    "pdg_gnn_init": (_i, [_vp, C.c_uint]),
    "pdg_gnn_num_params": (_i, []),
    "pdg_gnn_get_params": (_i, [_vp, _vp]),
    "pdg_gnn_set_params": (_i, [_vp, _vp]),
    "pdg_gnn_forward_backward": (_i, [_vp, _vp, _vp, _vp, _vp]),
    "pdg_gnn_train": (_i, [_vp, _vp, _i, _vp, _vp, _i, C.c_float, C.c_uint, C.POINTER(C.c_float)]),
    "pdg_quickgrid": (_i, [_i, _i, _i, _vp, _vp, _d, _i, _vp, _vp, _vp, C.POINTER(_ll)]),
    "pdg_radius_search": (_i, [_i, _vp, _d, _i, _vp, _vp, C.POINTER(_ll)]),
}

_LIB = None


def load(path: str = LIB_PATH):
    """Load libpdgpu.so and bind every declared symbol (raises if absent)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(path):
        raise OSError(f"libpdgpu.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _LIB = lib
    return lib


def exported_symbols(path: str = LIB_PATH):
    lib = C.CDLL(path)
    return {name for name in SIGNATURES if hasattr(lib, name)}
