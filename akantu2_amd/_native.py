"""ctypes loader of the gfx950 library ``libakantu_mi355x.so``.

The library is built in-tree by ``akantu2_amd.build.build_library()``.  There
is no fallback: if the library is missing or no gfx950 device is visible the
model constructors raise.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libakantu_mi355x.so")

# status codes (include/akantu_mi355x.h)
AKX_OK = 0
ERRORS = {-1: "invalid argument", -2: "HIP runtime error", -3: "negative jacobian",
          -4: "device out of memory", -5: "unsupported"}

TET4, HEX8, TRI3, QUAD4 = 0, 1, 2, 3
LAW_ELASTIC, LAW_ELASTIC_FD, LAW_NEOHOOKEAN, LAW_PLASTIC_LIH = 0, 1, 2, 3
LAW_USER, LAW_USER_FD = 4, 5  # computeStress supplied by the caller
PRECISION_FP64, PRECISION_FP32, PRECISION_FP64_FAST = 0, 1, 2
FIELD_DISPLACEMENT, FIELD_VELOCITY, FIELD_ACCELERATION = 0, 1, 2
FIELD_INTERNAL_FORCE, FIELD_EXTERNAL_FORCE, FIELD_MASS = 3, 4, 5
FIELD_BLOCKED_DOFS, FIELD_NODES = 6, 7
INTERNAL_GRAD_U, INTERNAL_STRESS, INTERNAL_INELASTIC_STRAIN = 0, 1, 2
INTERNAL_ISO_HARDENING = 3


class AkantuError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, 'error')} ({code}): {msg}")
        self.code = code


class JacobianError(AkantuError):
    """Mirror of the reference's IntegratorGaussException."""


_lib = None

_vp = C.c_void_p
_i64 = C.c_int64
_dp = C.POINTER(C.c_double)

# int (*)(void *user, int material_id, int64_t nb_qp, const double *grad_u,
#         double *stress)
COMPUTE_STRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.c_int64, _dp, _dp)


def _declare(lib):
    sigs = {
        "akx_last_error": (C.c_char_p, []),
        "akx_version": (C.c_int, []),
        "akx_device_count": (C.c_int, [C.POINTER(C.c_int)]),
        "akx_model_create": (C.c_int, [C.c_int, C.POINTER(_vp)]),
        "akx_model_destroy": (C.c_int, [_vp]),
        "akx_model_set_mesh": (C.c_int, [_vp, C.c_int, _i64, _vp, _i64, _vp]),
        "akx_model_add_material": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _vp, _i64,
                                             C.POINTER(C.c_int)]),
        "akx_model_set_element_parameters": (C.c_int, [_vp, C.c_int, _vp, _vp, _vp, _vp]),
        "akx_material_law_by_name": (C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_int)]),
        "akx_model_set_precision": (C.c_int, [_vp, C.c_int]),
        "akx_model_set_eigen_grad_u": (C.c_int, [_vp, C.c_int, _vp]),
        "akx_model_set_thermal": (C.c_int, [_vp, C.c_int, C.c_double, _vp]),
        "akx_model_set_stress_callback": (C.c_int, [_vp, COMPUTE_STRESS_FN, _vp]),
        "akx_model_set_device_stress": (C.c_int, [_vp, C.c_int, C.c_char_p, _vp, C.c_int]),
        "akx_fe_gradient_on_integration_points": (C.c_int, [_vp, _vp, C.c_int, _vp, _i64,
                                                            _vp]),
        "akx_fe_compute_btd": (C.c_int, [_vp, _vp, C.c_int, _vp, _i64, _vp]),
        "akx_fe_integrate": (C.c_int, [_vp, _vp, C.c_int, _vp, _i64, _vp]),
        "akx_dof_assemble_elemental_array": (C.c_int, [_vp, _vp, C.c_int, _vp, _i64,
                                                       C.c_double, _vp]),
        "akx_model_init_full": (C.c_int, [_vp]),
        "akx_model_get_stable_time_step": (C.c_int, [_vp, _dp]),
        "akx_model_set_time_step": (C.c_int, [_vp, C.c_double]),
        "akx_model_solve_step": (C.c_int, [_vp, C.c_char_p]),
        "akx_model_solve_steps": (C.c_int, [_vp, _i64]),
        "akx_model_assemble_internal_forces": (C.c_int, [_vp]),
        "akx_model_get_nb_nodes": (C.c_int, [_vp, C.POINTER(_i64)]),
        "akx_model_get_field": (C.c_int, [_vp, C.c_int, _vp]),
        "akx_model_set_field": (C.c_int, [_vp, C.c_int, _vp]),
        "akx_model_compute_all_stresses": (C.c_int, [_vp]),
        "akx_model_get_internal": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "akx_model_synchronize": (C.c_int, [_vp]),
        "akx_model_get_energy": (C.c_int, [_vp, C.c_char_p, C.c_int, _dp]),
        "akx_model_get_energy_noalloc": (C.c_int, [_vp, C.c_char_p, C.c_int, _dp]),
        "akx_model_apply_traction": (C.c_int, [_vp, C.c_int, _i64, _vp, _dp]),
        "akx_model_set_pipeline": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int)]),
        "akx_model_set_block_assembly": (C.c_int, [_vp, C.c_int]),
        "akx_model_get_block_assembly": (C.c_int, [_vp, C.POINTER(C.c_int)]),
        "akx_model_enable_kernel_timing": (C.c_int, [_vp, C.c_int]),
        "akx_model_get_kernel_timing": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(_i64)]),
        "akx_model_get_kernel_bytes": (C.c_int, [_vp, C.c_int, _dp]),
        "akx_calibration_copy": (C.c_int, [C.c_int, _i64, C.c_int, _dp]),
        "akx_batch_create": (C.c_int, [C.c_int, C.c_int, _i64, _i64, _i64, C.c_int,
                                       C.POINTER(_vp)]),
        "akx_batch_destroy": (C.c_int, [_vp]),
        "akx_batch_footprint": (C.c_int, [C.c_int, C.c_int, C.c_int, _i64, _i64,
                                          C.POINTER(_i64)]),
        "akx_batch_add_instance": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _i64, _vp, _i64,
                                             _vp, _vp, _vp, C.c_double,
                                             C.POINTER(C.c_int)]),
        "akx_batch_remove_instance": (C.c_int, [_vp, C.c_int]),
        "akx_batch_get_time_step": (C.c_int, [_vp, C.c_int, _dp, _dp]),
        "akx_batch_set_time_step": (C.c_int, [_vp, C.c_int, C.c_double]),
        "akx_batch_get_steps": (C.c_int, [_vp, C.c_int, C.POINTER(_i64)]),
        "akx_batch_solve_step": (C.c_int, [_vp, _i64]),
        "akx_batch_synchronize": (C.c_int, [_vp]),
        "akx_batch_get_field": (C.c_int, [_vp, C.c_int, C.c_int, _vp]),
        "akx_batch_set_launch": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, C.c_int]),
        "akx_batch_get_launch": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                           C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "akx_batch_autotune": (C.c_int, [_vp, C.c_int, C.POINTER(C.c_int),
                                         C.POINTER(C.c_int), C.POINTER(C.c_int), _dp]),
        "akx_batch_get_stats": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64),
                                          C.POINTER(_i64), C.POINTER(_i64)]),
        "akx_batch_capacity_left": (C.c_int, [_vp, C.POINTER(_i64), C.POINTER(_i64),
                                              C.POINTER(_i64), C.POINTER(_i64)]),
        "akx_batch_enable_kernel_timing": (C.c_int, [_vp, C.c_int]),
        "akx_batch_get_kernel_timing": (C.c_int, [_vp, C.c_int, _dp, C.POINTER(_i64)]),
        "akx_batch_get_kernel_bytes": (C.c_int, [_vp, C.c_int, _dp]),
    }
    for name, (res, args) in sigs.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def lib():
    """The loaded library (raises if it was not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with "
                "`python -c 'import akantu2_amd.build as b; b.build_library()'`")
        lib_ = C.CDLL(LIB_PATH)
        _declare(lib_)
        _lib = lib_
    return _lib


def check(status):
    if status != AKX_OK:
        msg = lib().akx_last_error().decode(errors="replace")
        if status == -3:
            raise JacobianError(status, msg)
        raise AkantuError(status, msg)


def calibration_copy(nbytes: int, repeats: int = 10, device: int = 0) -> float:
    """Average ms of a streaming device copy of nbytes (read nbytes, write
    nbytes)."""
    ms = C.c_double()
    check(lib().akx_calibration_copy(device, int(nbytes), int(repeats), C.byref(ms)))
    return ms.value


def device_count() -> int:
    n = C.c_int(0)
    check(lib().akx_device_count(C.byref(n)))
    return n.value
