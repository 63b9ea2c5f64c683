"""ctypes binding of lib/libflowvpm_hip.so (declared in include/flowvpm_hip.h).

The library is required: there is no CPU fallback.  If it is missing or
fails to load, every call raises.
"""
import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libflowvpm_hip.so")
HEADER_PATH = os.path.join(os.path.dirname(HERE), "include", "flowvpm_hip.h")

VPM_OK, VPM_EINVAL, VPM_ERANGE, VPM_ENOMEM, VPM_EDEVICE, VPM_ENAN, VPM_ERESET, VPM_ECONVERGE = \
    0, -1, -2, -3, -4, -5, -6, -7
STATUS_NAMES = {0: "VPM_OK", -1: "VPM_EINVAL", -2: "VPM_ERANGE", -3: "VPM_ENOMEM",
                -4: "VPM_EDEVICE", -5: "VPM_ENAN", -6: "VPM_ERESET", -7: "VPM_ECONVERGE"}
KERNELS = {"singular": 0, "gaussian": 1, "gaussianerf": 2, "winckelmans": 3}
PRECISIONS = {"f32c": 0, "f64": 1}
UJ_RESET, UJ_RESET_SFS, UJ_SFS, UJ_NO_SORT, UJ_F64_FAR = 1, 2, 4, 8, 16


class VPMError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


class Config(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("precision", C.c_int32),
                ("transposed", C.c_int32), ("reserved", C.c_int32)]


class StepConfig(C.Structure):
    _fields_ = [("uj", Config), ("f", C.c_double), ("g", C.c_double), ("Uinf", C.c_double * 3),
                ("sfs", C.c_int32), ("sfs_force_positive", C.c_int32), ("sfs_clip", C.c_int32),
                ("relax", C.c_int32),
                ("sfs_alpha", C.c_double), ("sfs_rlxf", C.c_double), ("sfs_minC", C.c_double),
                ("sfs_maxC", C.c_double), ("rlxf", C.c_double),
                ("viscous", C.c_int32), ("uj_flags", C.c_int32),
                ("nu", C.c_double), ("sgm0", C.c_double), ("beta", C.c_double), ("t_sgm", C.c_double),
                ("rbf_itmax", C.c_int32), ("rbf_iterror", C.c_int32), ("rbf_tol", C.c_double)]


_dptr = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_SIGS = {
    "vpm_version": (C.c_char_p, []),
    "vpm_last_error": (C.c_char_p, []),
    "vpm_device_count": (C.c_int, [C.POINTER(C.c_int)]),
    "vpm_create": (C.c_int, [C.c_int, C.c_int64, C.POINTER(C.c_void_p)]),
    "vpm_destroy": (C.c_int, [C.c_void_p]),
    "vpm_set_particles": (C.c_int, [C.c_void_p, _dptr, C.c_int64]),
    "vpm_get_particles": (C.c_int, [C.c_void_p, _dptr, C.c_int64, C.c_int32, C.c_int32]),
    "vpm_add_particles": (C.c_int, [C.c_void_p, _dptr, C.c_int64]),
    "vpm_remove_particle": (C.c_int, [C.c_void_p, C.c_int64]),
    "vpm_get_np": (C.c_int, [C.c_void_p, _i64p]),
    "vpm_synchronize": (C.c_int, [C.c_void_p]),
    "vpm_uj_direct": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_uint32]),
    "vpm_uj_direct_noalloc": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_uint32]),
    "vpm_uj_direct_fields": (C.c_int, [C.c_void_p, C.c_void_p, C.POINTER(Config)]),
    "vpm_p2p_ranges": (C.c_int, [C.c_void_p, C.POINTER(Config), _i64p, _i64p, _i64p, _i64p, C.c_int64]),
    "vpm_plan_ranges": (C.c_int, [C.c_void_p, _i64p, _i64p, _i64p, _i64p, C.c_int64, C.POINTER(C.c_void_p)]),
    "vpm_plan_destroy": (C.c_int, [C.c_void_p]),
    "vpm_p2p_plan": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_void_p]),
    "vpm_estr_plan": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_void_p]),
    "vpm_zeta_plan": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_void_p]),
    "vpm_estr_direct": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "vpm_reset_particles": (C.c_int, [C.c_void_p]),
    "vpm_reset_particles_sfs": (C.c_int, [C.c_void_p]),
    "vpm_update_particle_states": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, _dptr,
                                             C.c_double, C.c_double, C.c_double, C.c_int32]),
    "vpm_relax": (C.c_int, [C.c_void_p, C.c_double, C.c_int32]),
    "vpm_core_spreading": (C.c_int, [C.c_void_p, C.c_double, C.c_double, C.c_double, C.c_double]),
    "vpm_sfs_dynamic_before": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_double, C.c_double, C.c_double]),
    "vpm_sfs_dynamic_after": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_int32, C.c_int32]),
    "vpm_rk3_step": (C.c_int, [C.c_void_p, C.POINTER(StepConfig), C.c_double, C.c_int32]),
    "vpm_uj_adjoint": (C.c_int, [C.c_void_p, C.POINTER(Config), _dptr, _dptr, _dptr, _dptr, _dptr]),
    "vpm_estr_ranges": (C.c_int, [C.c_void_p, C.POINTER(Config), _i64p, _i64p, _i64p, _i64p, C.c_int64]),
    "vpm_zeta_direct": (C.c_int, [C.c_void_p, C.POINTER(Config)]),
    "vpm_zeta_ranges": (C.c_int, [C.c_void_p, C.POINTER(Config), _i64p, _i64p, _i64p, _i64p, C.c_int64]),
    "vpm_core_spreading_reset": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_double, C.c_int32, C.c_double,
                                           C.c_int32, C.POINTER(C.c_int32), _dptr]),
    "vpm_rbf_conjugategradient": (C.c_int, [C.c_void_p, C.POINTER(Config), C.c_int32, C.c_double, C.c_int32,
                                            C.POINTER(C.c_int32), _dptr]),
    "vpm_tape_create": (C.c_int, [C.c_void_p, C.POINTER(C.c_void_p)]),
    "vpm_tape_destroy": (C.c_int, [C.c_void_p]),
    "vpm_rk3_step_taped": (C.c_int, [C.c_void_p, C.POINTER(StepConfig), C.c_double, C.c_void_p]),
    "vpm_rk3_step_adjoint": (C.c_int, [C.c_void_p, _dptr]),
    "vpm_profile": (C.c_int, [C.c_void_p, C.c_int32]),
    "vpm_timing": (C.c_int, [C.c_void_p, C.c_int32, _dptr, _i64p]),
}

_lib = None


def header_symbols():
    """Function names declared in include/flowvpm_hip.h."""
    with open(HEADER_PATH) as fh:
        text = fh.read()
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\*?\s+\*?(vpm_\w+)\s*\(", text, re.M)))


def lib():
    """Load the HIP library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise OSError("libflowvpm_hip.so is not built (%s); run __graft_entry__.build()" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(status):
    if status != VPM_OK:
        raise VPMError(status, lib().vpm_last_error().decode())
    return status


def device_count():
    n = C.c_int(0)
    rc = lib().vpm_device_count(C.byref(n))
    return n.value if rc == VPM_OK else 0
