"""ctypes mirror of include/snac_hip.h (the C ABI of libsnac_hip.so).

Struct layouts and constants here must match the header field for field;
tests/test_abi.py checks the sizes against the compiled library.
"""
import ctypes as C
import os

SNH_OK = 0
SNH_EINVAL = -1
SNH_ENOMEM = -2
SNH_EHIP = -3
SNH_ESTATE = -4
SNH_EUNSUPPORTED = -5
SNH_ENUMERIC = -6

SNH_TETRA_COUNT = 10
SNH_MAX_SEGMENTS = 8
SNH_MAX_MATERIALS = 16

RHEOLOGY_ELASTIC = 0x1
RHEOLOGY_PLASTIC = 0x2
RHEOLOGY_MAXWELL = 0x4
RHEOLOGY_VISCOPLASTIC = 0x8

DT_CONSTANT = 0
DT_DYNAMIC = 1
DT_WAVE = 2

WALL_FRONT, WALL_BACK, WALL_LEFT, WALL_RIGHT, WALL_TOP, WALL_BOTTOM, WALL_ALL = range(7)
WALL_NAMES = {"front": WALL_FRONT, "back": WALL_BACK, "left": WALL_LEFT, "right": WALL_RIGHT,
              "top": WALL_TOP, "bottom": WALL_BOTTOM, "all": WALL_ALL}
VX, VY, VZ = 0x1, 0x2, 0x4

PRECISION_FP64 = 0
PRECISION_MIXED = 1

FIELD_REFVISC = 0
FIELD_COHESION_OFFSET = 1

_SEG = SNH_MAX_SEGMENTS + 1


class Material(C.Structure):
    _fields_ = [
        ("rheology", C.c_uint32),
        ("phsDensity", C.c_double),
        ("alpha", C.c_double),
        ("beta", C.c_double),
        ("lambda_", C.c_double),
        ("mu", C.c_double),
        ("nsegments", C.c_uint32),
        ("plstrain", C.c_double * _SEG),
        ("frictionAngle", C.c_double * _SEG),
        ("dilationAngle", C.c_double * _SEG),
        ("cohesion", C.c_double * _SEG),
        ("ten_off", C.c_double),
        ("vis_min", C.c_double),
        ("vis_max", C.c_double),
        ("refvisc", C.c_double),
        ("refsrate", C.c_double),
        ("reftemp", C.c_double),
        ("activationE", C.c_double),
        ("srexponent", C.c_double),
        ("thermal_conduct", C.c_double),
        ("heatCapacity", C.c_double),
    ]


class VelocityBC(C.Structure):
    _fields_ = [("wall", C.c_int32), ("mask", C.c_uint32), ("value", C.c_double * 3)]


class Config(C.Structure):
    _fields_ = [
        ("meshSizeI", C.c_uint32),
        ("meshSizeJ", C.c_uint32),
        ("meshSizeK", C.c_uint32),
        ("minX", C.c_double),
        ("minY", C.c_double),
        ("minZ", C.c_double),
        ("maxX", C.c_double),
        ("maxY", C.c_double),
        ("maxZ", C.c_double),
        ("dtType", C.c_int32),
        ("timeStep", C.c_double),
        ("gravity", C.c_double),
        ("demf", C.c_double),
        ("strain_inert", C.c_double),
        ("topo_kappa", C.c_double),
        ("topoGradCriterion", C.c_double),
        ("nmaterials", C.c_uint32),
        ("materials", C.POINTER(Material)),
        ("nbcs", C.c_uint32),
        ("bcs", C.POINTER(VelocityBC)),
        ("precision", C.c_int32),
    ]


class State(C.Structure):
    _fields_ = [
        ("coord", C.POINTER(C.c_double)),
        ("velocity", C.POINTER(C.c_double)),
        ("stress", C.POINTER(C.c_double)),
        ("plastic_strain", C.POINTER(C.c_double)),
        ("viscosity", C.POINTER(C.c_double)),
        ("volume", C.POINTER(C.c_double)),
        ("temperature", C.POINTER(C.c_double)),
    ]


class TemperatureBC(C.Structure):
    _fields_ = [("wall", C.c_int32), ("value", C.c_double)]


class Thermal(C.Structure):
    _fields_ = [
        ("computeThermalStress", C.c_int32),
        ("nbcs", C.c_uint32),
        ("bcs", C.POINTER(TemperatureBC)),
        ("temperature", C.POINTER(C.c_double)),
    ]


class Scalars(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("speedOfSound", C.c_double),
        ("minLengthScale", C.c_double),
        ("vmax", C.c_double),
        ("timeStep", C.c_uint32),
        ("topoGradMax", C.c_double),
    ]


class KernelTimes(C.Structure):
    _fields_ = [("elem_ms", C.c_double), ("node_ms", C.c_double), ("geo_ms", C.c_double), ("fin_ms", C.c_double),
                ("elem_n", C.c_uint32), ("node_n", C.c_uint32), ("geo_n", C.c_uint32), ("fin_n", C.c_uint32),
                ("heat_elem_ms", C.c_double), ("heat_node_ms", C.c_double),
                ("heat_elem_n", C.c_uint32), ("heat_node_n", C.c_uint32)]


ELEM_TET_PER_LANE = 0
ELEM_ELEMENT_PER_LANE = 1
ELEM_ELEMENT_PER_LANE_LOWREG = 2
ELEM_TET_PER_WAVE = 3
ELEM_VARIANTS = 4
NODE_PER_NODE = 0      # SNH_NODE_* of include/snac_hip.h
NODE_TILED = 1


NODE_TUNES = 4          # SNH_NODE_TUNES


class TuneReport(C.Structure):
    _fields_ = [("elem_ms", C.c_double * ELEM_VARIANTS), ("elem_variant", C.c_int32),
                ("node_ms", C.c_double * NODE_TUNES), ("node_variant", C.c_int32), ("node_kc", C.c_int32)]


class Diagnostics(C.Structure):
    _fields_ = [("strain_rate", C.POINTER(C.c_double)), ("stress", C.POINTER(C.c_double)),
                ("hydro_pressure", C.POINTER(C.c_double)), ("plastic_strain", C.POINTER(C.c_double)),
                ("force", C.POINTER(C.c_double))]


PARAM_LOG_VISCOSITY = 0
PARAM_COHESION = 1


ADJ_CORNER_CLAMP = 0x1


class AdjointConfig(C.Structure):
    _fields_ = [("param", C.c_int32), ("checkpoint_interval", C.c_int32), ("flags", C.c_int32),
                ("snapshots", C.c_int32)]


class AdjointTimes(C.Structure):
    _fields_ = [("forward_ms", C.c_double), ("replay_ms", C.c_double), ("reverse_ms", C.c_double),
                ("steps", C.c_int32), ("checkpoints", C.c_int32), ("replay_steps", C.c_int64)]


# Every function the header declares, with its ctypes signature.
_P = C.c_void_p
_I = C.c_int
SIGNATURES = {
    "snh_abi_version": (_I, []),
    "snh_last_error": (C.c_char_p, []),
    "snh_context_create": (_I, [C.POINTER(Config), C.POINTER(_P)]),
    "snh_context_destroy": (_I, [_P]),
    "snh_set_stream": (_I, [_P, _P]),
    "snh_set_element_material": (_I, [_P, C.POINTER(C.c_uint8), C.c_size_t]),
    "snh_set_element_field": (_I, [_P, _I, C.POINTER(C.c_double), C.c_size_t]),
    "snh_upload_state": (_I, [_P, C.POINTER(State)]),
    "snh_download_state": (_I, [_P, C.POINTER(State)]),
    "snh_hydrostatic_ic": (_I, [_P]),
    "snh_timestep_zero": (_I, [_P]),
    "snh_set_thermal": (_I, [_P, C.POINTER(Thermal)]),
    "snh_remesh": (_I, [_P, C.POINTER(State)]),
    "snh_loop_elements_energy": (_I, [_P]),
    "snh_loop_nodes_energy": (_I, [_P]),
    "snh_calc_stresses": (_I, [_P]),
    "snh_loop_nodes_momentum": (_I, [_P]),
    "snh_loop_elements_momentum": (_I, [_P]),
    "snh_sync": (_I, [_P]),
    "snh_step": (_I, [_P, _I]),
    "snh_get_scalars": (_I, [_P, C.POINTER(Scalars)]),
    "snh_synchronize": (_I, [_P]),
    "snh_profile": (_I, [_P, _I]),
    "snh_autotune": (_I, [_P, C.POINTER(TuneReport)]),
    "snh_set_elem_variant": (_I, [_P, C.c_int32]),
    "snh_set_node_variant": (_I, [_P, C.c_int32]),
    "snh_sync_state": (_I, [C.POINTER(Config), _P, _P, _P, _P]),
    "snh_measure_copy_bandwidth": (_I, [C.c_size_t, _I, C.POINTER(C.c_double)]),
    "snh_set_diagnostics": (_I, [_P, _I]),
    "snh_get_diagnostics": (_I, [_P, C.POINTER(Diagnostics)]),
    "snh_set_schedule": (_I, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                              C.POINTER(C.c_int8)]),
    "snh_record_schedule": (_I, [_P, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                 C.POINTER(C.c_int8)]),
    "snh_ensemble_create": (_I, [C.POINTER(_P), C.c_int32, C.c_int32, C.POINTER(_P)]),
    "snh_ensemble_step": (_I, [_P, C.c_int32]),
    "snh_ensemble_batched": (_I, [_P]),
    "snh_ensemble_synchronize": (_I, [_P]),
    "snh_ensemble_destroy": (_I, [_P]),
    "snh_adjoint_begin": (_I, [_P, C.POINTER(AdjointConfig)]),
    "snh_adjoint_forward": (_I, [_P, C.c_int32]),
    "snh_adjoint_reverse": (_I, [_P, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "snh_adjoint_get_times": (_I, [_P, C.POINTER(AdjointTimes)]),
    "snh_adjoint_end": (_I, [_P]),
    "snh_get_kernel_times": (_I, [_P, C.POINTER(KernelTimes)]),
}

ABI_VERSION = 6          # SNH_ABI_VERSION of include/snac_hip.h

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libsnac_hip.so")

_lib = None


def load_library(path=LIB_PATH):
    """Load libsnac_hip.so (the HIP path).  Raises if it is missing: there is
    no CPU fallback in the product."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"libsnac_hip.so not found at {path}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    lib = C.CDLL(path)
    lib.snh_abi_version.restype = C.c_int
    if lib.snh_abi_version() != ABI_VERSION:
        raise RuntimeError(f"{path} has ABI version {lib.snh_abi_version()}, this binding expects {ABI_VERSION}; "
                           "rebuild it with `python -c 'import __graft_entry__ as g; g.build()'`")
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class SnacHipError(RuntimeError):
    pass


def check(rc, lib=None):
    if rc != SNH_OK:
        lib = lib or _lib
        msg = lib.snh_last_error().decode() if lib is not None else ""
        raise SnacHipError(f"snac_hip error {rc}: {msg}")
    return rc
