"""ctypes mirror of include/newt_gpu.h.

The library is loaded from the package's _build directory; a missing
library raises instead of falling back to anything else.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libnewt_gpu.so")

# newt_status
OK = 0
STATUS_NAMES = {
    0: "OK", 1: "INVALID_CONFIG", 2: "UNSUPPORTED_PAIR", 3: "NOT_PD", 4: "NEWTON_FAIL",
    5: "OOM", 6: "DEVICE", 7: "INVALID_ARG", 8: "CAPACITY", 9: "UNSUPPORTED",
}
JOINT_FREE, JOINT_FIXED, JOINT_HINGE, JOINT_SLIDE, JOINT_BALL = range(5)
SHAPE_PLANE, SHAPE_SPHERE, SHAPE_BOX, SHAPE_CAPSULE, SHAPE_MESH = range(5)
INTEGRATOR_EULER, INTEGRATOR_IMPLICITFAST, INTEGRATOR_RK4 = range(3)
SOLVER_PENALTY, SOLVER_PGS, SOLVER_NEWTON = range(3)
CONE_PYRAMIDAL, CONE_ELLIPTIC = range(2)
PRECISION_FP64, PRECISION_MIXED, PRECISION_FP64_EXACT = range(3)
WORLD_PARKED = -1
(ERR_INVALID_CONFIG, ERR_UNSUPPORTED_PAIR, ERR_NOT_PD, ERR_NEWTON_FAIL, ERR_OOM, ERR_DEVICE, ERR_INVALID_ARG,
 ERR_CAPACITY, ERR_UNSUPPORTED) = range(1, 10)

D3 = C.c_double * 3
D4 = C.c_double * 4


class Options(C.Structure):
    _fields_ = [
        ("dt", C.c_double),
        ("gravity", D3),
        ("penetration_slop", C.c_double),
        ("integrator", C.c_int32),
        ("solver_mode", C.c_int32),
        ("iterations", C.c_int32),
        ("cone", C.c_int32),
    ]


class LinkDesc(C.Structure):
    _fields_ = [
        ("parent", C.c_int32),
        ("joint", C.c_int32),
        ("axis", D3),
        ("damping", C.c_double),
        ("armature", C.c_double),
        ("free_damping", C.c_double),
        ("has_range", C.c_int32),
        ("pad0", C.c_int32),
        ("range", C.c_double * 2),
        ("parent_pos", D3),
        ("parent_quat", D4),
        ("child_pos", D3),
        ("child_quat", D4),
        ("mass", C.c_double),
        ("inertia", C.c_double * 9),
        ("limit_custom", C.c_int32),
        ("limit_solimp_power", C.c_int32),
        ("limit_solref", C.c_double * 2),
        ("limit_solimp", C.c_double * 4),
        ("limit_stiffness", C.c_double),
        ("limit_damping", C.c_double),
    ]


class GeomDesc(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("tree", C.c_int32),
        ("link", C.c_int32),
        ("condim", C.c_int32),
        ("size", D3),
        ("pos", D3),
        ("quat", D4),
        ("friction", C.c_double),
        ("margin", C.c_double),
        ("gap", C.c_double),
        ("solref", C.c_double * 2),
        ("solimp", C.c_double * 4),
        ("solimp_power", C.c_int32),
        ("group", C.c_uint32),
        ("mask", C.c_uint32),
        ("body", C.c_int32),
        ("mesh", C.c_int32),
    ]


class BodyDesc(C.Structure):
    _fields_ = [
        ("mass", C.c_double),
        ("inertia", C.c_double * 9),
        ("position", D3),
        ("orientation", D4),
        ("linear_velocity", D3),
        ("angular_velocity_body", D3),
        ("has_max_linear_velocity", C.c_int32),
        ("has_max_angular_velocity", C.c_int32),
        ("max_linear_velocity", C.c_double),
        ("max_angular_velocity", C.c_double),
        ("gravity_scale", C.c_double),
    ]


class SensorDesc(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("tree", C.c_int32),
        ("link", C.c_int32),
        ("body", C.c_int32),
        ("geom", C.c_int32),
        ("offset", D3),
        ("quat", D4),
    ]


# newt_sensor_kind
(SENSOR_JOINT_POS, SENSOR_JOINT_VEL, SENSOR_FRAME_POS, SENSOR_FRAME_QUAT, SENSOR_GYRO, SENSOR_ACCELEROMETER,
 SENSOR_TOUCH, SENSOR_FORCE, SENSOR_TORQUE, SENSOR_VELOCIMETER, SENSOR_MAGNETOMETER, SENSOR_SUBTREE_COM,
 SENSOR_FRAME_LINVEL, SENSOR_FRAME_ANGVEL) = (0, 1, 4, 5, 6, 7, 8, 9, 10, 13, 14, 16, 17, 18)
SENSOR_DIMS = {SENSOR_JOINT_POS: 1, SENSOR_JOINT_VEL: 1, SENSOR_TOUCH: 1, SENSOR_FRAME_QUAT: 4}


def sensor_dim(kind: int) -> int:
    return SENSOR_DIMS.get(kind, 3)


class DeviceState(C.Structure):
    _fields_ = [
        ("qpos", C.c_void_p),
        ("qvel", C.c_void_p),
        ("qacc", C.c_void_p),
        ("qfrc", C.c_void_p),
        ("fat", C.c_void_p),
        ("fault", C.c_void_p),
        ("ncontact", C.c_void_p),
        ("npair", C.c_void_p),
    ]


class BatchOptions(C.Structure):
    _fields_ = [
        ("order", C.c_int32),
        ("wpb_cap", C.c_int32 * 3),
        ("use_graph", C.c_int32),
        ("branches", C.c_int32),
    ]


# Every symbol include/newt_gpu.h declares, with (restype, argtypes).
_P = C.c_void_p
_I = C.c_int32
_PI = C.POINTER(C.c_int32)
_PD = C.POINTER(C.c_double)
SIGNATURES = {
    "newt_last_error": (C.c_char_p, []),
    "newt_model_create": (_I, [C.POINTER(Options), C.POINTER(_P)]),
    "newt_model_add_tree": (_I, [_P, C.POINTER(LinkDesc), _I, _I, _PI]),
    "newt_model_add_geom": (_I, [_P, C.POINTER(GeomDesc), _PI]),
    "newt_model_add_mesh": (_I, [_P, C.POINTER(C.c_double), C.c_int32, C.POINTER(C.c_int32), C.c_int32, _PI]),
    "newt_model_add_body": (_I, [_P, C.POINTER(BodyDesc), _PI]),
    "newt_model_finalize": (_I, [_P]),
    "newt_model_dims": (_I, [_P, _PI, _PI, _PI, _PI]),
    "newt_model_destroy": (None, [_P]),
    "newt_data_create": (_I, [_P, _I, _I, C.POINTER(_P)]),
    "newt_data_destroy": (None, [_P]),
    "newt_data_set_state": (_I, [_P, _PD, _PD]),
    "newt_data_get_state": (_I, [_P, _PD, _PD]),
    "newt_data_set_state_async": (_I, [_P, _PD, _PD, _P]),
    "newt_data_set_worlds_async": (_I, [_P, _I, _I, _PD, _PD, _P]),
    "newt_data_park_worlds": (_I, [_P, _I, _I, _P]),
    "newt_data_get_worlds": (_I, [_P, _I, _I, _PD, _PD]),
    "newt_data_device_bytes": (_I, [_P, C.POINTER(C.c_uint64)]),
    "newt_data_get_qacc": (_I, [_P, _PD]),
    "newt_data_set_qfrc_applied": (_I, [_P, _PD]),
    "newt_data_get_qfrc_applied": (_I, [_P, _PD]),
    "newt_data_get_faults": (_I, [_P, _PI]),
    "newt_data_contact_capacity": (_I, [_P, _PI]),
    "newt_data_get_contacts": (_I, [_P, _PI, _PI, _PD]),
    "newt_data_pair_capacity": (_I, [_P, _PI]),
    "newt_data_get_pairs": (_I, [_P, _PI, _PI]),
    "newt_step": (_I, [_P, _P, _I, _P]),
    "newt_data_synchronize": (_I, [_P]),
    "newt_model_set_contact_capacity": (_I, [_P, _I]),
    "newt_model_from_json": (_I, [C.c_char_p, _I, C.POINTER(_P)]),
    "newt_model_counts": (_I, [_P, _PI, _PI, _PI]),
    "newt_model_get_options": (_I, [_P, C.POINTER(Options)]),
    "newt_model_get_tree": (_I, [_P, _I, C.POINTER(LinkDesc), _I, _PI, _PI]),
    "newt_model_get_geom": (_I, [_P, _I, C.POINTER(GeomDesc)]),
    "newt_model_get_body": (_I, [_P, _I, C.POINTER(BodyDesc)]),
    "newt_data_reset": (_I, [_P, C.POINTER(C.c_uint8), _P]),
    "newt_data_reset_to_keyframe": (_I, [_P, C.POINTER(C.c_uint8), _PD, _PD, _P]),
    "newt_data_snapshot": (_I, [_P, C.POINTER(_P)]),
    "newt_data_restore": (_I, [_P, _P, _P]),
    "newt_snapshot_destroy": (None, [_P]),
    "newt_batch_create": (_I, [C.POINTER(_P), _I, C.POINTER(BatchOptions), C.POINTER(_P)]),
    "newt_batch_step": (_I, [_P, _I, _P]),
    "newt_batch_launches": (_I, [_P, _PI]),
    "newt_batch_destroy": (None, [_P]),
    "newt_step_noalloc": (_I, [_P, C.POINTER(DeviceState), _I, _I, _I, _I, _P]),
    "newt_noalloc_max_worlds": (_I, [_P, _I, _PI]),
    "newt_model_add_sensor": (_I, [_P, C.POINTER(SensorDesc), _PI]),
    "newt_model_set_magnetic_field": (_I, [_P, _PD]),
    "newt_model_sensor_dim": (_I, [_P, _PI]),
    "newt_data_get_sensordata": (_I, [_P, _PD]),
}

_lib = None


def load_library(path: str = LIB_PATH):
    """Load libnewt_gpu.so and declare its signatures; raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(
            f"{path} is missing: build the HIP extension first (python -m tooling_amd.build)")
    # The library links the system HIP runtime (libamdhip64.so.7);
    # tooling_amd.hip drives streams and events through that same runtime.
    # A process that also uses torch must import torch first: torch ships its
    # own runtime under that soname, which then serves both.
    lib = C.CDLL(path)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


class NewtError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


def check(status: int):
    if status != OK:
        msg = load_library().newt_last_error()
        raise NewtError(status, msg.decode() if msg else "")
