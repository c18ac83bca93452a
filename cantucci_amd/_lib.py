"""ctypes binding of the C ABI in include/cantucci_hip.h.

The shared library is built in-tree by ``cantucci_amd/build.py`` (or
``__graft_entry__.build()``).  There is no CPU fallback: if the library is
missing, importing this module raises.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libcantucci_hip.so"

CTH_OK = 0
CTH_E_INVALID = 1
CTH_E_UNSUPPORTED = 2
CTH_E_DEVICE = 3
CTH_E_NOMEM = 4

CTH_SHAPE_MANDELBULB = 0
CTH_SHAPE_MANDELBULB_TRIG = 1
CTH_SHAPE_SPHERE = 2

CTH_MODE_PARITY = 0
CTH_MODE_FAST = 1
CTH_MODE_CONTRACT = 2

CTH_NORMALS_FD6 = 0
CTH_NORMALS_TANGENT = 1
CTH_NORMALS_DEFERRED = 2

CTH_SWEEP_FORWARD = 0
CTH_SWEEP_REVERSE = 1


class CthShape(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("power", C.c_float),
        ("max_iters", C.c_uint32),
        ("bailout", C.c_float),
        ("center", C.c_float * 3),
        ("radius", C.c_float),
    ]


class CthSpan(C.Structure):
    _fields_ = [("start", C.c_float * 3), ("end", C.c_float * 3)]


class CthCamera(C.Structure):
    _fields_ = [
        ("eye", C.c_float * 3),
        ("right", C.c_float * 3),
        ("up", C.c_float * 3),
        ("fwd", C.c_float * 3),
        ("tan_half_fov", C.c_float),
        ("aspect", C.c_float),
    ]


class CthRefCamera(C.Structure):
    """The reference's `Camera` + `Projection` (src/camera.rs:16-20,114-127)."""
    _fields_ = [
        ("position", C.c_float * 3),
        ("direction", C.c_float * 3),
        ("fov", C.c_float),
        ("aspect", C.c_float),
        ("near_plane", C.c_float),
        ("far_plane", C.c_float),
    ]


class CthLaunch(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("chunk", C.c_uint32),
                ("refill_min", C.c_uint32), ("dynamic", C.c_uint32)]


CTH_KERNEL_DE = 0
CTH_KERNEL_TRACE = 1
CTH_KERNEL_ADJOINT = 2
CTH_KERNEL_SPANS = 3
CTH_KERNEL_JOBS = 4
CTH_KERNEL_MESH = 5


class CthJob(C.Structure):
    _fields_ = [("n", C.c_uint32), ("power", C.c_float), ("max_iters", C.c_uint32), ("start", C.c_float * 3),
                ("end", C.c_float * 3), ("divisor", C.c_float)]


class CthError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"cantucci_hip error {code}: {msg}")
        self.code = code


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python cantucci_amd/build.py` "
            "(there is no CPU fallback)")
    # torch ships its own HIP runtime under the same SONAME (libamdhip64.so.7).
    # Loading torch first makes the dynamic linker bind this library to that
    # already-loaded runtime, so the process has one HIP runtime and torch's
    # device buffers and streams are valid for our kernels.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    return C.CDLL(str(LIB_PATH))


lib = _load()

_P = C.c_void_p
_S = C.c_size_t
_U = C.c_uint32
_F = C.c_float
_I = C.c_int
_SH = C.POINTER(CthShape)

_SIGS = {
    "cth_ctx_create": (_I, [_I, C.POINTER(_P)]),
    "cth_ctx_destroy": (None, [_P]),
    "cth_last_error": (C.c_char_p, []),
    "cth_ctx_set_mode": (_I, [_P, _I]),
    "cth_ctx_get_mode": (_I, [_P]),
    "cth_ctx_stream": (_P, [_P]),
    "cth_synchronize": (_I, [_P]),
    "cth_ctx_num_cus": (_I, [_P]),
    "cth_ctx_set_launch": (_I, [_P, _I, C.POINTER(CthLaunch)]),
    "cth_ctx_get_launch": (_I, [_P, _I, C.POINTER(CthLaunch)]),
    "cth_ctx_count_work": (_I, [_P, _P]),
    "cth_bounding_box": (_I, [_SH, C.POINTER(CthSpan)]),
    "cth_min_distance_batch": (_I, [_P, _SH, _P, _S, _P]),
    "cth_max_distance_batch": (_I, [_P, _SH, _P, _S, _P]),
    "cth_bounded_distance_batch": (_I, [_P, _SH, _P, _S, _P]),
    "cth_contains_batch": (_I, [_P, _SH, _P, _S, _P]),
    "cth_min_distance_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P]),
    "cth_de_state_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P]),
    "cth_de_grid_dev": (_I, [_P, _SH, _P, _P, _U, _F, _P, _P, _P]),
    "cth_sample_spans_dev": (_I, [_P, _SH, _P, _S, _U, _P, _P, _P, _P]),
    "cth_refine_dev": (_I, [_P, _SH, _P, _U, _U, _P, _P, _S, C.POINTER(_S), _P, _P]),
    "cth_sample_spans": (_I, [_P, _SH, _P, _S, _U, _P, _P]),
    "cth_sphere_trace_dev": (_I, [_P, _SH, _P, _I, _P, _S, _U, _F, _F, _F, _P, _P, _P, _P]),
    "cth_sphere_trace": (_I, [_P, _SH, _P, _I, _P, _S, _U, _F, _F, _F, _P, _P, _P]),
    "cth_de_vjp_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P, _P]),
    "cth_de_vjp_split_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P, _P, _P]),
    "cth_de_vjp_batch": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P, _I]),
    "cth_de_vjp_noalloc_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P, _P]),
    "cth_normals_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P]),
    "cth_de_grad_dev": (_I, [_P, _SH, _P, _P, _S, _P, _P, _P, _P, _P]),
    "cth_normals_fd_dev": (_I, [_P, _SH, _P, _P, _S, _F, _P, _P, _P, _P]),
    "cth_jobs_create": (_I, [_P, _I, _F, _P, _U, C.POINTER(_P)]),
    "cth_jobs_destroy": (None, [_P]),
    "cth_jobs_total": (_S, [_P]),
    "cth_jobs_offset": (_S, [_P, _U]),
    "cth_jobs_run_dev": (_I, [_P, _P, _P, _P, _P, _P, _P]),
    "cth_wall_clock_khz": (_I, [_P]),
    "cth_graph_begin": (_I, [_P]),
    "cth_graph_end": (_I, [_P, C.POINTER(_P)]),
    "cth_graph_launch": (_I, [_P, _P]),
    "cth_graph_destroy": (None, [_P]),
    "cth_render_depth_dev": (_I, [_P, _SH, _P, _U, _U, _U, _U, _F, _F, _F, _P, _P, _P]),
    "cth_get_focii_dev": (_I, [_P, _SH, _P, _U, _U, _U, _F, _P, _P, _P, _P, _P]),
    "cth_get_focii": (_I, [_P, _SH, C.POINTER(CthRefCamera), _U, _P, C.POINTER(_S)]),
    "cth_mesh_count_dev": (_I, [_P, _P, _S, _U, _P, _P, _P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), _P]),
    "cth_mesh_emit_dev": (_I, [_P, _SH, _P, _P, _S, _U, _I, _P, _P, C.c_uint64, _P, _P, _P]),
    "cth_mesh_vertices_dev": (_I, [_P, _SH, _P, C.c_uint64, _I, _P]),
    "cth_mesh_span": (_I, [_P, _SH, C.POINTER(CthSpan), _U, _I, _P, _S, _P, _S, C.POINTER(_S), C.POINTER(_S)]),
    "cth_mesh_spans": (_I, [_P, _SH, _P, _S, _U, _I, _P, _S, _P, _S, _P, _P, C.POINTER(_S), C.POINTER(_S)]),
    "cth_focus_split": (_I, [_P, _P, _S, _P, _S, _P, _P]),
    "cth_focus_split_dev": (_I, [_P, _P, _S, _P, _P, _S, _P, _P, _P]),
    "cth_split_leaves_dev": (_I, [_P, _P, _S, _P, _P, _S, C.POINTER(_S), _P]),
}


def _bind() -> None:
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name, None)
        if fn is None:
            continue
        fn.restype = res
        fn.argtypes = args


_bind()


def check(rc: int) -> None:
    if rc != CTH_OK:
        raise CthError(rc, (lib.cth_last_error() or b"").decode())


def make_shape(kind: int, power: float = 8.0, max_iters: int = 15, bailout: float = 2.0,
               center=(0.0, 0.0, 0.0), radius: float = 1.0) -> CthShape:
    s = CthShape()
    s.kind = kind
    s.power = power
    s.max_iters = max_iters
    s.bailout = bailout
    s.center[:] = list(center)
    s.radius = radius
    return s
