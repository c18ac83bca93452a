"""ctypes binding of oracle/liboracle.so (test infrastructure only).

Each wrapper names the reference function it restates; the C++ source
(oracle.cpp) cites file:line for every operation.
"""
from __future__ import annotations

import contextlib
import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "liboracle.so"

K_MANDELBULB = 0
K_MANDELBULB_TRIG = 1
K_SPHERE = 2

__all__ = [
    "OrcShape", "mandelbulb", "mandelbulb_trig", "sphere", "de_points", "de_points_f64", "de_grad_f64", "de_grad_split_f64",
    "de_grid", "sample_spans", "surface_nets", "mesh_batch", "focus_split", "create_spans", "refine", "ray_dirs", "focus_dirs", "get_focii", "sphere_trace", "fd_diff",
    "powi", "build", "K_MANDELBULB", "K_MANDELBULB_TRIG", "K_SPHERE", "default_threads", "ln_mode",
]


class OrcShape(C.Structure):
    _fields_ = [
        ("kind", C.c_int32),
        ("power", C.c_float),
        ("max_iters", C.c_uint32),
        ("bailout", C.c_float),
        ("center", C.c_float * 3),
        ("radius", C.c_float),
    ]


def build() -> Path:
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def _load() -> C.CDLL:
    if not LIB_PATH.exists():
        build()
    return C.CDLL(str(LIB_PATH))


_lib = _load()
_P = C.c_void_p
_S = C.c_size_t
_U = C.c_uint32
_F = C.c_float
_I = C.c_int
_SH = C.POINTER(OrcShape)
for _name, _res, _args in [
    ("orc_shape_size", _I, []),
    ("orc_de_points_f32", None, [_SH, _P, _P, _S, _P, _P, _P, _P, _I]),
    ("orc_de_points_f64", None, [_SH, _P, _P, _S, _P, _P, _I]),
    ("orc_de_grad_f64", None, [_SH, _P, _P, _S, _P, _P, _P, _I]),
    ("orc_de_grad_split_f64", None, [_SH, _P, _P, _S, _P, _P, _P, _I]),
    ("orc_de_grid_f32", None, [_SH, _P, _P, _U, _F, _P, _P, _I]),
    ("orc_sample_spans", None, [_SH, _P, _S, _U, _P, _P, _P, _I]),
    ("orc_surface_nets", None, [_SH, _P, _U, _P, _P, _S, _P, _S, C.POINTER(_S), C.POINTER(_S)]),
    ("orc_mesh_batch", None, [_SH, _P, _S, _U, _P, _I, _P, _P, _P, _S, _P, _S]),
    ("orc_create_spans", None, [_P, _P]),
    ("orc_focus_split", None, [_P, _S, _P, _P, _S, _P, _P]),
    ("orc_refine", _S, [_SH, _P, _U, _U, _P, _P, _S, _P, _I]),
    ("orc_ray_dirs", None, [_P, _U, _U, _U, _U, _P]),
    ("orc_focus_dirs", None, [_P, _U, _P]),
    ("orc_sphere_trace", None, [_SH, _P, _P, _S, _I, _U, _F, _F, _F, _P, _P, _P, _I]),
    ("orc_sphere_trace_marked", None, [_SH, _P, _P, _S, _I, _U, _F, _F, _F, _P, _P, _P, _P, _I]),
    ("orc_set_ln_mode", None, [_I]),
    ("orc_fd_diff", None, [_SH, _P, _P, _S, _F, _P, _I]),
    ("orc_powi_f32", _F, [_F, _I]),
]:
    _fn = getattr(_lib, _name)
    _fn.restype = _res
    _fn.argtypes = _args


def default_threads() -> int:
    return int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))


def _shape(kind, power=8.0, max_iters=15, bailout=2.0, center=(0.0, 0.0, 0.0), radius=1.0) -> OrcShape:
    s = OrcShape()
    s.kind = kind
    s.power = power
    s.max_iters = max_iters
    s.bailout = bailout
    s.center[:] = list(center)
    s.radius = radius
    return s


def mandelbulb(max_iters=15, bailout=2.0, power=8) -> OrcShape:
    """`Mandelbulb<P>::new(max_iters, bailout)` (src/shape/mandelbulb.rs:19-26)."""
    return _shape(K_MANDELBULB, float(power), max_iters, bailout)


def mandelbulb_trig(max_iters=15, bailout=2.0, power=8.0) -> OrcShape:
    """The trig map (src/shape/mandelbulb.rs:128-146) with a real power."""
    return _shape(K_MANDELBULB_TRIG, float(power), max_iters, bailout)


def sphere(center=(0.0, 0.0, 0.0), radius=1.0) -> OrcShape:
    """`Sphere::new` (src/shape/sphere.rs:13-19)."""
    return _shape(K_SPHERE, 0.0, 1, 0.0, center, radius)


def _f32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32))


def _ptr(a):
    return None if a is None else a.ctypes.data


def de_points(shape: OrcShape, pts, powers=None, threads=None):
    """`min_distance_from` over points (src/shape/mandelbulb.rs:59-79).
    Returns (d, iters, r, dr)."""
    pts = _f32(pts).reshape(-1, 3)
    n = len(pts)
    pw = None if powers is None else _f32(powers)
    d = np.empty(n, np.float32)
    it = np.empty(n, np.uint32)
    r = np.empty(n, np.float32)
    dr = np.empty(n, np.float32)
    _lib.orc_de_points_f32(C.byref(shape), _ptr(pts), _ptr(pw), n, _ptr(d), _ptr(it), _ptr(r), _ptr(dr),
                           threads or default_threads())
    return d, it, r, dr


def de_points_f64(shape: OrcShape, pts, powers=None, threads=None):
    """The same map in FP64 (accuracy oracle).  Returns (d, iters)."""
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64)).reshape(-1, 3)
    n = len(pts)
    pw = None if powers is None else _f32(powers)
    d = np.empty(n, np.float64)
    it = np.empty(n, np.uint32)
    _lib.orc_de_points_f64(C.byref(shape), _ptr(pts), _ptr(pw), n, _ptr(d), _ptr(it), threads or default_threads())
    return d, it


def de_grad_f64(shape: OrcShape, pts, powers=None, threads=None):
    """FP64 forward-mode dual numbers: (d, grad[n,4] = dd/dp xyz, dd/dpower, iters)."""
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64)).reshape(-1, 3)
    n = len(pts)
    pw = None if powers is None else _f32(powers)
    d = np.empty(n, np.float64)
    g = np.empty((n, 4), np.float64)
    it = np.empty(n, np.uint32)
    _lib.orc_de_grad_f64(C.byref(shape), _ptr(pts), _ptr(pw), n, _ptr(d), _ptr(g), _ptr(it),
                         threads or default_threads())
    return d, g, it


def de_grad_split_f64(shape: OrcShape, pts, powers=None, threads=None):
    """FP64 dual numbers with the point's two roles apart:
    (d, grad[n,7] = dd/dz0 xyz, dd/dpower, dd/dc xyz, iters); dd/dp = dd/dz0 + dd/dc."""
    pts = np.ascontiguousarray(np.asarray(pts, dtype=np.float64)).reshape(-1, 3)
    n = len(pts)
    pw = None if powers is None else _f32(powers)
    d = np.empty(n, np.float64)
    g = np.empty((n, 7), np.float64)
    it = np.empty(n, np.uint32)
    _lib.orc_de_grad_split_f64(C.byref(shape), _ptr(pts), _ptr(pw), n, _ptr(d), _ptr(g), _ptr(it),
                               threads or default_threads())
    return d, g, it


def de_grid(shape: OrcShape, start, end, n: int, divisor: float, threads=None):
    """Corner-formula grid (src/mesh/buffer.rs:78-83) in GridTable order."""
    s = _f32(start)
    e = _f32(end)
    d = np.empty(n ** 3, np.float32)
    it = np.empty(n ** 3, np.uint32)
    _lib.orc_de_grid_f32(C.byref(shape), _ptr(s), _ptr(e), n, float(divisor), _ptr(d), _ptr(it),
                         threads or default_threads())
    return d, it


def sample_spans(shape: OrcShape, spans, res: int, threads=None):
    """Naive surface nets step 1 + sign classification per span
    (src/mesh/buffer.rs:63-147).  Returns (dists[m,(res+1)^3], cross[m,words], iters)."""
    sp = _f32(spans).reshape(-1, 6)
    m = len(sp)
    per = (res + 1) ** 3
    words = (res ** 3 + 63) // 64
    d = np.empty((m, per), np.float32)
    cb = np.empty((m, words), np.uint64)
    it = np.empty((m, per), np.uint32)
    _lib.orc_sample_spans(C.byref(shape), _ptr(sp), m, res, _ptr(d), _ptr(cb), _ptr(it), threads or default_threads())
    return d, cb, it


def surface_nets(shape: OrcShape, span, res: int):
    """The whole `naive_surface_nets` (src/mesh/buffer.rs:58-391).
    Returns (dists, vertices[nv,7], indices)."""
    sp = _f32(span).reshape(6)
    nv = C.c_size_t()
    ni = C.c_size_t()
    d = np.empty((res + 1) ** 3, np.float32)
    _lib.orc_surface_nets(C.byref(shape), _ptr(sp), res, _ptr(d), None, 0, None, 0, C.byref(nv), C.byref(ni))
    v = np.empty((nv.value, 7), np.float32)
    ix = np.empty(ni.value, np.uint32)
    _lib.orc_surface_nets(C.byref(shape), _ptr(sp), res, _ptr(d), _ptr(v), nv.value, _ptr(ix), ni.value,
                          C.byref(nv), C.byref(ni))
    return d, v, ix


def mesh_batch(shape: OrcShape, spans, res: int, dists=None, threads=None):
    """Surface nets steps 2-3 (or the whole algorithm when `dists` is None)
    over a batch of spans.  Returns (voff[n+1], ioff[n+1], vertices[nv,7],
    indices) with span-local u32 indices, concatenated in span order."""
    sp = _f32(spans).reshape(-1, 6)
    n = sp.shape[0]
    d = None if dists is None else _f32(dists).reshape(n, (res + 1) ** 3)
    voff = np.zeros(n + 1, np.uint64)
    ioff = np.zeros(n + 1, np.uint64)
    t = threads or default_threads()
    _lib.orc_mesh_batch(C.byref(shape), _ptr(sp), n, res, None if d is None else _ptr(d), t, _ptr(voff),
                        _ptr(ioff), None, 0, None, 0)
    v = np.empty((int(voff[-1]), 7), np.float32)
    ix = np.empty(int(ioff[-1]), np.uint32)
    _lib.orc_mesh_batch(C.byref(shape), _ptr(sp), n, res, None if d is None else _ptr(d), t, _ptr(voff),
                        _ptr(ioff), _ptr(v), v.shape[0], _ptr(ix), ix.shape[0])
    return voff, ioff, v, ix


def focus_split(leaves, focii, eye, depth=None) -> np.ndarray:
    """ShapeMesh::update's split flags (src/mesh/mod.rs:91-104) for a flat leaf set."""
    lv = _f32(leaves).reshape(-1, 6)
    fc = _f32(focii).reshape(-1, 3)
    dp = None if depth is None else _f32(depth).reshape(-1)
    e = _f32(eye).reshape(3)
    flags = np.zeros(lv.shape[0], np.uint32)
    _lib.orc_focus_split(_ptr(lv), lv.shape[0], _ptr(fc), None if dp is None else _ptr(dp), fc.shape[0], _ptr(e),
                         _ptr(flags))
    return flags


def create_spans(span) -> np.ndarray:
    """`create_spans` (src/octree/mod.rs:315-329): 8 child spans."""
    sp = _f32(span).reshape(6)
    out = np.empty((8, 6), np.float32)
    _lib.orc_create_spans(_ptr(sp), _ptr(out))
    return out


def refine(shape: OrcShape, root, min_depth: int, max_depth: int, half_diag, threads=None, cap=None):
    """Level-synchronous refinement; returns (leaves[m,6], level_counts)."""
    r = _f32(root).reshape(6)
    hd = _f32(half_diag)
    lc = np.zeros(max_depth + 1, np.uint64)
    total = _lib.orc_refine(C.byref(shape), _ptr(r), min_depth, max_depth, _ptr(hd), None, 0, _ptr(lc),
                            threads or default_threads())
    leaves = np.empty((total, 6), np.float32)
    _lib.orc_refine(C.byref(shape), _ptr(r), min_depth, max_depth, _ptr(hd), _ptr(leaves), total, _ptr(lc),
                    threads or default_threads())
    return leaves, lc


def ray_dirs(pose, W: int, H: int, row0: int = 0, rows: int | None = None) -> np.ndarray:
    """Pinhole primary-ray directions for rows [row0, row0+rows)."""
    p = _f32(pose).reshape(14)
    rows = H - row0 if rows is None else rows
    out = np.empty((rows * W, 3), np.float32)
    _lib.orc_ray_dirs(_ptr(p), W, H, row0, rows, _ptr(out))
    return out


def focus_dirs(cam, focus_points: int = 5) -> np.ndarray:
    """The focus-ray directions of `ShapeMesh::get_focii` (src/mesh/mod.rs:205-231)
    for cam = (position xyz, unit direction xyz, fov, aspect, near, far)
    (src/camera.rs:16-20,114-127); ray x * fp + y of iter::square."""
    c = _f32(cam).reshape(10)
    out = np.empty((focus_points * focus_points, 3), np.float32)
    _lib.orc_focus_dirs(_ptr(c), focus_points, _ptr(out))
    return out


def get_focii(shape: OrcShape, cam, focus_points: int = 5, max_steps: int = 100, eps: float = 1e-6):
    """`ShapeMesh::get_focii` (src/mesh/mod.rs:205-243) without the filter:
    (positions after the last step (fp^2, 3), depth (+inf on a miss), steps,
    directions).  The reference returns positions[np.isfinite(depth)]."""
    c = _f32(cam).reshape(10)
    dirs = focus_dirs(c, focus_points)
    depth, steps, pos = sphere_trace(shape, c[:3][None], dirs, max_steps, eps_rel=0.0, t_max=float("inf"),
                                     eps_abs=eps, with_pos=True)
    return pos, depth, steps, dirs


def sphere_trace(shape: OrcShape, origins, dirs, max_steps=256, eps_rel=1e-4, t_max=10.0, eps_abs=0.0,
                 with_pos=False, threads=None, marked=False):
    """Sphere tracing (src/mesh/mod.rs:229-241; hit when d < eps_abs + eps_rel * t).
    With marked=True also returns, last, a bool per ray: its march took the
    final ln of a hard case (only under ln_mode("cr"))."""
    o = _f32(origins).reshape(-1, 3)
    dd = _f32(dirs).reshape(-1, 3)
    n = len(dd)
    stride = 0 if len(o) == 1 else 1
    depth = np.empty(n, np.float32)
    steps = np.empty(n, np.uint32)
    pos = np.empty((n, 3), np.float32) if with_pos else None
    hard = np.zeros(n, np.uint8) if marked else None
    _lib.orc_sphere_trace_marked(C.byref(shape), _ptr(o), _ptr(dd), n, stride, max_steps, float(eps_abs),
                                 float(eps_rel), float(t_max), _ptr(depth), _ptr(steps), _ptr(pos), _ptr(hard),
                                 threads or default_threads())
    out = (depth, steps, pos) if with_pos else (depth, steps)
    return out + (hard.astype(bool),) if marked else out


@contextlib.contextmanager
def ln_mode(mode: str):
    """Within the block the final formula's ln is glibc's logf ("glibc", the
    reference and the default) or the f64 log rounded to f32 ("cr": the
    correctly rounded f32 ln outside hard cases, which is what the device
    computes)."""
    if mode not in ("glibc", "cr"):
        raise ValueError("ln mode must be 'glibc' or 'cr'")
    _lib.orc_set_ln_mode(1 if mode == "cr" else 0)
    try:
        yield
    finally:
        _lib.orc_set_ln_mode(0)


def fd_diff(shape: OrcShape, pts, h: float, powers=None, threads=None) -> np.ndarray:
    """DE(p + h e_a) - DE(p - h e_a) per axis (src/mesh/buffer.rs:256-266 stencil)."""
    p = _f32(pts).reshape(-1, 3)
    pw = None if powers is None else _f32(powers)
    out = np.empty((len(p), 3), np.float32)
    _lib.orc_fd_diff(C.byref(shape), _ptr(p), _ptr(pw), len(p), float(h), _ptr(out), threads or default_threads())
    return out


def powi(x: float, n: int) -> float:
    return _lib.orc_powi_f32(x, n)
