"""Host-side mirror of cantucci's `Shape` trait (src/shape/mod.rs:24-105).

Each method keeps the reference's name, argument meaning and error
behaviour; the work runs on the GPU through the C ABI.  Where the reference
panics (`batch_max_distance_from` on a shape without an upper bound,
src/shape/util.rs:7-9) the mirror raises ``CthError`` with code
CTH_E_UNSUPPORTED.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import _lib
from ._lib import check, lib, make_shape


class Context:
    """One HIP stream plus work counters on one device (`cth_ctx`)."""

    def __init__(self, device: int = 0, mode: str = "parity"):
        h = C.c_void_p()
        check(lib.cth_ctx_create(device, C.byref(h)))
        self.handle = h
        self.set_mode(mode)

    def set_mode(self, mode: str) -> None:
        m = {"parity": _lib.CTH_MODE_PARITY, "fast": _lib.CTH_MODE_FAST, "contract": _lib.CTH_MODE_CONTRACT}[mode]
        check(lib.cth_ctx_set_mode(self.handle, m))
        self.mode = mode

    def set_launch(self, kernel: int, block: int, blocks_per_cu: int, chunk: int, refill_min: int,
                   dynamic: int = 0) -> None:
        l = _lib.CthLaunch(block, blocks_per_cu, chunk, refill_min, dynamic)
        check(lib.cth_ctx_set_launch(self.handle, kernel, C.byref(l)))

    def get_launch(self, kernel: int) -> dict:
        l = _lib.CthLaunch()
        check(lib.cth_ctx_get_launch(self.handle, kernel, C.byref(l)))
        return dict(block=l.block, blocks_per_cu=l.blocks_per_cu, chunk=l.chunk, refill_min=l.refill_min,
                    dynamic=l.dynamic)

    def count_work(self, counters) -> None:
        """Accumulate the refill launches' work into `counters` (a device
        tensor of 3 int64: loop bodies, bailed evaluations, evaluations), or
        stop with None."""
        check(lib.cth_ctx_count_work(self.handle, None if counters is None else counters.data_ptr()))

    @property
    def stream(self) -> int:
        return lib.cth_ctx_stream(self.handle) or 0

    @property
    def num_cus(self) -> int:
        return lib.cth_ctx_num_cus(self.handle)

    def synchronize(self) -> None:
        check(lib.cth_synchronize(self.handle))

    def close(self) -> None:
        if getattr(self, "handle", None):
            lib.cth_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_tls = threading.local()


def default_context() -> Context:
    """A per-thread context, like the reference's per-worker use of one
    `Arc<dyn Shape>` from many pool threads (src/mesh/mod.rs:60-63)."""
    ctx = getattr(_tls, "ctx", None)
    if ctx is None:
        ctx = Context()
        _tls.ctx = ctx
    return ctx


def _as_points(points) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(points, dtype=np.float32))
    if a.ndim == 1:
        a = a.reshape(1, 3)
    if a.ndim != 2 or a.shape[1] != 3:
        raise ValueError("points must have shape (n, 3)")
    return a


class Shape:
    """Mirror of the trait.  Subclasses set ``self.desc`` (a CthShape)."""

    desc: _lib.CthShape

    def __init__(self, ctx: Context | None = None):
        self._ctx = ctx

    @property
    def ctx(self) -> Context:
        return self._ctx or default_context()

    # fn min_distance_from(&self, p: Point3<f32>) -> f32          (:37)
    def min_distance_from(self, p) -> float:
        return float(self.batch_min_distance_from(_as_points(p))[0])

    # fn bounding_box(&self) -> Range<Point3<f32>>                (:39)
    def bounding_box(self):
        span = _lib.CthSpan()
        check(lib.cth_bounding_box(C.byref(self.desc), C.byref(span)))
        return tuple(span.start), tuple(span.end)

    # fn max_distance_from(&self, p) -> Option<f32>                (:62-64)
    def max_distance_from(self, p):
        return None

    # fn bounded_distance_from(&self, p) -> (f32, Option<f32>)     (:73-75)
    def bounded_distance_from(self, p):
        return self.min_distance_from(p), self.max_distance_from(p)

    # fn contains(&self, p) -> bool                                (:78-80)
    def contains(self, p) -> bool:
        return bool(self.batch_contains(_as_points(p))[0])

    # fn batch_min_distance_from(&self, &[Point3<f32>]) -> Vec<f32> (:89)
    def batch_min_distance_from(self, points) -> np.ndarray:
        pts = _as_points(points)
        out = np.empty(len(pts), np.float32)
        check(lib.cth_min_distance_batch(self.ctx.handle, C.byref(self.desc), pts.ctypes.data, len(pts),
                                         out.ctypes.data))
        return out

    # fn batch_max_distance_from(&self, &[Point3<f32>]) -> Vec<f32> (:96)
    def batch_max_distance_from(self, points) -> np.ndarray:
        pts = _as_points(points)
        out = np.empty(len(pts), np.float32)
        check(lib.cth_max_distance_batch(self.ctx.handle, C.byref(self.desc), pts.ctypes.data, len(pts),
                                         out.ctypes.data))
        return out

    # fn batch_bounded_distance_from(&self, ...) -> Vec<(f32, f32)>  (:103)
    def batch_bounded_distance_from(self, points) -> np.ndarray:
        pts = _as_points(points)
        out = np.empty((len(pts), 2), np.float32)
        check(lib.cth_bounded_distance_batch(self.ctx.handle, C.byref(self.desc), pts.ctypes.data, len(pts),
                                             out.ctypes.data))
        return out

    def batch_contains(self, points) -> np.ndarray:
        pts = _as_points(points)
        out = np.empty(len(pts), np.uint8)
        check(lib.cth_contains_batch(self.ctx.handle, C.byref(self.desc), pts.ctypes.data, len(pts),
                                     out.ctypes.data))
        return out.astype(bool)


class Mandelbulb(Shape):
    """`Mandelbulb<P>` (src/shape/mandelbulb.rs:12-33)."""

    def __init__(self, max_iters: int, bailout: float, power: int = 8, ctx: Context | None = None):
        super().__init__(ctx)
        if max_iters < 1:  # assert!(max_iters >= 1)  (:20)
            raise ValueError("max_iters must be >= 1")
        self.desc = make_shape(_lib.CTH_SHAPE_MANDELBULB, float(power), int(max_iters), float(bailout))

    @classmethod
    def classic(cls, max_iters: int, bailout: float, ctx: Context | None = None) -> "Mandelbulb":
        """`Mandelbulb::<8>::classic` (src/shape/mandelbulb.rs:29-33)."""
        return cls(max_iters, bailout, 8, ctx)


class MandelbulbTrig(Shape):
    """The trigonometric map with a real power (powf)."""

    def __init__(self, max_iters: int, bailout: float, power: float, ctx: Context | None = None):
        super().__init__(ctx)
        if max_iters < 1:
            raise ValueError("max_iters must be >= 1")
        self.desc = make_shape(_lib.CTH_SHAPE_MANDELBULB_TRIG, float(power), int(max_iters), float(bailout))


class Sphere(Shape):
    """`Sphere` (src/shape/sphere.rs:6-52)."""

    def __init__(self, center, radius: float, ctx: Context | None = None):
        super().__init__(ctx)
        self.desc = make_shape(_lib.CTH_SHAPE_SPHERE, 0.0, 1, 0.0, tuple(center), float(radius))

    def max_distance_from(self, p):
        return self.min_distance_from(p)
