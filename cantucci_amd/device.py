"""Device-tensor entry points: torch supplies HBM buffers and the stream, the
C ABI's `_dev` functions do the work (asynchronously on torch's current
stream)."""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib
from ._lib import check, lib
from .shape import Context, Shape


def _ptr(t: torch.Tensor | None) -> int | None:
    if t is None:
        return None
    if not t.is_cuda or not t.is_contiguous():
        raise ValueError("expected a contiguous device tensor")
    return t.data_ptr()


def _stream() -> int:
    return torch.cuda.current_stream().cuda_stream


def _f32(t: torch.Tensor) -> torch.Tensor:
    if t.dtype != torch.float32:
        raise TypeError("expected float32")
    return t


def _points(pts: torch.Tensor) -> torch.Tensor:
    """(n, 3) float32 points, made contiguous.  The kernels read 3 floats per
    point, so any other shape would send them past the buffer."""
    if pts.dim() != 2 or pts.shape[1] != 3:
        raise ValueError(f"points must have shape (n, 3), got {tuple(pts.shape)}")
    return _f32(pts).contiguous()


def _per_point(t: torch.Tensor | None, n: int, name: str) -> torch.Tensor | None:
    """An optional float32 tensor of one value per point (per-point powers,
    incoming gradients)."""
    if t is None:
        return None
    if t.dtype != torch.float32:
        raise TypeError(f"{name} must be float32, got {t.dtype}")
    if t.numel() != n:
        raise ValueError(f"{name} must hold {n} values (one per point), got {t.numel()}")
    return t.contiguous()


def de_points(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, with_iters: bool = False,
              ctx: Context | None = None):
    """Distances of (n, 3) f32 points; optionally the per-point iteration count."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    out = torch.empty(n, dtype=torch.float32, device=pts.device)
    iters = torch.empty(n, dtype=torch.int32, device=pts.device) if with_iters else None
    check(lib.cth_min_distance_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(out),
                                   _ptr(iters), _stream()))
    return (out, iters) if with_iters else out


def de_state(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, ctx: Context | None = None):
    """(r, dr, iters): the loop state the distance is formed from."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    r = torch.empty(n, dtype=torch.float32, device=pts.device)
    dr = torch.empty_like(r)
    it = torch.empty(n, dtype=torch.int32, device=pts.device)
    check(lib.cth_de_state_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(r), _ptr(dr),
                               _ptr(it), _stream()))
    return r, dr, it


def de_grid(shape: Shape, start, end, n: int, divisor: float, with_iters: bool = False, device="cuda",
            ctx: Context | None = None):
    """n^3 grid in GridTable order (x slowest, z fastest)."""
    ctx = ctx or shape.ctx
    out = torch.empty(n * n * n, dtype=torch.float32, device=device)
    iters = torch.empty(n * n * n, dtype=torch.int32, device=device) if with_iters else None
    s = (C.c_float * 3)(*start)
    e = (C.c_float * 3)(*end)
    check(lib.cth_de_grid_dev(ctx.handle, C.byref(shape.desc), s, e, n, float(divisor), _ptr(out), _ptr(iters),
                              _stream()))
    return (out, iters) if with_iters else out


def sample_spans(shape: Shape, spans: torch.Tensor, res: int, with_cross: bool = True, with_iters: bool = False,
                 ctx: Context | None = None, out: torch.Tensor | None = None, cross: torch.Tensor | None = None):
    """Corner distances ((n, (res+1)^3) f32), crossing words ((n, ceil(res^3/64))
    int64) and optionally per-corner iteration counts."""
    ctx = ctx or shape.ctx
    spans = _f32(spans).contiguous()
    m = spans.shape[0]
    per = (res + 1) ** 3
    words = (res ** 3 + 63) // 64
    dists = out if out is not None else torch.empty((m, per), dtype=torch.float32, device=spans.device)
    if with_cross and cross is None:
        cross = torch.empty((m, words), dtype=torch.int64, device=spans.device)
    if not with_cross:
        cross = None
    iters = torch.empty((m, per), dtype=torch.int32, device=spans.device) if with_iters else None
    check(lib.cth_sample_spans_dev(ctx.handle, C.byref(shape.desc), _ptr(spans), m, res, _ptr(dists), _ptr(cross),
                                   _ptr(iters), _stream()))
    if with_iters:
        return dists, cross, iters
    return dists, cross


_NORMALS = {"fd6": _lib.CTH_NORMALS_FD6, "tangent": _lib.CTH_NORMALS_TANGENT,
            "deferred": _lib.CTH_NORMALS_DEFERRED}


def mesh(shape: Shape, spans: torch.Tensor, dists: torch.Tensor, res: int, normals: str = "fd6",
         ctx: Context | None = None, cross: torch.Tensor | None = None):
    """Naive surface nets, steps 2-3, over leaves already sampled by
    `sample_spans` at the same `res`.  Returns (vert_offsets (n+1,) int64,
    idx_offsets (n+1,) int64, vertices (nv, 7) f32 in the `Vertex` layout,
    indices (ni,) int32 holding u32 leaf-local vertex numbers).  `cross`
    (optional (n, ceil(res^3/64)) int64) receives the crossing words."""
    ctx = ctx or shape.ctx
    spans = _f32(spans).contiguous()
    dists = _f32(dists).contiguous()
    m = spans.shape[0]
    if dists.numel() != m * (res + 1) ** 3:
        raise ValueError("dists must hold (res+1)^3 distances per span")
    voff = torch.empty(m + 1, dtype=torch.int64, device=spans.device)
    ioff = torch.empty(m + 1, dtype=torch.int64, device=spans.device)
    nv = C.c_uint64()
    ni = C.c_uint64()
    if cross is not None and cross.numel() != m * ((res ** 3 + 63) // 64):
        raise ValueError("cross must hold ceil(res^3/64) words per span")
    check(lib.cth_mesh_count_dev(ctx.handle, _ptr(dists), m, res, _ptr(voff), _ptr(ioff), _ptr(cross), C.byref(nv),
                                 C.byref(ni), _stream()))
    verts = torch.empty((max(nv.value, 1), 7), dtype=torch.float32, device=spans.device)
    idx = torch.empty(max(ni.value, 1), dtype=torch.int32, device=spans.device)
    check(lib.cth_mesh_emit_dev(ctx.handle, C.byref(shape.desc), _ptr(spans), _ptr(dists), m, res,
                                _NORMALS[normals], _ptr(voff), _ptr(ioff), nv.value, _ptr(verts), _ptr(idx),
                                _stream()))
    return voff, ioff, verts[:nv.value], idx[:ni.value]


def mesh_vertices(shape: Shape, verts: torch.Tensor, normals: str = "fd6", ctx: Context | None = None):
    """The vertex pass of surface nets over (nv, 7) records that `mesh` wrote
    with normals="deferred": fills each normal and distance_from_surface in
    place."""
    ctx = ctx or shape.ctx
    verts = _f32(verts)
    if verts.dim() != 2 or verts.shape[1] != 7 or not verts.is_contiguous():
        raise ValueError("vertices must be a contiguous (n, 7) float32 tensor")
    check(lib.cth_mesh_vertices_dev(ctx.handle, C.byref(shape.desc), _ptr(verts), verts.shape[0], _NORMALS[normals],
                                    _stream()))
    return verts


def focus_split(ctx: Context, leaves: torch.Tensor, focii: torch.Tensor, eye, depth: torch.Tensor | None = None):
    """The reference's split rule (ShapeMesh::update): a uint32 flag per leaf,
    1 where a focus point lies in the leaf and the eye is nearer than twice
    the leaf's width.  `depth` (optional) skips focus rays that missed."""
    leaves = _f32(leaves).contiguous()
    focii = _f32(focii).contiguous()
    flags = torch.empty(leaves.shape[0], dtype=torch.int32, device=leaves.device)
    e = (C.c_float * 3)(*[float(v) for v in eye])
    check(lib.cth_focus_split_dev(ctx.handle, _ptr(leaves), leaves.shape[0], _ptr(focii),
                                  _ptr(depth.contiguous()) if depth is not None else None, focii.shape[0], e,
                                  _ptr(flags), _stream()))
    return flags


def split_leaves(ctx: Context, leaves: torch.Tensor, flags: torch.Tensor) -> torch.Tensor:
    """Replace flagged leaves by their 8 create_spans children, in place order."""
    leaves = _f32(leaves).contiguous()
    n = C.c_size_t()
    check(lib.cth_split_leaves_dev(ctx.handle, _ptr(leaves), leaves.shape[0], _ptr(flags), None, 0, C.byref(n),
                                   _stream()))
    out = torch.empty((max(n.value, 1), 6), dtype=torch.float32, device=leaves.device)
    check(lib.cth_split_leaves_dev(ctx.handle, _ptr(leaves), leaves.shape[0], _ptr(flags), _ptr(out), out.shape[0],
                                   C.byref(n), _stream()))
    return out[:n.value]


def refine(shape: Shape, root, min_depth: int, max_depth: int, half_diag, cap: int | None = None,
           leaves: torch.Tensor | None = None, ctx: Context | None = None, device="cuda"):
    """Level-synchronous refinement; returns (leaves (m, 6) f32 on device, level_counts)."""
    import numpy as np
    ctx = ctx or shape.ctx
    r = (C.c_float * 6)(*root)
    hd = np.ascontiguousarray(np.asarray(half_diag, np.float32))
    lc = np.zeros(max_depth + 1, np.uint64)
    n = C.c_size_t()
    if leaves is None:
        check(lib.cth_refine_dev(ctx.handle, C.byref(shape.desc), r, min_depth, max_depth, hd.ctypes.data, None, 0,
                                 C.byref(n), lc.ctypes.data, _stream()))
        leaves = torch.empty((max(n.value, 1), 6), dtype=torch.float32, device=device)
    check(lib.cth_refine_dev(ctx.handle, C.byref(shape.desc), r, min_depth, max_depth, hd.ctypes.data,
                             _ptr(leaves), leaves.shape[0], C.byref(n), lc.ctypes.data, _stream()))
    return leaves[:n.value], lc


def sphere_trace(shape: Shape, origins: torch.Tensor, dirs: torch.Tensor, max_steps: int = 100, eps_abs: float = 1e-6,
                 eps_rel: float = 0.0, t_max: float = float("inf"), with_pos: bool = False,
                 ctx: Context | None = None):
    """Sphere-trace explicit rays; returns (depth, steps[, hit_pos])."""
    ctx = ctx or shape.ctx
    origins = _f32(origins).contiguous().reshape(-1, 3)
    dirs = _f32(dirs).contiguous()
    n = dirs.shape[0]
    per_ray = 1 if origins.shape[0] == n and n > 1 else 0
    depth = torch.empty(n, dtype=torch.float32, device=dirs.device)
    steps = torch.empty(n, dtype=torch.int32, device=dirs.device)
    pos = torch.empty((n, 3), dtype=torch.float32, device=dirs.device) if with_pos else None
    check(lib.cth_sphere_trace_dev(ctx.handle, C.byref(shape.desc), _ptr(origins), per_ray, _ptr(dirs), n, max_steps,
                                   eps_abs, eps_rel, t_max, _ptr(depth), _ptr(steps), _ptr(pos), _stream()))
    return (depth, steps, pos) if with_pos else (depth, steps)


def render_depth(shape: Shape, cams: torch.Tensor, width: int, height: int, max_steps: int = 256,
                 eps_abs: float = 0.0, eps_rel: float = 1e-4, t_max: float = 10.0, ctx: Context | None = None,
                 depth: torch.Tensor | None = None, steps: torch.Tensor | None = None):
    """Depth images of (ncams, 14) f32 camera records; returns (depth, steps) of shape (ncams, H, W)."""
    ctx = ctx or shape.ctx
    cams = _f32(cams).contiguous()
    m = cams.shape[0]
    if depth is None:
        depth = torch.empty((m, height, width), dtype=torch.float32, device=cams.device)
    if steps is None:
        steps = torch.empty((m, height, width), dtype=torch.int32, device=cams.device)
    check(lib.cth_render_depth_dev(ctx.handle, C.byref(shape.desc), _ptr(cams), m, width, height, max_steps, eps_abs,
                                   eps_rel, t_max, _ptr(depth), _ptr(steps), _stream()))
    return depth, steps


def get_focii(shape: Shape, cams: torch.Tensor, focus_points: int = 5, max_steps: int = 100, eps: float = 1e-6,
              with_dirs: bool = False, ctx: Context | None = None):
    """ShapeMesh::get_focii for (ncams, 10) f32 reference cameras (position,
    unit direction, fov, aspect, near, far), rays built and marched on the
    device.  Returns (focii (ncams, fp^2, 3), depth (ncams, fp^2; +inf on a
    miss), steps[, dirs (ncams, fp^2, 3)])."""
    ctx = ctx or shape.ctx
    cams = _f32(cams).contiguous()
    if cams.dim() != 2 or cams.shape[1] != 10:
        raise ValueError(f"cameras must have shape (n, 10), got {tuple(cams.shape)}")
    m, n = cams.shape[0], focus_points * focus_points
    focii = torch.empty((m, n, 3), dtype=torch.float32, device=cams.device)
    depth = torch.empty((m, n), dtype=torch.float32, device=cams.device)
    steps = torch.empty((m, n), dtype=torch.int32, device=cams.device)
    dirs = torch.empty((m, n, 3), dtype=torch.float32, device=cams.device) if with_dirs else None
    check(lib.cth_get_focii_dev(ctx.handle, C.byref(shape.desc), _ptr(cams), m, focus_points, max_steps, eps,
                                _ptr(focii), _ptr(depth), _ptr(steps), _ptr(dirs), _stream()))
    return (focii, depth, steps, dirs) if with_dirs else (focii, depth, steps)


def de_vjp(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, grad_d: torch.Tensor | None = None,
           with_power: bool = False, ctx: Context | None = None):
    """(d, grad_p[, grad_power]) from the fused forward + reverse sweep."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    grad_d = _per_point(grad_d, n, "grad_d")
    d = torch.empty(n, dtype=torch.float32, device=pts.device)
    gp = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    gpow = torch.empty(n, dtype=torch.float32, device=pts.device) if with_power else None
    check(lib.cth_de_vjp_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(grad_d), _ptr(d),
                             _ptr(gp), _ptr(gpow), _stream()))
    return (d, gp, gpow) if with_power else (d, gp)


def de_vjp_split(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None,
                 grad_d: torch.Tensor | None = None, with_power: bool = False, ctx: Context | None = None):
    """(d, grad_z0, grad_c[, grad_power]): the reverse sweep's gradient with the
    point's two roles apart, the orbit's start z0 and the constant c of every
    body (grad_p = grad_z0 + grad_c)."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    grad_d = _per_point(grad_d, n, "grad_d")
    d = torch.empty(n, dtype=torch.float32, device=pts.device)
    gz = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    gc = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    gpow = torch.empty(n, dtype=torch.float32, device=pts.device) if with_power else None
    check(lib.cth_de_vjp_split_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(grad_d),
                                   _ptr(d), _ptr(gz), _ptr(gc), _ptr(gpow), _stream()))
    return (d, gz, gc, gpow) if with_power else (d, gz, gc)


def de_grad(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, grad_d: torch.Tensor | None = None,
            with_power: bool = False, ctx: Context | None = None):
    """(d, grad_p[, grad_power]) in forward (tangent-linear) mode."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    grad_d = _per_point(grad_d, n, "grad_d")
    d = torch.empty(n, dtype=torch.float32, device=pts.device)
    gp = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    gpow = torch.empty(n, dtype=torch.float32, device=pts.device) if with_power else None
    check(lib.cth_de_grad_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(grad_d), _ptr(d),
                              _ptr(gp), _ptr(gpow), _stream()))
    return (d, gp, gpow) if with_power else (d, gp)


def normals(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, ctx: Context | None = None,
            out: torch.Tensor | None = None, d: torch.Tensor | None = None):
    """Unit normals (n, 3) from one fused launch (parity forward + tangent
    Jacobian); returns (d, normals)."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    if d is None:
        d = torch.empty(n, dtype=torch.float32, device=pts.device)
    check(lib.cth_normals_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, _ptr(d), _ptr(out),
                              _stream()))
    return d, out


def normals_fd(shape: Shape, pts: torch.Tensor, h: float, powers: torch.Tensor | None = None,
               ctx: Context | None = None, out: torch.Tensor | None = None, d: torch.Tensor | None = None,
               grad: torch.Tensor | None = None):
    """6-tap central-difference normals; returns (d, normals)."""
    ctx = ctx or shape.ctx
    pts = _points(pts)
    n = pts.shape[0]
    powers = _per_point(powers, n, "powers")
    if out is None:
        out = torch.empty((n, 3), dtype=torch.float32, device=pts.device)
    if d is None:
        d = torch.empty(n, dtype=torch.float32, device=pts.device)
    check(lib.cth_normals_fd_dev(ctx.handle, C.byref(shape.desc), _ptr(pts), _ptr(powers), n, float(h), _ptr(d),
                                 _ptr(out), _ptr(grad), _stream()))
    return d, out


class MandelbulbDistance(torch.autograd.Function):
    """torch.autograd.Function over the C ABI: forward = cth_min_distance_dev,
    backward = the vector-Jacobian product grad_d * (dd/dp, dd/dpower), with
    the forward iteration recomputed per lane.  Two kernels compute it:
    "forward" (default) carries the orbit's Jacobian through every loop body
    in one launch and needs no workspace (cth_de_grad_dev); "reverse" records
    the orbit and sweeps back through the power map (cth_de_vjp_dev).  With 3
    point coordinates and a power as the only inputs, the forward sweep does
    less work: the reverse sweep has to re-evaluate every transcendental of
    the map on its way back."""

    @staticmethod
    def forward(ctx_, pts, powers, shape, sweep):
        ctx_.shape = shape
        ctx_.sweep = sweep
        ctx_.save_for_backward(pts, powers)
        return de_points(shape, pts.detach(), None if powers is None else powers.detach())

    @staticmethod
    def backward(ctx_, grad_d):
        pts, powers = ctx_.saved_tensors
        shape = ctx_.shape
        fn = de_grad if ctx_.sweep == "forward" else de_vjp
        _, gp, gpow = fn(shape, pts.detach(), None if powers is None else powers.detach(), grad_d.contiguous(),
                         with_power=True)
        return gp, (gpow if (powers is not None and ctx_.needs_input_grad[1]) else None), None, None


def distance(shape: Shape, pts: torch.Tensor, powers: torch.Tensor | None = None, sweep: str = "forward"
             ) -> torch.Tensor:
    """Differentiable distance estimate; `sweep` picks the backward kernel
    ("forward": tangent-linear, "reverse": orbit record + reverse sweep)."""
    if sweep not in ("forward", "reverse"):
        raise ValueError("sweep must be 'forward' or 'reverse'")
    return MandelbulbDistance.apply(pts, powers, shape, sweep)


class JobSet:
    """A batch of independent grid jobs (`cth_jobset`)."""

    def __init__(self, ctx: Context, kind: int, bailout: float, jobs):
        arr = (_lib.CthJob * len(jobs))()
        for a, j in zip(arr, jobs):
            a.n = j["n"]
            a.power = j["power"]
            a.max_iters = j["max_iters"]
            a.start[:] = list(j.get("start", (-1.5, -1.5, -1.5)))
            a.end[:] = list(j.get("end", (1.5, 1.5, 1.5)))
            a.divisor = j.get("divisor", j["n"] - 1.0)
        h = C.c_void_p()
        check(lib.cth_jobs_create(ctx.handle, kind, bailout, arr, len(jobs), C.byref(h)))
        self.ctx = ctx
        self.handle = h
        self.njobs = len(jobs)
        self.total = lib.cth_jobs_total(h)
        self.offsets = [lib.cth_jobs_offset(h, i) for i in range(self.njobs)] + [self.total]

    def run(self, dists: torch.Tensor, normals: torch.Tensor | None, done: torch.Tensor, start: torch.Tensor,
            stream: int | None = None):
        check(lib.cth_jobs_run_dev(self.ctx.handle, self.handle, _ptr(dists), _ptr(normals), _ptr(done), _ptr(start),
                                   _stream() if stream is None else stream))

    def close(self):
        if self.handle:
            lib.cth_jobs_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """HIP graph captured from a callable that issues `_dev` calls on `stream`."""

    def __init__(self, fn, stream: torch.cuda.Stream):
        self.stream = stream
        with torch.cuda.stream(stream):
            check(lib.cth_graph_begin(stream.cuda_stream))
            try:
                fn()
            finally:
                h = C.c_void_p()
                rc = lib.cth_graph_end(stream.cuda_stream, C.byref(h))
            check(rc)
        self.handle = h

    def launch(self):
        check(lib.cth_graph_launch(self.handle, self.stream.cuda_stream))

    def close(self):
        if getattr(self, "handle", None):
            lib.cth_graph_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
