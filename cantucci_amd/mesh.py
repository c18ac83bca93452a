"""Mesh extraction mirror: `MeshBuffer::generate_for_box`
(src/mesh/buffer.rs:23-42), whose three steps run on the device through
cth_mesh_span (corner sampling, vertex placement + normals, quads),
`ShapeMesh::get_focii` (src/mesh/mod.rs:205-243) through cth_get_focii, and
`ShapeMesh::new` / `update` (src/mesh/mod.rs:45-170), which mesh a batch of
octree leaves through cth_mesh_spans."""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .shape import Shape

# `Vertex` (src/mesh/mod.rs:255-261): position[3], normal[3], distance_from_surface.
VERTEX_DTYPE = np.dtype([("position", np.float32, 3), ("normal", np.float32, 3),
                         ("distance_from_surface", np.float32)])


@dataclass
class Timings:
    """The counts of the reference's `Timings` (src/mesh/buffer.rs:375-381)."""
    vertices: int
    faces: int


@dataclass
class MeshBuffer:
    vertices: np.ndarray  # VERTEX_DTYPE records
    indices: np.ndarray   # u32, local to this buffer

    @classmethod
    def generate_for_box(cls, span, shape: Shape, resolution: int, normals: str = "fd6"):
        """`MeshBuffer::generate_for_box(span, shape, resolution)`: returns
        (MeshBuffer, Timings).  Raises CthError where the reference asserts
        (start < end per axis, power-of-two resolution)."""
        sp = _lib.CthSpan()
        sp.start[:] = [float(v) for v in span[0]]
        sp.end[:] = [float(v) for v in span[1]]
        mode = {"fd6": _lib.CTH_NORMALS_FD6, "tangent": _lib.CTH_NORMALS_TANGENT}[normals]
        nv = C.c_size_t()
        ni = C.c_size_t()
        h = shape.ctx.handle
        check(lib.cth_mesh_span(h, C.byref(shape.desc), C.byref(sp), resolution, mode, None, 0, None, 0,
                                C.byref(nv), C.byref(ni)))
        verts = np.empty(nv.value, VERTEX_DTYPE)
        idx = np.empty(ni.value, np.uint32)
        check(lib.cth_mesh_span(h, C.byref(shape.desc), C.byref(sp), resolution, mode,
                                verts.ctypes.data if nv.value else None, nv.value,
                                idx.ctypes.data if ni.value else None, ni.value, C.byref(nv), C.byref(ni)))
        return cls(verts, idx), Timings(nv.value, ni.value // 6)


def get_focii(shape: Shape, position, direction, fov: float, aspect: float, near_plane: float, far_plane: float,
              focus_points: int = 5) -> np.ndarray:
    """`ShapeMesh::get_focii(camera, focus_points)` for the camera
    `Camera::new(position, direction, Projection { fov, aspect, near, far })`
    (src/camera.rs:26-34,135-144): the hit points, (k, 3) f32 in ray order.
    Rays are built and marched on the device."""
    d = np.asarray(direction, np.float32)
    if not d.any():
        raise ValueError("direction must not be zero (Camera::new asserts it)")
    cam = _lib.CthRefCamera()
    cam.position[:] = [float(v) for v in position]
    # Camera::new stores the normalised direction (v * (1 / |v|), in f32)
    m = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
    cam.direction[:] = [float(v) for v in d * (np.float32(1.0) / m)]
    cam.fov, cam.aspect, cam.near_plane, cam.far_plane = fov, aspect, near_plane, far_plane
    out = np.empty((focus_points * focus_points, 3), np.float32)
    k = C.c_size_t()
    check(lib.cth_get_focii(shape.ctx.handle, C.byref(shape.desc), C.byref(cam), focus_points, out.ctypes.data,
                            C.byref(k)))
    return out[:k.value]


@dataclass
class Camera:
    """`Camera::new(position, direction, projection)` (src/camera.rs:26-34,
    108-144): the direction is stored normalised, v * (1 / |v|) in f32, and
    must not be zero (the reference asserts it)."""
    position: tuple
    direction: tuple
    fov: float
    aspect: float
    near_plane: float
    far_plane: float

    def __post_init__(self):
        d = np.asarray(self.direction, np.float32)
        if not d.any():
            raise ValueError("direction must not be zero (Camera::new asserts it)")
        m = np.sqrt((d[0] * d[0] + d[1] * d[1]) + d[2] * d[2])
        self.direction = tuple(float(v) for v in d * (np.float32(1.0) / m))
        self.position = tuple(float(v) for v in np.asarray(self.position, np.float32))


def create_spans(span):
    """`create_spans` (src/octree/mod.rs:315-329): the 8 children of
    ((sx, sy, sz), (ex, ey, ez)), child bit 2 = upper x half, bit 1 = y,
    bit 0 = z, around start + (end - start) / 2 in f32."""
    s = np.asarray(span[0], np.float32)
    e = np.asarray(span[1], np.float32)
    c = s + (e - s) / np.float32(2.0)
    out = []
    for i in range(8):
        hi = np.array([(i >> 2) & 1, (i >> 1) & 1, i & 1], bool)
        out.append((tuple(float(v) for v in np.where(hi, c, s)), tuple(float(v) for v in np.where(hi, e, c))))
    return out


class ShapeMesh:
    """`ShapeMesh` (src/mesh/mod.rs:22-178) without its rendering half: the
    leaves of an octree over the shape's bounding box, each with the
    MeshBuffer of its span at resolution 64 once meshed.  The reference
    meshes empty leaves on a CPU thread pool and picks the buffers up in later
    frames; update() meshes all of them in one batch on the device before it
    returns.  Leaves are kept depth-first, a split leaf replaced by its 8
    children in create_spans order."""
    RESOLUTION = 64   # (:131)
    FOCUS_POINTS = 5  # (:86)

    def __init__(self, shape: Shape):
        # ShapeMesh::new (:45-74): the bounding box split twice, 64 leaves
        self.shape = shape
        self.leaves = [[leaf, None] for child in create_spans(shape.bounding_box()) for leaf in create_spans(child)]
        self.batch_timings = Timings(0, 0)

    def get_focii(self, camera: Camera, focus_points: int = 5) -> np.ndarray:
        return get_focii(self.shape, camera.position, camera.direction, camera.fov, camera.aspect,
                         camera.near_plane, camera.far_plane, focus_points)

    def update(self, camera: Camera) -> None:
        """ShapeMesh::update (:82-170): every meshed leaf that holds a focus
        point the camera is nearer to than twice the leaf's width is split
        (its mesh dropped); then every empty leaf is meshed."""
        focii = np.ascontiguousarray(self.get_focii(camera, self.FOCUS_POINTS), np.float32)
        spans = np.array([[*a, *b] for (a, b), _ in self.leaves], np.float32)
        flags = np.zeros(len(spans), np.uint32)
        eye = np.asarray(camera.position, np.float32)
        check(lib.cth_focus_split(self.shape.ctx.handle, spans.ctypes.data, len(spans),
                                  focii.ctypes.data if len(focii) else None, len(focii), eye.ctypes.data,
                                  flags.ctypes.data))
        nxt = []
        for (span, buf), f in zip(self.leaves, flags):
            if f and buf is not None:  # only leaves showing a mesh are split (:94)
                nxt.extend([child, None] for child in create_spans(span))
            else:
                nxt.append([span, buf])
        self.leaves = nxt
        self._mesh_empty_leaves()

    def _mesh_empty_leaves(self) -> None:
        todo = [i for i, (_, buf) in enumerate(self.leaves) if buf is None]
        if not todo:
            return
        spans = np.array([[*self.leaves[i][0][0], *self.leaves[i][0][1]] for i in todo], np.float32)
        h, n = self.shape.ctx.handle, len(todo)
        nv, ni = C.c_size_t(), C.c_size_t()
        check(lib.cth_mesh_spans(h, C.byref(self.shape.desc), spans.ctypes.data, n, self.RESOLUTION,
                                 _lib.CTH_NORMALS_FD6, None, 0, None, 0, None, None, C.byref(nv), C.byref(ni)))
        verts = np.empty(nv.value, VERTEX_DTYPE)
        idx = np.empty(ni.value, np.uint32)
        voff = np.empty(n + 1, np.uint64)
        ioff = np.empty(n + 1, np.uint64)
        check(lib.cth_mesh_spans(h, C.byref(self.shape.desc), spans.ctypes.data, n, self.RESOLUTION,
                                 _lib.CTH_NORMALS_FD6, verts.ctypes.data if nv.value else None, nv.value,
                                 idx.ctypes.data if ni.value else None, ni.value, voff.ctypes.data, ioff.ctypes.data,
                                 C.byref(nv), C.byref(ni)))
        for k, i in enumerate(todo):
            a, b, c, d = int(voff[k]), int(voff[k + 1]), int(ioff[k]), int(ioff[k + 1])
            self.leaves[i][1] = MeshBuffer(verts[a:b].copy(), idx[c:d].copy())
            self.batch_timings = Timings(self.batch_timings.vertices + (b - a),
                                         self.batch_timings.faces + (d - c) // 6)
