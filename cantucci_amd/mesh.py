"""Mesh extraction mirror: `MeshBuffer::generate_for_box`
(src/mesh/buffer.rs:23-42), whose three steps run on the device through
cth_mesh_span (corner sampling, vertex placement + normals, quads), and
`ShapeMesh::get_focii` (src/mesh/mod.rs:205-243) through cth_get_focii."""
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
