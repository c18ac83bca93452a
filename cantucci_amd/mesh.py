"""Mesh extraction mirror: `MeshBuffer::generate_for_box`
(src/mesh/buffer.rs:23-42), whose three steps run on the device through
cth_mesh_span (corner sampling, vertex placement + normals, quads)."""
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
