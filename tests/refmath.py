"""Independent closed forms used to pin the oracle (FP64 numpy).

These are *not* restatements of the reference's f32 operation sequence: they
use other formulations of the same mathematics (complex powers, spherical
coordinates) so that a transposed term, a wrong coefficient or a wrong sign in
the oracle shows up as a disagreement.
"""
from __future__ import annotations

import ctypes

import numpy as np

_libm = ctypes.CDLL("libm.so.6")
_libm.logf.restype = ctypes.c_float
_libm.logf.argtypes = [ctypes.c_float]


def glibc_logf(x: np.ndarray) -> np.ndarray:
    """glibc's logf (what Rust's f32::ln calls on x86_64-linux-gnu)."""
    return np.array([_libm.logf(float(v)) for v in np.asarray(x, np.float32).ravel()], np.float32).reshape(
        np.shape(x))


def p8_map_complex(x, y, z):
    """The reference's P=8 polynomial map in complex form.

    With w = sqrt(x^2 + y^2):
      a  = Re((w + i z)^8) / w^8
      x' = a * (Re((x + i y)^8) - 2 y^8)     (the `- y8` term, src/shape/mandelbulb.rs:187)
      y' = a * Im((x + i y)^8)
      z' = -Im((w + i z)^8)
    """
    x = np.asarray(x, np.float64)
    y = np.asarray(y, np.float64)
    z = np.asarray(z, np.float64)
    w = np.sqrt(x * x + y * y)
    q = (w + 1j * z) ** 8
    c = (x + 1j * y) ** 8
    a = q.real / w ** 8
    return a * (c.real - 2 * y ** 8), a * c.imag, -q.imag


def latitude_map(x, y, z, n=8):
    """r^n (cos n t cos n f, cos n t sin n f, sin n t), t measured from the xy-plane."""
    r = np.sqrt(x * x + y * y + z * z)
    t = np.arcsin(z / r)
    f = np.arctan2(y, x)
    R = r ** n
    return R * np.cos(n * t) * np.cos(n * f), R * np.cos(n * t) * np.sin(n * f), R * np.sin(n * t)


def trig_map(x, y, z, P):
    """Textbook triplex power (polar angle from +z): r^P (sin Pt cos Pf, sin Pt sin Pf, cos Pt)."""
    r = np.sqrt(x * x + y * y + z * z)
    t = np.arccos(z / r)
    f = np.arctan2(y, x)
    R = r ** P
    return R * np.sin(P * t) * np.cos(P * f), R * np.sin(P * t) * np.sin(P * f), R * np.cos(P * t)


def de_f64(p, mapfn, P, max_iters, bailout, c=None):
    """Distance estimate 0.5 ln(r) r / dr with dr' = P r^(P-1) dr + 1, in FP64,
    for an arbitrary map (vectorised, with the reference's loop structure).
    The orbit starts at p and adds c in every body; c defaults to p, as in the
    reference."""
    p = np.asarray(p, np.float64)
    x, y, z = p[:, 0].copy(), p[:, 1].copy(), p[:, 2].copy()
    c = p if c is None else np.asarray(c, np.float64)
    px, py, pz = c[:, 0], c[:, 1], c[:, 2]
    dr = np.ones_like(px)
    r = np.zeros_like(px)
    active = np.ones(px.shape, bool)
    iters = np.zeros(px.shape, np.int64)
    P = np.broadcast_to(np.asarray(P, np.float64), px.shape)
    with np.errstate(all="ignore"):
        for _ in range(max_iters):
            rr = np.sqrt(x * x + y * y + z * z)
            r = np.where(active, rr, r)
            active &= ~(rr > bailout)
            dr = np.where(active, P * rr ** (P - 1) * dr + 1, dr)
            nx, ny, nz = mapfn(x, y, z, P)
            x = np.where(active, nx + px, x)
            y = np.where(active, ny + py, y)
            z = np.where(active, nz + pz, z)
            iters += active
        d = 0.5 * np.log(r) * r / dr
    return d, iters


def cr_logf(x, band: float = 2.0 ** -18):
    """ln of f32 inputs correctly rounded to f32, and a mask of the "hard"
    inputs, whose ln lies within `band` f32 ulps of a rounding boundary (a
    midpoint between two neighbouring f32 values).

    The ln is taken in f64 (error below 2^-26 f32 ulps), so outside the mask
    its rounding to f32 is the correctly rounded value.  Inside the mask any
    implementation whose error bound is below `band` f32 ulps (the device's
    f64 logarithm is one, with a relative error below 2^-43, i.e. 2^-19 ulps)
    may round either way."""
    x = np.asarray(x, np.float32)
    with np.errstate(all="ignore"):
        y = np.log(x.astype(np.float64))
        f = y.astype(np.float32)
        nb = np.nextafter(f, np.where(y > f, np.float32(np.inf), np.float32(-np.inf)))
        mid = (f.astype(np.float64) + nb.astype(np.float64)) * 0.5
        gap = np.abs(nb.astype(np.float64) - f.astype(np.float64))
        hard = np.isfinite(y) & (y != f) & (np.abs(y - mid) < band * gap)
    return f, hard
