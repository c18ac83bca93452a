"""Synthetic workloads (the five benchmark configurations), seeded and
deterministic.  Randomness comes from SplitMix64 (Steele, Lea & Flood,
OOPSLA 2014), which is counter based: output i is mix(seed + (i+1) * gamma),
so a stream is generated vectorised.  Uniform floats take the top 24 bits.
"""
from __future__ import annotations

import math

import numpy as np

GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed: int, n: int, offset: int = 0) -> np.ndarray:
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + i * GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, n: int, lo=0.0, hi=1.0, offset: int = 0) -> np.ndarray:
    """f64 uniforms in [lo, hi) from the top 24 bits (exactly f32 representable on [0,1))."""
    u = (splitmix64(seed, n, offset) >> np.uint64(40)).astype(np.float64) * (1.0 / (1 << 24))
    return lo + (hi - lo) * u


# ---------------------------------------------------------------------------
# Plain grid: 64^3 over [-1.5, 1.5]^3, p = s + (e - s) * (i / 63)
# ---------------------------------------------------------------------------
GRID64 = dict(n=64, start=(-1.5, -1.5, -1.5), end=(1.5, 1.5, 1.5), divisor=63.0, power=8, max_iters=15, bailout=2.0)

# ---------------------------------------------------------------------------
# Octree workload: refinement to depth 8 then 8^3 spans per surface leaf
# ---------------------------------------------------------------------------
OCTREE = dict(root=(-1.5, -1.5, -1.5, 1.5, 1.5, 1.5), min_depth=2, max_depth=8, res=8, power=8, max_iters=15,
          bailout=2.0)


def half_diagonals(width: float = 3.0, max_depth: int = 8) -> np.ndarray:
    """sqrt(3) * w_d / 2 per depth d, w_d = width / 2^d, rounded once to f32."""
    return np.array([math.sqrt(3.0) * (width / 2 ** d) / 2.0 for d in range(max_depth + 1)], np.float32)


# ---------------------------------------------------------------------------
# Camera tracing: 1920x1080 primary rays for 32 orbit poses
# ---------------------------------------------------------------------------
TRACE_1080P = dict(W=1920, H=1080, poses=32, seed=0, fov_deg=60.0, max_steps=256, eps_rel=1e-4, t_max=10.0, power=8,
          max_iters=15, bailout=2.0)

THETA_SAFE = 0.02  # src/camera.rs:10


def look_at_basis(eye, target, up=(0.0, 0.0, 1.0)):
    """Right-handed look-at basis (cgmath look_at_rh, src/camera.rs:69-75):
    f = normalize(target - eye), s = normalize(f x up), u = s x f."""
    eye = np.asarray(eye, np.float64)
    f = np.asarray(target, np.float64) - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, np.asarray(up, np.float64))
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    return s, u, f


def orbit_poses(n=32, seed=0, W=1920, H=1080, fov_deg=60.0) -> np.ndarray:
    """n camera records {eye[3], right[3], up[3], fwd[3], tan_half_fov, aspect} (f32).
    Eye at radius U[2.5, 4] in a uniform direction with the polar angle clamped
    to [0.02, pi - 0.02] (src/camera.rs:10,107-109); target = origin + U[-0.1, 0.1]^3."""
    u = uniform(seed, 6 * n).reshape(n, 6)
    out = np.zeros((n, 14), np.float32)
    for i in range(n):
        radius = 2.5 + 1.5 * u[i, 0]
        cos_t = 1.0 - 2.0 * u[i, 1]
        theta = min(max(math.acos(cos_t), THETA_SAFE), math.pi - THETA_SAFE)
        phi = 2.0 * math.pi * u[i, 2]
        eye = radius * np.array([math.sin(theta) * math.cos(phi), math.sin(theta) * math.sin(phi), math.cos(theta)])
        target = -0.1 + 0.2 * u[i, 3:6]
        s, up, f = look_at_basis(eye, target)
        out[i, 0:3] = eye
        out[i, 3:6] = s
        out[i, 6:9] = up
        out[i, 9:12] = f
        out[i, 12] = math.tan(math.radians(fov_deg) / 2.0)
        out[i, 13] = W / H
    return out


# ---------------------------------------------------------------------------
# Shell normals: 2^22 points in the shell 0.9 < |p| < 1.3, power U[6, 10]
# ---------------------------------------------------------------------------
SHELL_NORMALS = dict(n=1 << 22, seed=1, r_lo=0.9, r_hi=1.3, p_lo=6.0, p_hi=10.0, max_iters=15, bailout=2.0, fd_h=1e-3)


def shell_points(n: int, seed: int = 1, r_lo=0.9, r_hi=1.3, p_lo=6.0, p_hi=10.0):
    """Direction uniform on S^2, radius cbrt(U (r_hi^3 - r_lo^3) + r_lo^3); power U[p_lo, p_hi]."""
    u = uniform(seed, 4 * n).reshape(n, 4)
    cos_t = 1.0 - 2.0 * u[:, 0]
    sin_t = np.sqrt(np.maximum(0.0, 1.0 - cos_t * cos_t))
    phi = 2.0 * np.pi * u[:, 1]
    rad = np.cbrt(u[:, 2] * (r_hi ** 3 - r_lo ** 3) + r_lo ** 3)
    pts = np.stack([rad * sin_t * np.cos(phi), rad * sin_t * np.sin(phi), rad * cos_t], 1).astype(np.float32)
    powers = (p_lo + (p_hi - p_lo) * u[:, 3]).astype(np.float32)
    return pts, powers


# ---------------------------------------------------------------------------
# Mixed jobs: 64 independent grid jobs of mixed size, power and iteration cap
# ---------------------------------------------------------------------------
MIXED_JOBS = dict(jobs=64, seed=5, sizes=(32, 48, 64, 96, 128), p_lo=2.0, p_hi=12.0, cap_lo=8, cap_hi=30, bailout=2.0)


def mixed_jobs(jobs=64, seed=5, sizes=(32, 48, 64, 96, 128), p_lo=2.0, p_hi=12.0, cap_lo=8, cap_hi=30):
    """Per job: (N, power, max_iters).  N from `sizes`, power U[p_lo, p_hi], cap U{cap_lo..cap_hi}."""
    u = uniform(seed, 3 * jobs).reshape(jobs, 3)
    N = np.array([sizes[min(int(v * len(sizes)), len(sizes) - 1)] for v in u[:, 0]], np.uint32)
    P = (p_lo + (p_hi - p_lo) * u[:, 1]).astype(np.float32)
    cap = np.array([cap_lo + min(int(v * (cap_hi - cap_lo + 1)), cap_hi - cap_lo) for v in u[:, 2]], np.uint32)
    return N, P, cap
