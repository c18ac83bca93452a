// Point sources and result sinks for the lane-refill work loop.
//
// Grid and span sources generate coordinates from the index (4 output bytes
// per point instead of 16 with explicit points), using the reference's corner
// formula p = start + across * (i / res) (src/mesh/buffer.rs:78-80).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "de_device.h"

namespace cth {

// Division of n < 2^31 by a runtime constant d via a 32-bit multiply-high:
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = 2^32 (2^s - d) / d + 1
// (the round-up method of Granlund and Montgomery, "Division by invariant
// integers using multiplication", PLDI 1994).
struct FastDiv {
    uint32_t d, m, s;
};

inline FastDiv make_fastdiv(uint32_t d) {
    FastDiv f;
    f.d = d;
    uint32_t s = 0;
    while ((1ull << s) < d) ++s;
    f.s = s;
    f.m = (uint32_t)((((1ull << s) - d) << 32) / d + 1);
    if (d == 1) f.m = 0;  // umulhi(n, 0) + n == n
    return f;
}

__device__ __forceinline__ uint32_t fdiv(uint32_t n, const FastDiv& f) {
    return (__umulhi(n, f.m) + n) >> f.s;
}

// ---------------------------------------------------------------------------
// Sources
// ---------------------------------------------------------------------------

// Explicit AoS points (cgmath Point3<f32>, 12 bytes) with an optional
// per-point power.
struct SrcPoints {
    const float* pts;
    const float* powers;
    float power;
    __device__ __forceinline__ void get(uint32_t i, float& x, float& y, float& z, float& p) const {
        x = pts[3 * (size_t)i];
        y = pts[3 * (size_t)i + 1];
        z = pts[3 * (size_t)i + 2];
        p = powers ? powers[i] : power;
    }
};

// n^3 grid, GridTable order: p = s + a * (i / divisor) per axis, with an
// IEEE division (or an exact multiply when divisor is a power of two).
struct SrcGrid {
    float s[3], a[3];
    float divisor, inv;  // inv != 0 only when divisor is a power of two
    FastDiv n, n2;
    float power;
    __device__ __forceinline__ float coord(int ax, uint32_t i) const {
        float v = (inv != 0.0f) ? (float)i * inv : div_rn((float)i, divisor);
        return s[ax] + a[ax] * v;
    }
    __device__ __forceinline__ void get(uint32_t idx, float& x, float& y, float& z, float& p) const {
        uint32_t i = fdiv(idx, n2);
        uint32_t rem = idx - i * n2.d;
        uint32_t j = fdiv(rem, n);
        uint32_t k = rem - j * n.d;
        x = coord(0, i);
        y = coord(1, j);
        z = coord(2, k);
        p = power;
    }
};

// Octree-leaf spans sampled like naive surface nets step 1
// (src/mesh/buffer.rs:64-83): the span grows by one cell on each side,
// overflow = (e - s) / res; s' = s + -overflow; e' = e + overflow; then
// corner (i,j,k) = s' + (e' - s') * ((i,j,k) / res).  res is a power of two
// (asserted by the reference, :39) so `/ res` is an exact multiply by 1/res.
struct SrcSpans {
    const float* spans;  // 6 floats per span
    float inv_res;
    FastDiv per, n, n2;  // (res+1)^3, res+1, (res+1)^2
    float power;
    __device__ __forceinline__ void get(uint32_t idx, float& x, float& y, float& z, float& p) const {
        uint32_t leaf = fdiv(idx, per);
        uint32_t local = idx - leaf * per.d;
        uint32_t i = fdiv(local, n2);
        uint32_t rem = local - i * n2.d;
        uint32_t j = fdiv(rem, n);
        uint32_t k = rem - j * n.d;
        const float* sp = spans + 6 * (size_t)leaf;
        uint32_t ijk[3] = {i, j, k};
        float c[3];
#pragma unroll
        for (int ax = 0; ax < 3; ++ax) {
            float s0 = sp[ax], e0 = sp[3 + ax];
            float of = (e0 - s0) * inv_res;
            float s1 = s0 + -of;
            float e1 = e0 + of;
            float across = e1 - s1;
            c[ax] = s1 + across * ((float)ijk[ax] * inv_res);
        }
        x = c[0];
        y = c[1];
        z = c[2];
        p = power;
    }
};

// ---------------------------------------------------------------------------
// Sinks.  wants_k(): whether put() reads the loop-body count k (the work
// loop only keeps a finished point's count for sinks that do).
// ---------------------------------------------------------------------------
struct SinkDist {
    float* out;
    uint32_t* iters;
    __device__ __forceinline__ bool wants_k() const { return iters != nullptr; }
    template <class E>
    __device__ __forceinline__ void put(uint32_t i, float d, uint32_t k, float, float, const E&) const {
        out[i] = d;
        if (iters) iters[i] = k;
    }
};

struct SinkState {
    float* r;
    float* dr;
    uint32_t* iters;
    __device__ __forceinline__ bool wants_k() const { return true; }
    template <class E>
    __device__ __forceinline__ void put(uint32_t i, float, uint32_t k, float rr, float drr, const E&) const {
        r[i] = rr;
        dr[i] = drr;
        iters[i] = k;
    }
};

// Gradient sink for the tangent evaluators (tangent.h): distance, unit normal,
// grad_d-scaled gradient with respect to p and the power.
struct SinkGrad {
    float* d;
    float* normals;
    float* grad_p;
    float* grad_power;
    const float* grad_d;
    uint32_t max_iters;
    __device__ __forceinline__ bool wants_k() const { return true; }
    template <class E>
    __device__ __forceinline__ void put(uint32_t i, float dv, uint32_t k, float r, float, const E& ev) const {
        float g[4];
        ev.grad(r, k == max_iters, g);
        if (d) d[i] = dv;
        if (normals) {
            float m = __builtin_sqrtf(__builtin_fmaf(g[0], g[0], __builtin_fmaf(g[1], g[1], g[2] * g[2])));
            float im = 1.0f / m;
            normals[3 * (size_t)i] = g[0] * im;
            normals[3 * (size_t)i + 1] = g[1] * im;
            normals[3 * (size_t)i + 2] = g[2] * im;
        }
        float s = grad_d ? grad_d[i] : 1.0f;
        if (grad_p) {
            grad_p[3 * (size_t)i] = g[0] * s;
            grad_p[3 * (size_t)i + 1] = g[1] * s;
            grad_p[3 * (size_t)i + 2] = g[2] * s;
        }
        if (grad_power) grad_power[i] = g[3] * s;
    }
};

}  // namespace cth
