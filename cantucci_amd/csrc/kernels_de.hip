// Distance-estimation kernels: explicit points, regular grids, octree-leaf
// spans with sign classification, plus the small kernels behind the Shape
// trait's derived queries.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "work_loop.h"

namespace cth {

// One wave per 64-cell word: lane l classifies cell w*64 + l of its span.
// A cell crosses the surface iff its 8 corners do not all share the IEEE sign
// bit (`is_sign_positive`, src/mesh/buffer.rs:130-141).
__global__ __launch_bounds__(256) void k_cross_bits(const float* __restrict__ dists, uint32_t words_per_span,
                                                    uint32_t total_words, uint32_t res_log2, uint64_t* cross) {
    uint32_t wave = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
    if (wave >= total_words) return;
    uint32_t lane = threadIdx.x & 63u;
    uint32_t span = wave / words_per_span;
    uint32_t word = wave - span * words_per_span;
    uint32_t res = 1u << res_log2;
    uint32_t n = res + 1;
    uint32_t cell = word * 64u + lane;
    bool crosses = false;
    if (cell < (res << (2 * res_log2))) {
        uint32_t x = cell >> (2 * res_log2);
        uint32_t y = (cell >> res_log2) & (res - 1);
        uint32_t z = cell & (res - 1);
        const float* D = dists + (size_t)span * n * n * n;
        uint32_t base = (x * n + y) * n + z;
        uint32_t s0 = __float_as_uint(D[base]) >> 31;
        uint32_t acc = 0;
#pragma unroll
        for (int c = 1; c < 8; ++c) {
            uint32_t off = ((c >> 2) & 1) * n * n + ((c >> 1) & 1) * n + (c & 1);
            acc |= (__float_as_uint(D[base + off]) >> 31) ^ s0;
        }
        crosses = acc != 0;
    }
    uint64_t bits = wave_mask(crosses);
    if (lane == 0) cross[wave] = bits;
}

__global__ void k_contains_sphere(const float* pts, uint32_t n, float cx, float cy, float cz, float radius,
                                  uint8_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // (center - p).magnitude2() <= radius * radius  (src/shape/sphere.rs:23-26)
    float dx = cx - pts[3 * (size_t)i], dy = cy - pts[3 * (size_t)i + 1], dz = cz - pts[3 * (size_t)i + 2];
    out[i] = ((dx * dx + dy * dy) + dz * dz) <= (radius * radius) ? 1 : 0;
}

__global__ void k_less_than_zero(const float* d, uint32_t n, uint8_t* out) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = d[i] < 0.0f ? 1 : 0;
}

__global__ void k_dup_pairs(const float* d, uint32_t n, float* pairs) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        float v = d[i];
        pairs[2 * (size_t)i] = v;
        pairs[2 * (size_t)i + 1] = v;
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------

static inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

int launch_de_points(Ctx& c, const cth_shape& s, const float* pts, const float* powers, uint32_t n, float* out,
                     uint32_t* iters, float* r_out, float* dr_out, hipStream_t st) {
    if (n == 0) return CTH_OK;
    SrcPoints src{pts, powers, s.power};
    if (s.kind == CTH_SHAPE_SPHERE) {
        if (r_out) return fail(CTH_E_UNSUPPORTED, "loop state requested for a sphere");
        hipLaunchKernelGGL((k_sphere<SrcPoints>), dim3(cdiv(n, 256)), dim3(256), 0, st, src, n, s.center[0],
                           s.center[1], s.center[2], s.radius, out, iters);
        return check_hip(hipGetLastError(), "sphere kernel launch");
    }
    MandelUniform u;
    int family = 0;
    int rc = resolve_mandel(s, u, family);
    if (rc != CTH_OK) return rc;
    if (r_out) {
        SinkState sink{r_out, dr_out, iters};
        return run_refill(c, family, u, src, sink, n, st, c.shape_de);
    }
    SinkDist sink{out, iters};
    return run_refill(c, family, u, src, sink, n, st, c.shape_de);
}

static bool is_pow2_float(float v, float& inv) {
    int e = 0;
    float m = frexpf(v, &e);
    if (v > 0.0f && m == 0.5f) {
        inv = ldexpf(1.0f, 1 - e);
        return true;
    }
    inv = 0.0f;
    return false;
}

int launch_de_grid(Ctx& c, const cth_shape& s, const float* start, const float* end, uint32_t n, float divisor,
                   float* out, uint32_t* iters, hipStream_t st) {
    uint64_t total = (uint64_t)n * n * n;
    if (total == 0) return CTH_OK;
    if (total >= (1ull << 31)) return fail(CTH_E_INVALID, "grid larger than 2^31 points");
    SrcGrid src;
    for (int a = 0; a < 3; ++a) {
        src.s[a] = start[a];
        src.a[a] = end[a] - start[a];
    }
    src.divisor = divisor;
    is_pow2_float(divisor, src.inv);
    src.n = make_fastdiv(n);
    src.n2 = make_fastdiv(n * n);
    src.power = s.power;
    if (s.kind == CTH_SHAPE_SPHERE) {
        hipLaunchKernelGGL((k_sphere<SrcGrid>), dim3(cdiv(total, 256)), dim3(256), 0, st, src, (uint32_t)total,
                           s.center[0], s.center[1], s.center[2], s.radius, out, iters);
        return check_hip(hipGetLastError(), "sphere grid launch");
    }
    MandelUniform u;
    int family = 0;
    int rc = resolve_mandel(s, u, family);
    if (rc != CTH_OK) return rc;
    SinkDist sink{out, iters};
    return run_refill(c, family, u, src, sink, (uint32_t)total, st, c.shape_de);
}

int launch_sample_spans(Ctx& c, const cth_shape& s, const float* spans, uint32_t nspans, uint32_t res, float* dists,
                        uint64_t* cross, uint32_t* iters, hipStream_t st) {
    if (nspans == 0) return CTH_OK;
    if (res == 0 || (res & (res - 1)) != 0) return fail(CTH_E_INVALID, "resolution must be a power of two");
    if (res > 512) return fail(CTH_E_INVALID, "resolution above 512");
    uint64_t n = res + 1;
    uint64_t per = n * n * n;
    uint64_t total = per * nspans;
    if (total >= (1ull << 31)) return fail(CTH_E_INVALID, "span batch larger than 2^31 corners");
    SrcSpans src;
    src.spans = spans;
    src.inv_res = 1.0f / (float)res;
    src.per = make_fastdiv((uint32_t)per);
    src.n = make_fastdiv((uint32_t)n);
    src.n2 = make_fastdiv((uint32_t)(n * n));
    src.power = s.power;
    int rc = CTH_OK;
    if (s.kind == CTH_SHAPE_SPHERE) {
        hipLaunchKernelGGL((k_sphere<SrcSpans>), dim3(cdiv(total, 256)), dim3(256), 0, st, src, (uint32_t)total,
                           s.center[0], s.center[1], s.center[2], s.radius, dists, iters);
        rc = check_hip(hipGetLastError(), "sphere span launch");
    } else {
        MandelUniform u;
        int family = 0;
        rc = resolve_mandel(s, u, family);
        if (rc != CTH_OK) return rc;
        SinkDist sink{dists, iters};
        rc = run_refill(c, family, u, src, sink, (uint32_t)total, st, c.shape_spans);
    }
    if (rc != CTH_OK || cross == nullptr) return rc;
    uint32_t res_log2 = 0;
    while ((1u << res_log2) < res) ++res_log2;
    uint64_t cells = (uint64_t)res * res * res;
    uint32_t words = (uint32_t)((cells + 63) / 64);
    uint64_t total_words = (uint64_t)words * nspans;
    if (total_words >= (1ull << 31)) return fail(CTH_E_INVALID, "too many crossing words");
    hipLaunchKernelGGL(k_cross_bits, dim3(cdiv(total_words, 4)), dim3(256), 0, st, dists, words,
                       (uint32_t)total_words, res_log2, cross);
    return check_hip(hipGetLastError(), "cross-bits launch");
}

int launch_contains_sphere(Ctx&, const cth_shape& s, const float* pts, uint32_t n, uint8_t* out, hipStream_t st) {
    if (n == 0) return CTH_OK;
    hipLaunchKernelGGL(k_contains_sphere, dim3(cdiv(n, 256)), dim3(256), 0, st, pts, n, s.center[0], s.center[1],
                       s.center[2], s.radius, out);
    return check_hip(hipGetLastError(), "contains launch");
}

int launch_less_than_zero(Ctx&, const float* d, uint32_t n, uint8_t* out, hipStream_t st) {
    if (n == 0) return CTH_OK;
    hipLaunchKernelGGL(k_less_than_zero, dim3(cdiv(n, 256)), dim3(256), 0, st, d, n, out);
    return check_hip(hipGetLastError(), "sign launch");
}

int launch_dup_pairs(Ctx&, const float* d, uint32_t n, float* pairs, hipStream_t st) {
    if (n == 0) return CTH_OK;
    hipLaunchKernelGGL(k_dup_pairs, dim3(cdiv(n, 256)), dim3(256), 0, st, d, n, pairs);
    return check_hip(hipGetLastError(), "pairs launch");
}

}  // namespace cth
