// The reference's own refinement policy, batched: ShapeMesh::update splits
// the leaf around each focus point when the camera is nearer to that point
// than twice the leaf's width (src/mesh/mod.rs:91-104); the focus points are
// the hits of the focus rays of get_focii (src/mesh/mod.rs:205-243).
//
// A set of leaves that tiles the octree is given as a flat array of spans.
// The leaf around a point -- Octree::leaf_around_mut, a descent through
// half-open child spans (src/octree/mod.rs:25-31, 75-89) -- is the one leaf
// whose half-open span contains it, since create_spans children share their
// parent's centre exactly.  One thread per leaf tests every focus point; the
// split leaves are then replaced in place by their 8 children.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "internal.h"

namespace cth {

namespace {

// SpanExt::contains (src/octree/mod.rs:25-31): s <= p < e on every axis.
__device__ __forceinline__ bool span_contains(const float* sp, float x, float y, float z) {
    return sp[0] <= x && sp[1] <= y && sp[2] <= z && x < sp[3] && y < sp[4] && z < sp[5];
}

// flags[i] = 1 when some focus point lies in leaf i and
// |eye - focus| < 2 |e.x - s.x|  (cgmath distance: sqrt((dx^2 + dy^2) + dz^2)).
// A focus point with a non-finite depth (a ray that missed; get_focii drops
// those) is skipped.
__global__ void k_focus_split(const float* __restrict__ leaves, uint32_t n, const float* __restrict__ focii,
                              const float* __restrict__ depth, uint32_t nf, float ex, float ey, float ez,
                              uint32_t* __restrict__ flags) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* sp = leaves + 6 * (size_t)i;
    const float threshold = 2.0f * fabsf(sp[3] - sp[0]);
    uint32_t split = 0;
    for (uint32_t f = 0; f < nf; ++f) {
        if (depth && !__builtin_isfinite(depth[f])) continue;
        const float fx = focii[3 * f], fy = focii[3 * f + 1], fz = focii[3 * f + 2];
        if (!span_contains(sp, fx, fy, fz)) continue;
        const float dx = ex - fx, dy = ey - fy, dz = ez - fz;
        const float dist = sqrt_rn((dx * dx + dy * dy) + dz * dz);
        if (dist < threshold) split = 1;
    }
    flags[i] = split;
}

// Leaf i becomes 8 children at 8 * ... or stays one leaf; `incl` is the
// inclusive prefix sum of the output counts (1 or 8).  Children follow
// create_spans (src/octree/mod.rs:315-329): centre = s + (e - s) / 2, child
// bit 2 = +x half, bit 1 = +y, bit 0 = +z.
__global__ void k_split_in_place(const float* __restrict__ leaves, const uint32_t* __restrict__ flags,
                                 const uint32_t* __restrict__ incl, uint32_t n, float* __restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = t >> 3, c = t & 7u;
    if (i >= n) return;
    const float* sp = leaves + 6 * (size_t)i;
    const uint32_t cnt = flags[i] ? 8u : 1u;
    if (c >= cnt) return;
    float* o = out + 6 * ((size_t)(incl[i] - cnt) + c);
    if (cnt == 1u) {
#pragma unroll
        for (int a = 0; a < 6; ++a) o[a] = sp[a];
        return;
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float s = sp[a], e = sp[3 + a];
        const float ctr = s + (e - s) * 0.5f;
        const bool hi = (c >> (2 - a)) & 1u;
        o[a] = hi ? ctr : s;
        o[3 + a] = hi ? e : ctr;
    }
}

__global__ void k_out_counts(const uint32_t* __restrict__ flags, uint32_t n, uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) counts[i] = flags[i] ? 8u : 1u;
}

inline uint32_t cdiv(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

}  // namespace

int launch_focus_split(Ctx&, const float* leaves, uint32_t n, const float* focii, const float* depth, uint32_t nf,
                       const float eye[3], uint32_t* flags, hipStream_t st) {
    if (n == 0) return CTH_OK;
    hipLaunchKernelGGL(k_focus_split, dim3(cdiv(n, 256)), dim3(256), 0, st, leaves, n, focii, depth, nf, eye[0], eye[1],
                       eye[2], flags);
    return check_hip(hipGetLastError(), "focus split launch");
}

int split_leaves(Ctx& c, const float* leaves, uint32_t n, const uint32_t* flags, float* out, size_t cap,
                 size_t* nout, hipStream_t st) {
    *nout = 0;
    if (n == 0) return CTH_OK;
    uint32_t* counts = nullptr;
    uint32_t* incl = nullptr;
    if (c.scratch(WS_FLAGS, n * sizeof(uint32_t), (void**)&counts) ||
        c.scratch(WS_OFFS, n * sizeof(uint32_t), (void**)&incl))
        return fail(CTH_E_NOMEM, "split scratch allocation failed");
    hipLaunchKernelGGL(k_out_counts, dim3(cdiv(n, 256)), dim3(256), 0, st, flags, n, counts);
    int rc = check_hip(hipGetLastError(), "split counts launch");
    size_t temp_bytes = 0;
    void* temp = nullptr;
    if (!rc) rc = check_hip(hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, counts, incl, (int)n, st), "scan size");
    if (!rc && c.scratch(WS_SCAN, temp_bytes, &temp)) rc = fail(CTH_E_NOMEM, "scan scratch allocation failed");
    if (!rc) rc = check_hip(hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, counts, incl, (int)n, st), "scan");
    if (!rc) rc = check_hip(hipMemcpyAsync(c.pinned, incl + (n - 1), 4, hipMemcpyDeviceToHost, st), "D2H count");
    if (!rc) rc = check_hip(hipStreamSynchronize(st), "sync");
    if (rc) return rc;
    const size_t total = *reinterpret_cast<volatile uint32_t*>(c.pinned);
    *nout = total;
    if (!out) return CTH_OK;
    if (cap < total) return fail(CTH_E_INVALID, "output holds fewer spans than the split produces");
    hipLaunchKernelGGL(k_split_in_place, dim3(cdiv((uint64_t)n * 8, 256)), dim3(256), 0, st, leaves, flags, incl, n,
                       out);
    return check_hip(hipGetLastError(), "split launch");
}

}  // namespace cth
