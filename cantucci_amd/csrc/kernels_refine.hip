// Level-synchronous octree refinement on the GPU.
//
// The reference refines a CPU octree one leaf at a time (ShapeMesh::update,
// src/mesh/mod.rs:82-178, child spans from create_spans,
// src/octree/mod.rs:315-329).  Here a whole level is one batch: the centre
// distance of every node at depth d is evaluated by the lane-refill work loop,
// the split flags are prefix-summed, and every split node writes its 8
// children (create_spans order) into the next level's array.  Parents stay in
// level order, so the level-d array equals the oracle's breadth-first order.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <vector>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "work_loop.h"

namespace cth {

// Node centre: start + (end - start) / 2 (SpanExt::center, src/octree/mod.rs:21-23);
// halving is exact, so `* 0.5f` equals the reference's `/ 2.0`.
struct SrcCenters {
    const float* spans;
    float power;
    __device__ __forceinline__ void get(uint32_t i, float& x, float& y, float& z, float& p) const {
        const float* sp = spans + 6 * (size_t)i;
        x = sp[0] + (sp[3] - sp[0]) * 0.5f;
        y = sp[1] + (sp[4] - sp[1]) * 0.5f;
        z = sp[2] + (sp[5] - sp[2]) * 0.5f;
        p = power;
    }
};

// Split rule of the refinement policy: |DE(centre)| < half diagonal, or NaN.
struct SinkSplit {
    uint32_t* flags;
    float half_diag;
    __device__ __forceinline__ bool wants_k() const { return false; }
    template <class E>
    __device__ __forceinline__ void put(uint32_t i, float d, uint32_t, float, float, const E&) const {
        flags[i] = ((d != d) || (fabsf(d) < half_diag)) ? 1u : 0u;
    }
};

__global__ void k_sphere_split(const float* nodes, uint32_t n, float cx, float cy, float cz, float radius,
                               float half_diag, uint32_t* flags) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    SrcCenters src{nodes, 0.0f};
    float x, y, z, p;
    src.get(i, x, y, z, p);
    const float c[3] = {cx, cy, cz};
    float d = sphere_de(x, y, z, c, radius);
    flags[i] = ((d != d) || (fabsf(d) < half_diag)) ? 1u : 0u;
}

__global__ void k_fill_u32(uint32_t* p, uint32_t n, uint32_t v) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Fill kernel for code that may be captured into a graph (a recorded memset
// node is avoided there).
int launch_fill_u32(uint32_t* p, size_t n, uint32_t v, hipStream_t st) {
    const size_t step = size_t(1) << 30;
    for (size_t off = 0; off < n; off += step) {
        uint32_t m = (uint32_t)((n - off) < step ? (n - off) : step);
        hipLaunchKernelGGL(k_fill_u32, dim3((m + 255) / 256), dim3(256), 0, st, p + off, m, v);
        int rc = check_hip(hipGetLastError(), "fill launch");
        if (rc) return rc;
    }
    return CTH_OK;
}

// One thread per (node, child): child c of a split node i lands at
// 8 * (incl[i] - 1) + c, incl being the inclusive prefix sum of the split
// flags.  create_spans (src/octree/mod.rs:315-329): centre
// = start + (end - start) / 2; child bit 2 = +x half, bit 1 = +y, bit 0 = +z.
__global__ void k_emit_children(const float* __restrict__ nodes, const uint32_t* __restrict__ flags,
                                const uint32_t* __restrict__ incl, uint32_t n, float* __restrict__ children) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t i = t >> 3, c = t & 7u;
    if (i >= n || !flags[i]) return;
    const float* sp = nodes + 6 * (size_t)i;
    float* out = children + 6 * ((size_t)(incl[i] - 1u) * 8 + c);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        float s = sp[a], e = sp[3 + a];
        float ctr = s + (e - s) * 0.5f;
        bool hi = (c >> (2 - a)) & 1u;
        out[a] = hi ? ctr : s;
        out[3 + a] = hi ? e : ctr;
    }
}

__global__ void k_emit_leaves(const float* __restrict__ nodes, const uint32_t* __restrict__ flags,
                              const uint32_t* __restrict__ incl, uint32_t n, float* __restrict__ leaves,
                              uint64_t cap) {
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t i = t / 6, a = t - i * 6;
    if (i >= n || !flags[i] || incl[i] - 1u >= cap) return;
    leaves[6 * (size_t)(incl[i] - 1u) + a] = nodes[6 * (size_t)i + a];
}

static inline uint32_t cdiv32(uint64_t a, uint64_t b) { return (uint32_t)((a + b - 1) / b); }

// Level-synchronous refinement.  Per level: the split flags (forced, sphere
// closed form or a DE refill launch), an inclusive scan of the flags, one
// 4-byte read-back of the split count into pinned host memory, and the
// children.  Node, flag, offset and scan buffers are the context's scratch
// slots (ping-pong for the nodes), so a level allocates nothing.
int refine_levels(Ctx& c, const cth_shape& s, const float* root, uint32_t min_depth, uint32_t max_depth,
                  const float* half_diag, float* leaves, size_t cap, size_t* nleaves, uint64_t* level_counts,
                  hipStream_t st) {
    MandelUniform u{};
    int family = FAM_SPHERE;
    if (s.kind != CTH_SHAPE_SPHERE) {
        int rc = resolve_mandel(s, u, family);
        if (rc) return rc;
    }
    const uint64_t kMaxNodes = 1ull << 28;
    float* cur = nullptr;
    int rc = c.scratch(WS_NODES_A, 6 * sizeof(float), (void**)&cur);
    if (rc) return fail(rc, "refinement scratch allocation failed");
    rc = check_hip(hipMemcpyAsync(cur, root, 6 * sizeof(float), hipMemcpyHostToDevice, st), "H2D root");
    int slot = WS_NODES_A;
    uint64_t n = 1;
    *nleaves = 0;
    for (uint32_t depth = 0; depth <= max_depth && rc == CTH_OK; ++depth) {
        if (level_counts) level_counts[depth] = n;
        if (n == 0) {
            for (uint32_t d2 = depth + 1; d2 <= max_depth && level_counts; ++d2) level_counts[d2] = 0;
            break;
        }
        uint32_t* flags = nullptr;
        uint32_t* incl = nullptr;
        if (c.scratch(WS_FLAGS, n * sizeof(uint32_t), (void**)&flags) ||
            c.scratch(WS_OFFS, n * sizeof(uint32_t), (void**)&incl)) {
            rc = fail(CTH_E_NOMEM, "refinement scratch allocation failed");
            break;
        }
        if (depth < min_depth) {
            hipLaunchKernelGGL(k_fill_u32, dim3(cdiv32(n, 256)), dim3(256), 0, st, flags, (uint32_t)n, 1u);
            rc = check_hip(hipGetLastError(), "fill launch");
        } else if (family == FAM_SPHERE) {
            hipLaunchKernelGGL(k_sphere_split, dim3(cdiv32(n, 256)), dim3(256), 0, st, cur, (uint32_t)n, s.center[0],
                               s.center[1], s.center[2], s.radius, half_diag[depth], flags);
            rc = check_hip(hipGetLastError(), "sphere split launch");
        } else {
            SrcCenters src{cur, s.power};
            SinkSplit sink{flags, half_diag[depth]};
            // A level has at most a few million centres: 64-point chunks
            // dealt statically spread it over the whole chip without the
            // claim atomics that a dynamic queue would serialise on.
            LaunchShape ls = c.shape_spans;
            ls.chunk = 64;
            ls.dynamic = 0;
            rc = run_refill(c, family, u, src, sink, (uint32_t)n, st, ls);
        }
        if (rc) break;
        size_t temp_bytes = 0;
        rc = check_hip(hipcub::DeviceScan::InclusiveSum(nullptr, temp_bytes, flags, incl, (int)n, st), "scan size");
        void* temp = nullptr;
        if (!rc && c.scratch(WS_SCAN, temp_bytes, &temp)) rc = fail(CTH_E_NOMEM, "scan scratch allocation failed");
        if (!rc) rc = check_hip(hipcub::DeviceScan::InclusiveSum(temp, temp_bytes, flags, incl, (int)n, st), "scan");
        if (!rc)
            rc = check_hip(hipMemcpyAsync(c.pinned, incl + (n - 1), sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                           "D2H split count");
        if (!rc) rc = check_hip(hipStreamSynchronize(st), "sync");
        if (rc) break;
        const uint64_t split = *reinterpret_cast<volatile uint32_t*>(c.pinned);
        if (depth == max_depth) {
            *nleaves = split;
            if (leaves && split) {
                hipLaunchKernelGGL(k_emit_leaves, dim3(cdiv32(n * 6, 256)), dim3(256), 0, st, cur, flags, incl,
                                   (uint32_t)n, leaves, (uint64_t)cap);
                rc = check_hip(hipGetLastError(), "leaves launch");
            }
            break;
        }
        const uint64_t next_n = split * 8;
        if (next_n > kMaxNodes) {
            rc = fail(CTH_E_INVALID, "refinement exceeds 2^28 nodes in one level");
            break;
        }
        // The next level goes to the other node slot; growing that slot
        // waits for the device, so `cur` stays valid.
        const int next_slot = (slot == WS_NODES_A) ? WS_NODES_B : WS_NODES_A;
        float* next = nullptr;
        if (c.scratch(next_slot, next_n * 6 * sizeof(float) + 4, (void**)&next)) {
            rc = fail(CTH_E_NOMEM, "refinement scratch allocation failed");
            break;
        }
        if (next_n) {
            hipLaunchKernelGGL(k_emit_children, dim3(cdiv32(n * 8, 256)), dim3(256), 0, st, cur, flags, incl,
                               (uint32_t)n, next);
            rc = check_hip(hipGetLastError(), "children launch");
        }
        cur = next;
        slot = next_slot;
        n = next_n;
    }
    int rs = check_hip(hipStreamSynchronize(st), "sync");
    return rc ? rc : rs;
}

}  // namespace cth
