// Launch plumbing shared by the kernel units: the lane-refill kernel
// template and the persistent-grid sizing.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include "internal.h"
#include "work_loop.h"

namespace cth {

// Ev::kWavesPerSimd: the occupancy the evaluator's register budget is fitted
// to (the P8 evaluators fit 64 VGPRs, 8 waves per SIMD, without spilling).
// kCount: the counting instantiation (work tallies, see WorkTally).
template <class Ev, class Src, class Sink, bool kCount = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(Ev::kWavesPerSimd))) void k_refill(
    Src src, MandelUniform u, Sink sink, WorkQueue q) {
    refill_loop<Ev, kCount>(src, u, sink, q);
}

// Sphere: no iteration, one point per thread.
template <class Src>
__global__ __launch_bounds__(256) void k_sphere(Src src, uint32_t n, float cx, float cy, float cz, float radius,
                                                float* out, uint32_t* iters) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x, y, z, p;
    src.get(i, x, y, z, p);
    const float c[3] = {cx, cy, cz};
    out[i] = sphere_de(x, y, z, c, radius);
    if (iters) iters[i] = 0;
}

inline WorkQueue make_queue(Ctx& c, uint32_t total, const LaunchShape& ls, uint32_t& blocks, hipStream_t st) {
    uint32_t waves_per_block = (uint32_t)ls.block / 64u;
    uint64_t want = (uint64_t)c.num_cus * (uint64_t)ls.blocks_per_cu;
    uint64_t need = ((uint64_t)total + (uint64_t)ls.chunk * waves_per_block - 1) / ((uint64_t)ls.chunk * waves_per_block);
    blocks = (uint32_t)std::max<uint64_t>(1, std::min(want, need));
    WorkQueue q;
    uint32_t* pair = c.take_counters(st);
    q.next = pair;
    q.done = pair + 1;
    q.waves = blocks * waves_per_block;
    q.total = total;
    // At least ~8 chunks per wave, so that small launches still balance:
    // chunk = clamp(total / (8 waves), 64, configured), a multiple of 64.
    uint64_t fair = (uint64_t)total / ((uint64_t)q.waves * 8u);
    uint64_t ch = std::min<uint64_t>(ls.chunk, std::max<uint64_t>(64, fair));
    q.chunk = (uint32_t)(ch & ~63ull);
    q.refill_min = ls.refill_min < 1 ? 1 : (ls.refill_min > 64 ? 64 : ls.refill_min);
    q.dynamic = ls.dynamic ? 1u : 0u;
    q.stats = c.stats;
    return q;
}

template <class Src, class Sink>
inline int run_refill(Ctx& c, int family, const MandelUniform& u, const Src& src, const Sink& sink,
                      uint32_t total, hipStream_t st, const LaunchShape& ls) {
    uint32_t blocks = 0;
    WorkQueue q = make_queue(c, total, ls, blocks, st);
    dim3 grid(blocks), block(ls.block);
    if (family == FAM_P8) {
        with_p8(c.mode, [&](auto ev) {
            using Ev = decltype(ev);
            if (q.stats)
                hipLaunchKernelGGL((k_refill<Ev, Src, Sink, true>), grid, block, 0, st, src, u, sink, q);
            else
                hipLaunchKernelGGL((k_refill<Ev, Src, Sink>), grid, block, 0, st, src, u, sink, q);
        });
    } else if (q.stats) {
        hipLaunchKernelGGL((k_refill<TrigExact, Src, Sink, true>), grid, block, 0, st, src, u, sink, q);
    } else {
        hipLaunchKernelGGL((k_refill<TrigExact, Src, Sink>), grid, block, 0, st, src, u, sink, q);
    }
    return check_hip(hipGetLastError(), "refill kernel launch");
}

}  // namespace cth
