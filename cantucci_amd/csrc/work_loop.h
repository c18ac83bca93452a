// Persistent lane-refill work loop for wave64.
//
// The Mandelbulb iteration count varies wildly between neighbouring points
// (on the 64^3 grid 73% of points bail after one body and 8% run all 15), so
// a wave that gives each lane one point idles most lanes while its slowest
// point finishes.  Here every lane owns one point at a time; when its point
// finishes (bailout or iteration cap) the lane writes the result and takes
// the next index from the wave's chunk, so all lanes keep executing loop
// bodies.  Chunks of `chunk` consecutive indices are taken from a global
// counter with one atomic per wave; the last wave to leave resets the
// counters, so a launch leaves them at zero (safe for graph replay).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "de_device.h"

namespace cth {

struct WorkQueue {
    uint32_t* next;    // next unclaimed index
    uint32_t* done;    // waves that have left the loop
    uint32_t waves;    // waves in the launch
    uint32_t total;    // indices in [0, total)
    uint32_t chunk;    // indices claimed per atomic
};

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One atomic per wave; the value is broadcast from the first active lane.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* ctr, uint32_t amount) {
    uint64_t active = __ballot(1);
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t v = 0;
    if ((int)__lane_id() == leader) v = atomicAdd(ctr, amount);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, leader);
}

// Called once per wave after its last claim; the final wave zeroes both
// counters.  Every other wave has already made its final claim by then.
__device__ __forceinline__ void wave_retire(const WorkQueue& q) {
    uint64_t active = __ballot(1);
    int leader = __ffsll((unsigned long long)active) - 1;
    if ((int)__lane_id() == leader) {
        uint32_t prev = atomicAdd(q.done, 1u);
        if (prev == q.waves - 1) {
            atomicExch(q.next, 0u);
            atomicExch(q.done, 0u);
        }
    }
}

// Drive evaluator `Ev` over the indices of `q`, reading points from `src`
// and writing results through `sink`.
//   src.get(i, x, y, z, power)
//   sink.put(i, d, iters, r, dr)
template <class Ev, class Src, class Sink>
__device__ __forceinline__ void refill_loop(const Src& src, const MandelUniform& u, const Sink& sink,
                                            const WorkQueue& q) {
    Ev ev;
    bool has = false;  // the lane holds a point that needs another loop body
    uint32_t idx = 0, k = 0;
    float r = 0.0f;
    uint32_t c_next = 0, c_end = 0;  // wave-uniform window of claimed indices
    bool exhausted = false;

    for (;;) {
        // Hand new points to empty lanes.  A point that bails before its first
        // body (|p| > bailout) is finished on the spot and the lane refilled.
        while (!exhausted) {
            uint64_t need = __ballot(!has);
            if (need == 0) break;
            if (c_next >= c_end) {
                uint32_t b = wave_claim(q.next, q.chunk);
                if (b >= q.total) {
                    exhausted = true;
                    break;
                }
                c_next = b;
                c_end = (q.total - b < q.chunk) ? q.total : b + q.chunk;
            }
            uint32_t avail = c_end - c_next;
            uint32_t nneed = (uint32_t)__popcll(need);
            if (!has) {
                uint32_t rank = lane_rank(need);
                if (rank < avail) {
                    idx = c_next + rank;
                    float px, py, pz, pw;
                    src.get(idx, px, py, pz, pw);
                    ev.start(px, py, pz, pw, u);
                    k = 0;
                    r = ev.mag();
                    if (r > u.bailout) {
                        sink.put(idx, ev.finish(r), 0u, r, ev.dr);
                    } else {
                        has = true;
                    }
                }
            }
            c_next += (nneed < avail) ? nneed : avail;
        }
        if (__ballot(has) == 0) break;
        if (has) {
            ev.body(r, u);
            ++k;
            bool done = (k == u.max_iters);
            if (!done) {
                float rn = ev.mag();
                if (rn > u.bailout) done = true;
                r = rn;
            }
            // At the iteration cap, r is the magnitude computed before the last
            // body (the reference's off-by-one, src/shape/mandelbulb.rs:65-78).
            if (done) {
                sink.put(idx, ev.finish(r), k, r, ev.dr);
                has = false;
            }
        }
    }
    wave_retire(q);
}

}  // namespace cth
