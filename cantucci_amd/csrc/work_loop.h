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
    uint32_t refill_min;  // idle lanes that trigger a write-back + refill (1..64)
    uint32_t dynamic;     // 1: claim chunks with an atomic; 0: static interleave
    // Work counters (nullable): the launch adds its loop bodies, its
    // evaluations that ended by bailout, and its evaluations to
    // stats[0], stats[1], stats[2] (the counting instantiations only).
    unsigned long long* stats;
};

// Per-wave work tally of a refill kernel (wave-uniform, scalar registers):
// lane masks of the lanes that ran a loop body, that ended an evaluation by
// bailout, that started an evaluation.  Only the counting instantiation of a
// kernel (kOn, launched while a context counts work) tallies: even scalar
// adds in the tick cost measurable time (11% on the span kernel), and the
// counts depend only on the inputs, not on which instantiation ran.
template <bool kOn>
struct WorkTally {
    uint64_t bodies = 0, bails = 0, evals = 0;
    __device__ __forceinline__ void body(uint64_t m) {
        if constexpr (kOn) bodies += (uint64_t)__popcll(m);
    }
    __device__ __forceinline__ void bail(uint64_t m) {
        if constexpr (kOn) bails += (uint64_t)__popcll(m);
    }
    __device__ __forceinline__ void eval(uint64_t m) {
        if constexpr (kOn) evals += (uint64_t)__popcll(m);
    }
    __device__ __forceinline__ void flush(unsigned long long* stats) const {
        if constexpr (kOn) {
            if (!stats) return;
            const uint64_t active = wave_mask(true);
            if ((int)__lane_id() == __ffsll((unsigned long long)active) - 1) {
                atomicAdd(stats, (unsigned long long)bodies);
                atomicAdd(stats + 1, (unsigned long long)bails);
                atomicAdd(stats + 2, (unsigned long long)evals);
            }
        }
    }
};

__device__ __forceinline__ uint32_t lane_rank(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// One atomic per wave; the value is broadcast from the first active lane.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* ctr, uint32_t amount) {
    uint64_t active = wave_mask(1);
    int leader = __ffsll((unsigned long long)active) - 1;
    uint32_t v = 0;
    if ((int)__lane_id() == leader) v = atomicAdd(ctr, amount);
    return (uint32_t)__builtin_amdgcn_readlane((int)v, leader);
}

// Wave index within the launch (read from the first lane so that the compiler
// knows it is wave-uniform and keeps the chunk window in scalar registers).
__device__ __forceinline__ uint32_t wave_id() {
    return blockIdx.x * (blockDim.x >> 6) + (uint32_t)__builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// Next chunk of the wave: either the next unclaimed chunk of the global
// counter (dynamic) or, statically, chunks w, w + W, w + 2W, ... of the
// index space (W waves), which needs no atomics and spreads every wave over
// the whole domain so that slow and fast regions average out.
__device__ __forceinline__ uint32_t next_chunk(const WorkQueue& q, uint32_t& round) {
    if (q.dynamic) return wave_claim(q.next, q.chunk);
    uint32_t b = (round * q.waves + wave_id()) * q.chunk;
    ++round;
    return b;
}

// Called once per wave after its last claim; the final wave zeroes both
// counters.  Every other wave has already made its final claim by then.
__device__ __forceinline__ void wave_retire(const WorkQueue& q) {
    uint64_t active = wave_mask(1);
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
//   sink.put(i, d, iters, r, dr, ev)  (ev: the evaluator, for derivative sinks)
//
// Each tick runs one loop body on every lane (results committed where the
// lane's point is live).  A lane whose point finishes holds (r, dr) until at
// least `q.refill_min` lanes of the wave are idle; then the held points are
// written and the idle lanes take new indices.  Batching the refill keeps the
// per-tick control flow to a few scalar instructions.
//
// Which lanes are live or hold a finished point is kept in wave-uniform lane
// masks (scalar registers).  A compare's lane mask is free to combine with
// them, and the per-lane view is the same mask
// (__builtin_amdgcn_inverse_ballot_w64).  A point's body count is the wave's
// tick counter (scalar) minus the tick at which the lane took the point, so
// the iteration cap costs one compare per tick and no per-lane counter.
//
// Evaluators with Ev::kLastDr end a point that reaches the iteration cap
// after max_iters - 1 bodies: of its max_iters-th body only the dr update is
// live (Ev::last_dr), and it is applied when the point is written (`last_m`
// marks those lanes).  r is then the magnitude that body would have read,
// which is the magnitude the reference ends with.
template <class Ev, bool kCount, class Src, class Sink>
__device__ __forceinline__ void refill_loop(const Src& src, const MandelUniform& u, const Sink& sink,
                                            const WorkQueue& q) {
    Ev ev;
    const uint64_t wave = wave_mask(true);  // the wave's lanes
    uint64_t live_m = 0;  // lanes whose point needs another loop body
    uint64_t hold_m = 0;  // lanes holding a finished point not yet written
    uint64_t last_m = 0;  // held points whose last body's dr update is due (kLastDr)
    uint32_t tick = 0;    // loop bodies run by the wave
    uint32_t idx = 0;
    uint32_t t0 = 0;      // tick at which the lane took its point
    uint32_t kh = 0;      // body count of a held point (kept if the sink reads it)
    float r = 0.0f;
    uint32_t c_next = 0, c_end = 0;  // wave-uniform window of claimed indices
    uint32_t round = 0;
    bool exhausted = false;
    WorkTally<kCount> tally;

    for (;;) {
        if ((uint32_t)__popcll(wave & ~live_m) >= q.refill_min) {
            if (__builtin_amdgcn_inverse_ballot_w64(hold_m)) {
                if constexpr (Ev::kLastDr) {
                    if (__builtin_amdgcn_inverse_ballot_w64(last_m)) ev.last_dr(r);
                }
                sink.put(idx, ev.finish(r), kh, r, ev.dr, ev);
            }
            hold_m = 0;
            last_m = 0;
            // Hand new points to idle lanes.  A point that bails before its
            // first body (|p| > bailout) is written on the spot and the lane
            // refilled again.
            while (!exhausted) {
                const uint64_t need = wave & ~live_m;
                if (need == 0) break;
                if (c_next >= c_end) {
                    uint32_t b = next_chunk(q, round);
                    if (b >= q.total) {
                        exhausted = true;
                        break;
                    }
                    c_next = b;
                    c_end = (q.total - b < q.chunk) ? q.total : b + q.chunk;
                }
                const uint32_t avail = c_end - c_next;
                const uint32_t nneed = (uint32_t)__popcll(need);
                const uint32_t rank = lane_rank(need);
                const bool take = __builtin_amdgcn_inverse_ballot_w64(need) && rank < avail;
                if (take) {
                    idx = c_next + rank;
                    float px, py, pz, pw;
                    src.get(idx, px, py, pz, pw);
                    ev.start(px, py, pz, pw, u);
                    t0 = tick;
                    r = ev.mag();
                }
                const uint64_t took = wave_mask(take);
                const uint64_t bailed = wave_mask(r > u.bailout) & took;
                tally.eval(took);
                tally.bail(bailed);
                // With kLastDr and max_iters == 1 a point that does not bail
                // at once is at its cap before any full body.
                const uint64_t capped = (Ev::kLastDr && u.max_iters == 1u) ? took & ~bailed : 0;
                tally.body(capped);
                if (__builtin_amdgcn_inverse_ballot_w64(bailed | capped)) {
                    const bool c = __builtin_amdgcn_inverse_ballot_w64(capped);
                    if constexpr (Ev::kLastDr) {
                        if (c) ev.last_dr(r);
                    }
                    sink.put(idx, ev.finish(r), c ? 1u : 0u, r, ev.dr, ev);
                }
                live_m |= took & ~(bailed | capped);
                c_next += (nneed < avail) ? nneed : avail;
            }
            if (live_m == 0) break;
            tick_guard(ev, u);  // the started points' squares
        }
        const bool live = __builtin_amdgcn_inverse_ballot_w64(live_m);
        uint64_t done_m;
        if constexpr (Ev::kLastDr) {
            tick_body(ev, r, u, live);
            ++tick;
            tally.body(live_m);
            const float rn = tick_mag(ev, u);
            const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m;
            tally.bail(bail_m);
            // Every lane takes rn: a lane without a live point ran the body
            // without committing it, so rn is its last magnitude again, bit
            // for bit (a held point's r, the one its finish needs).
            r = rn;
            // Lanes that have now run max_iters - 1 bodies without bailing:
            // their max_iters-th body is the dr update alone (max_iters >= 2
            // here; max_iters == 1 ends at the start).
            const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m & ~bail_m;
            tally.body(cap_m);
            done_m = cap_m | bail_m;
            last_m |= cap_m;
            if (done_m != 0 && sink.wants_k()) {
                if (__builtin_amdgcn_inverse_ballot_w64(done_m))
                    kh = tick - t0 + (__builtin_amdgcn_inverse_ballot_w64(cap_m) ? 1u : 0u);
            }
        } else {
            // Lanes making their max_iters-th body in this tick (max_iters >= 1).
            const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m;
            tick_body(ev, r, u, live);
            ++tick;
            tally.body(live_m);
            const float rn = tick_mag(ev, u);
            const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m & ~cap_m;
            tally.bail(bail_m);
            // At the iteration cap, r stays the magnitude computed before the
            // last body (the reference's off-by-one, src/shape/mandelbulb.rs:65-78).
            if (__builtin_amdgcn_inverse_ballot_w64(live_m & ~cap_m)) r = rn;
            done_m = cap_m | bail_m;
            if (done_m != 0 && sink.wants_k()) {
                if (__builtin_amdgcn_inverse_ballot_w64(done_m)) kh = tick - t0;
            }
        }
        live_m &= ~done_m;
        hold_m |= done_m;
    }
    tally.flush(q.stats);
    if (q.dynamic) wave_retire(q);
}

}  // namespace cth
