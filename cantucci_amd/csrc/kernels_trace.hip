// Sphere tracing with two-level lane refill.
//
// The reference marches 25 focus rays per frame on the CPU
// (ShapeMesh::get_focii, src/mesh/mod.rs:205-243): d = DE(pos); pos += dir * d;
// stop when d < EPSILON, at most 100 steps, no escape test.  Here every lane
// owns one ray; each ray step is one distance evaluation whose loop bodies run
// in the same branch-free ticks as the plain distance kernel.  When a lane's
// evaluation finishes, the lane applies the step (pos += dir * d, t += d) and
// either starts the ray's next evaluation or writes the ray and takes a new
// one, so lanes stay busy although both the evaluation lengths and the ray
// step counts differ wildly.  The hit test is d < eps_abs + eps_rel * t
// (eps_rel = 0 gives the reference's absolute test; eps_abs = 0 the relative
// one), a ray escapes when t > t_max (+inf: never, like the reference).
#include <hip/hip_runtime.h>

#include "internal.h"
#include "launch.h"
#include "point_sources.h"
#include "work_loop.h"

namespace cth {

struct TraceParams {
    uint32_t max_steps;
    float eps_abs, eps_rel, t_max;
};

// Pinhole camera rays generated from the pixel index:
//   u = ((2 (x + 0.5)) / W - 1) * tan_half * aspect
//   v = (1 - (2 (y + 0.5)) / H) * tan_half
//   d = (fwd + right * u) + up * v,  dir = d * (1 / |d|)
// (the focus rays of get_focii are this construction on the near plane,
// src/mesh/mod.rs:209-231, src/camera.rs:69-103,186-189).
struct RaysCamera {
    const cth_camera* cams;
    uint32_t W, H;
    FastDiv wh, w;
    __device__ __forceinline__ void get(uint32_t idx, float* o, float* d) const {
        uint32_t cam = fdiv(idx, wh);
        uint32_t pix = idx - cam * wh.d;
        uint32_t y = fdiv(pix, w);
        uint32_t x = pix - y * w.d;
        const cth_camera& c = cams[cam];
        float th = c.tan_half_fov, asp = c.aspect;
        float u = ((div_rn(2.0f * ((float)x + 0.5f), (float)W)) - 1.0f) * th * asp;
        float v = (1.0f - div_rn(2.0f * ((float)y + 0.5f), (float)H)) * th;
        float dd[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) dd[a] = (c.fwd[a] + c.right[a] * u) + c.up[a] * v;
        float m = sqrt_rn((dd[0] * dd[0] + dd[1] * dd[1]) + dd[2] * dd[2]);
        float inv = div_rn(1.0f, m);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            d[a] = dd[a] * inv;
            o[a] = c.eye[a];
        }
    }
};

// Explicit rays: origins (one shared or one per ray) and unit directions.
struct RaysExplicit {
    const float* origins;
    const float* dirs;
    int per_ray_origin;
    __device__ __forceinline__ void get(uint32_t idx, float* o, float* d) const {
        const float* op = origins + (per_ray_origin ? 3 * (size_t)idx : 0);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            o[a] = op[a];
            d[a] = dirs[3 * (size_t)idx + a];
        }
    }
};

struct TraceOut {
    float* depth;
    uint32_t* steps;
    float* hit_pos;  // nullable: position after the last step
};

template <class Ev, class Rays>
__global__ __launch_bounds__(256) void k_trace(
    Rays rays, MandelUniform u, TraceParams tp, TraceOut out, WorkQueue q) {
    Ev ev;
    // Wave-uniform lane masks, as in refill_loop (work_loop.h): lanes with an
    // evaluation in progress, and lanes whose evaluation finished and whose
    // step is not applied yet.  `has_ray` is only read at the batched refill.
    const uint64_t wave = wave_mask(true);
    uint64_t live_m = 0, pend_m = 0;
    bool has_ray = false;
    uint32_t idx = 0, nsteps = 0;
    uint32_t tick = 0, t0 = 0;  // wave's body count; tick at which the lane's evaluation started
    float r = 0.0f, t = 0.0f;
    float pos[3] = {0.0f, 0.0f, 0.0f}, dir[3] = {0.0f, 0.0f, 0.0f};
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;

    // One sphere-tracing step of the lane's ray by distance d.  When the ray
    // ends (hit, escape or the step limit) its outputs are written and the
    // lane has no ray.
    auto step = [&](float d) {
        pos[0] = pos[0] + dir[0] * d;
        pos[1] = pos[1] + dir[1] * d;
        pos[2] = pos[2] + dir[2] * d;
        t = t + d;
        ++nsteps;
        bool hit = d < (tp.eps_abs + tp.eps_rel * t);
        if (hit || t > tp.t_max || nsteps == tp.max_steps) {
            out.depth[idx] = hit ? t : __builtin_inff();
            out.steps[idx] = nsteps;
            if (out.hit_pos) {
                out.hit_pos[3 * (size_t)idx] = pos[0];
                out.hit_pos[3 * (size_t)idx + 1] = pos[1];
                out.hit_pos[3 * (size_t)idx + 2] = pos[2];
            }
            has_ray = false;
        }
    };

    for (;;) {
        if ((uint32_t)__popcll(wave & ~live_m) >= q.refill_min) {
            // step_m: lanes whose evaluation finished (after loop bodies, or
            // with none when the point is beyond the bailout radius) and whose
            // step is not applied yet; new_m: lanes that just took a ray.
            // Steps are applied and the next evaluations started over and
            // over while some lane has a zero-body step; new rays are fetched
            // only when none has, so that rays are generated in batches.
            uint64_t step_m = pend_m, new_m = 0;
            for (;;) {
                if (__builtin_amdgcn_inverse_ballot_w64(step_m)) step(ev.finish(r));
                const uint64_t start_m = (step_m | new_m) & wave_mask(has_ray);
                if (__builtin_amdgcn_inverse_ballot_w64(start_m)) {
                    ev.start(pos[0], pos[1], pos[2], u.power, u);
                    t0 = tick;
                    r = ev.mag();
                }
                step_m = wave_mask(r > u.bailout) & start_m;  // zero loop bodies: finished already
                live_m |= start_m & ~step_m;
                new_m = 0;
                if (step_m != 0) continue;
                // New rays for lanes without one.
                const uint64_t need = wave_mask(!has_ray);
                if (exhausted || need == 0) break;
                if (c_next >= c_end) {
                    const uint32_t b = next_chunk(q, round);
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
                const bool take = !has_ray && rank < avail;
                if (take) {
                    idx = c_next + rank;
                    rays.get(idx, pos, dir);
                    t = 0.0f;
                    nsteps = 0;
                    has_ray = true;
                }
                new_m = wave_mask(take);
                c_next += (nneed < avail) ? nneed : avail;
            }
            pend_m = 0;
            if (live_m == 0) break;
        }
        const bool live = __builtin_amdgcn_inverse_ballot_w64(live_m);
        const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m;  // the max_iters-th body
        ev.body(r, u, live);
        ++tick;
        const float rn = ev.mag();
        const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m & ~cap_m;
        if (__builtin_amdgcn_inverse_ballot_w64(live_m & ~cap_m)) r = rn;
        live_m &= ~(cap_m | bail_m);
        pend_m |= cap_m | bail_m;
    }
    if (q.dynamic) wave_retire(q);
}

template <class Rays>
static int run_trace(Ctx& c, const cth_shape& s, const Rays& rays, uint32_t total, const TraceParams& tp,
                     const TraceOut& out, hipStream_t st) {
    if (total == 0) return CTH_OK;
    MandelUniform u{};
    int family = FAM_SPHERE;
    if (s.kind == CTH_SHAPE_SPHERE) {
        u.max_iters = 1;
        u.bailout = 0.0f;
        for (int a = 0; a < 3; ++a) u.center[a] = s.center[a];
        u.radius = s.radius;
    } else {
        int rc = resolve_mandel(s, u, family);
        if (rc) return rc;
    }
    uint32_t blocks = 0;
    WorkQueue q = make_queue(c, total, c.shape_trace, blocks);
    dim3 grid(blocks), block(c.shape_trace.block);
    if (family == FAM_SPHERE) {
        hipLaunchKernelGGL((k_trace<SphereEv, Rays>), grid, block, 0, st, rays, u, tp, out, q);
    } else if (family == FAM_P8) {
        if (c.mode == CTH_MODE_FAST)
            hipLaunchKernelGGL((k_trace<P8Fast, Rays>), grid, block, 0, st, rays, u, tp, out, q);
        else
            hipLaunchKernelGGL((k_trace<P8Exact, Rays>), grid, block, 0, st, rays, u, tp, out, q);
    } else {
        hipLaunchKernelGGL((k_trace<TrigExact, Rays>), grid, block, 0, st, rays, u, tp, out, q);
    }
    return check_hip(hipGetLastError(), "trace launch");
}

int launch_trace_camera(Ctx& c, const cth_shape& s, const cth_camera* cams, uint32_t ncams, uint32_t W, uint32_t H,
                        uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                        uint32_t* steps, hipStream_t st) {
    uint64_t total = (uint64_t)ncams * W * H;
    if (total >= (1ull << 31)) return fail(CTH_E_INVALID, "more than 2^31 rays in one launch");
    RaysCamera rays{cams, W, H, make_fastdiv(W * H), make_fastdiv(W)};
    TraceParams tp{max_steps, eps_abs, eps_rel, t_max};
    return run_trace(c, s, rays, (uint32_t)total, tp, TraceOut{depth, steps, nullptr}, st);
}

int launch_trace_rays(Ctx& c, const cth_shape& s, const float* origins, int per_ray_origin, const float* dirs,
                      uint32_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                      uint32_t* steps, float* hit_pos, hipStream_t st) {
    RaysExplicit rays{origins, dirs, per_ray_origin};
    TraceParams tp{max_steps, eps_abs, eps_rel, t_max};
    return run_trace(c, s, rays, n, tp, TraceOut{depth, steps, hit_pos}, st);
}

}  // namespace cth
