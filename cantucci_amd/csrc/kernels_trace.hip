// Sphere tracing with lane refill.
//
// The reference marches 25 focus rays per frame on the CPU
// (ShapeMesh::get_focii, src/mesh/mod.rs:205-243): d = DE(pos); pos += dir * d;
// stop when d < EPSILON, at most 100 steps, no escape test.  Here every lane
// owns one ray; each ray step is one distance evaluation whose loop bodies run
// in the same branch-free ticks as the plain distance kernel.  When a lane's
// evaluation finishes, the lane applies the step (pos += dir * d, t += d) in
// the same iteration and either starts the ray's next evaluation or writes
// the ray, so lanes stay busy although both the evaluation lengths and the
// ray step counts differ wildly.  (On the camera workload an evaluation runs
// 1.6 loop bodies on average and a third run none, so the steps are as much
// of the work as the bodies.)  The hit test is d < eps_abs + eps_rel * t
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

// The focus rays of ShapeMesh::get_focii (src/mesh/mod.rs:205-231), built
// per lane from the reference camera (src/camera.rs:69-103,186-189) with the
// reference's f32 operations (the library never contracts):
//   f = normalize((pos + dir) - pos), s = normalize(cross(f, up)),
//   u = cross(s, f)                          look_at_rh, up = (0, 0, 1)
//   height = (2 near) tan(fov 0.5), width = height aspect
//   top_left = (-width / 2, -height / 2, -near), bottom_right = -top_left.xy
//   center = (top_left + (x width / fp, y height / fp, 0)) + (br - tl) / (2 fp)
//   p = ((s cx + u cy) + (-f) cz) + pos      the inverse view matrix's columns
//   dir = normalize(p - pos)
// cgmath's normalize is v * (1 / |v|) with |v| = sqrt((x x + y y) + z z).  The
// tangent is f64 tan rounded to f32: the correctly rounded value, as glibc's
// tanf gives for the fields of view tested.
struct RaysFocii {
    const cth_ref_camera* cams;
    uint32_t fp;
    FastDiv per_cam, side;  // fp * fp, fp
    __device__ __forceinline__ static void normalize3(float* v) {
        const float m = sqrt_rn((v[0] * v[0] + v[1] * v[1]) + v[2] * v[2]);
        const float inv = div_rn(1.0f, m);
        v[0] = v[0] * inv;
        v[1] = v[1] * inv;
        v[2] = v[2] * inv;
    }
    __device__ __forceinline__ static void cross3(const float* a, const float* b, float* o) {
        o[0] = a[1] * b[2] - a[2] * b[1];
        o[1] = a[2] * b[0] - a[0] * b[2];
        o[2] = a[0] * b[1] - a[1] * b[0];
    }
    __device__ __forceinline__ void get(uint32_t idx, float* o, float* d) const {
        const uint32_t cam = fdiv(idx, per_cam);
        const uint32_t cell = idx - cam * per_cam.d;
        const uint32_t x = fdiv(cell, side);
        const uint32_t y = cell - x * side.d;
        const cth_ref_camera& c = cams[cam];
        const float up[3] = {0.0f, 0.0f, 1.0f};
        float f[3], s[3], u[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) f[a] = (c.position[a] + c.direction[a]) - c.position[a];
        normalize3(f);
        cross3(f, up, s);
        normalize3(s);
        cross3(s, f, u);
        const float height = (2.0f * c.near_plane) * (float)tan((double)(c.fov * 0.5f));
        const float width = height * c.aspect;
        const float tl[3] = {div_rn(-width, 2.0f), div_rn(-height, 2.0f), -c.near_plane};
        const float br[3] = {div_rn(width, 2.0f), div_rn(height, 2.0f), -c.near_plane};
        const float ffp = (float)fp;
        const float off[3] = {(float)x * div_rn(width, ffp), (float)y * div_rn(height, ffp), 0.0f};
        float cc[3];
#pragma unroll
        for (int a = 0; a < 3; ++a) cc[a] = (tl[a] + off[a]) + div_rn(br[a] - tl[a], 2.0f * ffp);
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const float p = ((s[a] * cc[0] + u[a] * cc[1]) + (-f[a]) * cc[2]) + c.position[a] * 1.0f;
            d[a] = p - c.position[a];
            o[a] = c.position[a];
        }
        normalize3(d);
    }
};

// Directions of a ray source, one ray per thread (the focus rays' own
// output, for checking the construction).
template <class Rays>
__global__ __launch_bounds__(256) void k_ray_dirs(Rays rays, uint32_t n, float* dirs) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[3], d[3];
    rays.get(i, o, d);
    dirs[3 * (size_t)i] = d[0];
    dirs[3 * (size_t)i + 1] = d[1];
    dirs[3 * (size_t)i + 2] = d[2];
}

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

// Lanes whose evaluation has finished take their step together once this
// many are due (or no lane is running a loop body).  On the 32 x 1080p
// render, thresholds of 8, 16, 32, 40, 48, 56 and 64 gave 8.7, 8.2, 7.8-8.0,
// 7.7-8.0, 7.9-8.2, 8.3-8.6 and 9.8 ms, against 9.0 ms stepping at once.
constexpr uint32_t kStepMin = 32;

template <class Ev, class Rays, bool kCount = false>
__global__ __launch_bounds__(256) void k_trace(
    Rays rays, MandelUniform u, TraceParams tp, TraceOut out, WorkQueue q) {
    Ev ev;
    // Every lane that holds a ray is in one of two states, kept as
    // wave-uniform lane masks: running loop bodies of a distance evaluation
    // (live_m), or holding a finished evaluation whose step is due (step_m:
    // an evaluation from beyond the bailout radius runs no body, so it is
    // finished as soon as it starts).  One iteration of the loop runs one
    // loop body on the live lanes and then, once kStepMin lanes are due or
    // none is live, the step of every due lane: pos += dir d, t += d, the end
    // tests, and the next evaluation's start.  An evaluation's step waits at
    // most until half the wave's lanes are due; it never waits for the rest
    // of the wave.  Lanes whose ray ended take new rays in batches of at
    // least refill_min (ray generation is the costly part of a start).
    const uint64_t wave = wave_mask(true);
    uint64_t live_m = 0, step_m = 0, ray_m = 0;
    bool has_ray = false;
    uint32_t idx = 0, nsteps = 0;
    uint32_t tick = 0, t0 = 0;  // wave's body count; tick at which the lane's evaluation started
    float r = 0.0f, t = 0.0f;
    float pos[3] = {0.0f, 0.0f, 0.0f}, dir[3] = {0.0f, 0.0f, 0.0f};
    uint32_t c_next = 0, c_end = 0, round = 0;
    bool exhausted = false;
    WorkTally<kCount> tally;

    // Start the lane's next evaluation at pos; zero loop bodies when |pos|
    // is beyond the bailout radius (its step is then due at once).
    auto start = [&]() {
        ev.start(pos[0], pos[1], pos[2], u.power, u);
        t0 = tick;
        r = ev.mag();
    };
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
        // New rays for the lanes without one.
        uint64_t need = wave & ~ray_m;
        if (!exhausted && (uint32_t)__popcll(need) >= q.refill_min) {
            while (need != 0) {
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
                const bool take = __builtin_amdgcn_inverse_ballot_w64(need) && rank < avail;
                if (take) {
                    idx = c_next + rank;
                    rays.get(idx, pos, dir);
                    t = 0.0f;
                    nsteps = 0;
                    has_ray = true;
                    start();
                }
                const uint64_t took = wave_mask(take);
                const uint64_t zero = wave_mask(r > u.bailout) & took;
                tally.eval(took);
                tally.bail(zero);
                live_m |= took & ~zero;
                step_m |= zero;
                ray_m |= took;
                need &= ~took;
                c_next += (nneed < avail) ? nneed : avail;
            }
        }
        // Every lane with a ray is live or has a step due; none left means
        // the index space is exhausted (a wave without rays always fetches).
        if (ray_m == 0) break;
        uint64_t done_m = 0;
        if (live_m != 0) {
            const bool live = __builtin_amdgcn_inverse_ballot_w64(live_m);
            const uint64_t cap_m = wave_mask(t0 == tick + 1u - u.max_iters) & live_m;  // the max_iters-th body
            ev.body(r, u, live);
            ++tick;
            tally.body(live_m);
            const float rn = ev.mag();
            const uint64_t bail_m = wave_mask(rn > u.bailout) & live_m & ~cap_m;
            tally.bail(bail_m);
            // at the cap r stays the magnitude before the last body
            // (src/shape/mandelbulb.rs:65-78)
            if (__builtin_amdgcn_inverse_ballot_w64(live_m & ~cap_m)) r = rn;
            done_m = cap_m | bail_m;
            live_m &= ~done_m;
        }
        // The step runs for at least kStepMin lanes at a time: due lanes wait
        // (as step_m) while other lanes are live.  An evaluation averages
        // under two loop bodies, so stepping every iteration would run the
        // step's ~80 instructions for a third of the wave almost every time.
        const uint64_t go_m = done_m | step_m;
        if (go_m != 0 && live_m != 0 && (uint32_t)__popcll(go_m) < kStepMin) {
            step_m = go_m;
        } else if (go_m != 0) {
            if (__builtin_amdgcn_inverse_ballot_w64(go_m)) {
                step(ev.finish(r));
                if (has_ray) start();
            }
            const uint64_t cont = go_m & wave_mask(has_ray);
            const uint64_t zero = wave_mask(r > u.bailout) & cont;
            tally.eval(cont);
            tally.bail(zero);
            ray_m &= ~(go_m & ~cont);
            live_m |= cont & ~zero;
            step_m = zero;
        }
    }
    tally.flush(q.stats);
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
    WorkQueue q = make_queue(c, total, c.shape_trace, blocks, st);
    dim3 grid(blocks), block(c.shape_trace.block);
    if (family == FAM_SPHERE) {
        hipLaunchKernelGGL((k_trace<SphereEv, Rays>), grid, block, 0, st, rays, u, tp, out, q);
    } else if (family == FAM_P8) {
        with_p8(c.mode, [&](auto ev) {
            using Ev = decltype(ev);
            if (q.stats)
                hipLaunchKernelGGL((k_trace<Ev, Rays, true>), grid, block, 0, st, rays, u, tp, out, q);
            else
                hipLaunchKernelGGL((k_trace<Ev, Rays>), grid, block, 0, st, rays, u, tp, out, q);
        });
    } else if (q.stats) {
        hipLaunchKernelGGL((k_trace<TrigExact, Rays, true>), grid, block, 0, st, rays, u, tp, out, q);
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

int launch_trace_focii(Ctx& c, const cth_shape& s, const cth_ref_camera* cams, uint32_t ncams, uint32_t fp,
                       uint32_t max_steps, float eps, float* focii, float* depth, uint32_t* steps, float* dirs,
                       hipStream_t st) {
    const uint64_t total = (uint64_t)ncams * fp * fp;
    if (total >= (1ull << 31)) return fail(CTH_E_INVALID, "more than 2^31 focus rays in one launch");
    if (total == 0) return CTH_OK;
    RaysFocii rays{cams, fp, make_fastdiv(fp * fp), make_fastdiv(fp)};
    if (dirs) {
        const uint32_t n = (uint32_t)total;
        hipLaunchKernelGGL((k_ray_dirs<RaysFocii>), dim3((n + 255) / 256), dim3(256), 0, st, rays, n, dirs);
        if (int rc = check_hip(hipGetLastError(), "focus direction launch")) return rc;
    }
    TraceParams tp{max_steps, eps, 0.0f, __builtin_inff()};
    return run_trace(c, s, rays, (uint32_t)total, tp, TraceOut{depth, steps, focii}, st);
}

int launch_trace_rays(Ctx& c, const cth_shape& s, const float* origins, int per_ray_origin, const float* dirs,
                      uint32_t n, uint32_t max_steps, float eps_abs, float eps_rel, float t_max, float* depth,
                      uint32_t* steps, float* hit_pos, hipStream_t st) {
    RaysExplicit rays{origins, dirs, per_ray_origin};
    TraceParams tp{max_steps, eps_abs, eps_rel, t_max};
    return run_trace(c, s, rays, n, tp, TraceOut{depth, steps, hit_pos}, st);
}

}  // namespace cth
